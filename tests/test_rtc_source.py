"""CPU: the kernel source the library embeds for runtime specialisation
compiles with hipRTC for gfx950 in every variant the launcher can ask for
(no device needed to compile)."""
import ctypes as C
import os
import re

import pytest

from conftest import ROOT

INC = os.path.join(ROOT, "cusna_amd", "csrc", "rtc_source.inc")
HIPRTC = "/opt/rocm/lib/libhiprtc.so"


def _source(name: str = "kRtcPeriodSource") -> bytes:
    text = open(INC).read()
    m = re.search(name + r'\[\] = R"CUSNASRC\((.*?)\)CUSNASRC"', text, re.S)
    assert m, f"rtc_source.inc has no embedded {name}"
    return m.group(1).encode()


def _compile(src: bytes, defines: dict) -> tuple:
    L = C.CDLL(HIPRTC)
    prog = C.c_void_p()
    assert L.hiprtcCreateProgram(C.byref(prog), C.c_char_p(src), b"k.hip", 0, None, None) == 0
    opts = [b"--offload-arch=gfx950", b"-O3", b"-std=c++17", b"-ffp-contract=off", b"-Xclang",
            b"-target-feature", b"-Xclang", b"-packed-fp32-ops"]
    opts += [f"-D{k}={v}".encode() for k, v in defines.items()]
    arr = (C.c_char_p * len(opts))(*opts)
    rc = L.hiprtcCompileProgram(prog, len(opts), arr)
    size = C.c_size_t()
    L.hiprtcGetProgramLogSize(prog, C.byref(size))
    log = C.create_string_buffer(size.value + 1)
    L.hiprtcGetProgramLog(prog, log)
    L.hiprtcDestroyProgram(C.byref(prog))
    return rc, log.value.decode(errors="replace")


CASES = {
    # config-2 shape: simple effects, 2 waves x 4 words, cached change statistics
    "c2": dict(SAOM_WMAX=8, SAOM_NW=2, SAOM_REPLAY=0, SAOM_SIMPLE=1, SAOM_FIXED_K=7,
               SAOM_FIXED_TYPES="0,1,2,3,10,11,12", SAOM_FIXED_ETYPES="0,0,0,0,0,0,0",
               SAOM_FIXED_DYN="0,0,0,0,0,0,0", SAOM_CACHE_D=1),
    # general effects (interaction, gwesp, creation), recomputed statistics
    "general": dict(SAOM_WMAX=2, SAOM_NW=1, SAOM_REPLAY=0, SAOM_SIMPLE=0, SAOM_FIXED_K=5,
                    SAOM_FIXED_TYPES="0,1,24,5,32", SAOM_FIXED_ETYPES="0,1,0,0,0",
                    SAOM_FIXED_DYN="0,0,0,0,0", SAOM_CACHE_D=0),
    # streamed network (word bound 32), behaviour-dynamic covariate
    "stream": dict(SAOM_WMAX=32, SAOM_NW=16, SAOM_REPLAY=0, SAOM_SIMPLE=1, SAOM_FIXED_K=4,
                   SAOM_FIXED_TYPES="0,1,2,12", SAOM_FIXED_ETYPES="0,0,0,0",
                   SAOM_FIXED_DYN="0,0,0,1", SAOM_CACHE_D=1),
    # exact word bounds (not powers of two), rate effects compiled in
    "exact": dict(SAOM_WMAX=18, SAOM_NW=9, SAOM_REPLAY=0, SAOM_SIMPLE=1, SAOM_FIXED_K=3,
                  SAOM_FIXED_TYPES="0,1,2", SAOM_FIXED_ETYPES="0,0,0", SAOM_FIXED_DYN="0,0,0",
                  SAOM_CACHE_D=1, SAOM_HAS_RATE=1),
}

MULTI_CASES = {
    # two networks, cross effects, cached statistics
    "multi_lds": dict(SAOM_WMAX=8, SAOM_NW=2, SAOM_FIXED_K=8,
                      SAOM_FIXED_TYPES="0,1,2,33,0,1,37,34", SAOM_FIXED_ETYPES="0,0,0,0,0,0,0,1",
                      SAOM_FIXED_DYN="0,0,0,0,0,0,0,0", SAOM_FIXED_NETOF="0,0,0,0,1,1,1,1",
                      SAOM_FIXED_NETREF="-1,-1,-1,1,-1,-1,0,0", SAOM_CACHE_D=1),
    # streamed, four networks, recomputed statistics
    "multi_stream": dict(SAOM_WMAX=32, SAOM_NW=16, SAOM_FIXED_K=5,
                         SAOM_FIXED_TYPES="0,36,0,35,5", SAOM_FIXED_ETYPES="0,0,0,0,0",
                         SAOM_FIXED_DYN="0,0,0,0,0", SAOM_FIXED_NETOF="0,0,1,2,3",
                         SAOM_FIXED_NETREF="-1,3,-1,0,-1", SAOM_CACHE_D=0),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_embedded_source_compiles(name):
    if not os.path.exists(HIPRTC):
        pytest.skip("hipRTC not installed")
    if not os.path.exists(INC):
        pytest.skip("library not built (rtc_source.inc is generated by the build)")
    rc, log = _compile(_source(), CASES[name])
    assert rc == 0, log[-3000:]


@pytest.mark.parametrize("name", sorted(MULTI_CASES))
def test_embedded_multi_source_compiles(name):
    if not os.path.exists(HIPRTC):
        pytest.skip("hipRTC not installed")
    if not os.path.exists(INC):
        pytest.skip("library not built (rtc_source.inc is generated by the build)")
    rc, log = _compile(_source("kRtcMultiSource"), MULTI_CASES[name])
    assert rc == 0, log[-3000:]
