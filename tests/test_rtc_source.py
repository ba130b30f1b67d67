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


def _compile(src: bytes, defines: dict, want_code: bool = False) -> tuple:
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
    code = b""
    if want_code and rc == 0:
        csize = C.c_size_t()
        assert L.hiprtcGetCodeSize(prog, C.byref(csize)) == 0
        buf = C.create_string_buffer(csize.value)
        assert L.hiprtcGetCode(prog, buf) == 0
        code = buf.raw
    L.hiprtcDestroyProgram(C.byref(prog))
    if want_code:
        return rc, log.value.decode(errors="replace"), code
    return rc, log.value.decode(errors="replace")


def _kernel_metadata(elf: bytes) -> list:
    """amdhsa.kernels of a gfx950 code object: the msgpack descriptor in its
    NT_AMDGPU_METADATA note (ELF64 section headers, SHT_NOTE sections)."""
    import struct

    import msgpack

    assert elf[:4] == b"\x7fELF" and elf[4] == 2
    shoff, = struct.unpack_from("<Q", elf, 0x28)
    shentsize, shnum = struct.unpack_from("<HH", elf, 0x3A)
    for q in range(shnum):
        sh_type, = struct.unpack_from("<I", elf, shoff + q * shentsize + 4)
        if sh_type != 7:  # SHT_NOTE
            continue
        off, size = struct.unpack_from("<QQ", elf, shoff + q * shentsize + 0x18)
        pos = off
        while pos + 12 <= off + size:
            namesz, descsz, ntype = struct.unpack_from("<III", elf, pos)
            name = elf[pos + 12:pos + 12 + namesz].rstrip(b"\0")
            dpos = pos + 12 + ((namesz + 3) & ~3)
            if name == b"AMDGPU" and ntype == 32:
                return msgpack.unpackb(elf[dpos:dpos + descsz], raw=False)["amdhsa.kernels"]
            pos = dpos + ((descsz + 3) & ~3)
    raise AssertionError("no AMDGPU metadata note")


CASES = {
    # n = 500, seven effects: simple effects, 2 waves x 4 words, cached change statistics
    "n500": dict(SAOM_WMAX=8, SAOM_NW=2, SAOM_REPLAY=0, SAOM_SIMPLE=1, SAOM_FIXED_K=7,
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

# twelve effects in four waves x 2 words: the change statistics cached
CASES["k12"] = dict(SAOM_WMAX=8, SAOM_NW=4, SAOM_REPLAY=0, SAOM_SIMPLE=0, SAOM_FIXED_K=12,
                    SAOM_FIXED_TYPES="0,1,2,3,4,5,6,8,10,11,12,24",
                    SAOM_FIXED_ETYPES=",".join(["0"] * 12), SAOM_FIXED_DYN=",".join(["0"] * 12),
                    SAOM_CACHE_D=1)

# the 500-actor form the launcher prefers: four waves x 2 words, actor count,
# behaviour-effect count and flags compiled in, four-wave register budget
CASES["n500_w4"] = dict(CASES["n500"], SAOM_NW=4, SAOM_FIXED_N=500, SAOM_FIXED_KB=0,
                        SAOM_FIXED_FLAGS=1, SAOM_HAS_RATE=0, SAOM_WAVES_PER_EU=4)

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


@pytest.mark.parametrize("which,name", [("period", "n500"), ("period", "k12"),
                                        ("period", "n500_w4"), ("multi", "multi_lds")])
def test_specialised_kernels_keep_state_in_registers(which, name):
    """The specialised kernels walk a fixed effect list with a compile-time
    index (not a '#pragma unroll' loop, which the compiler refuses to unroll
    from about eight effects and then indexes the cached change statistics
    through scratch memory): no scratch, no spilled VGPRs."""
    if not os.path.exists(HIPRTC):
        pytest.skip("hipRTC not installed")
    if not os.path.exists(INC):
        pytest.skip("library not built (rtc_source.inc is generated by the build)")
    src, defs = ((_source(), CASES[name]) if which == "period"
                 else (_source("kRtcMultiSource"), MULTI_CASES[name]))
    rc, log, code = _compile(src, defs, want_code=True)
    assert rc == 0, log[-3000:]
    (k,) = _kernel_metadata(code)
    assert k[".private_segment_fixed_size"] == 0, k[".private_segment_fixed_size"]
    assert k[".vgpr_spill_count"] == 0
