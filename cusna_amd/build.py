"""Build of the gfx950 library ``cusna_amd/_lib/libcusna_amd.so``.

Every HIP source is compiled by hipcc for gfx950 into an object file (in
parallel), then linked into one shared library exporting the C ABI of
include/libcusna.h. Device code is compiled with -ffp-contract=off so that no
multiply-add is fused: the change statistics and linear predictors must round
as the reference's separate multiply and add do.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "_lib")
LIB = os.path.join(OUT_DIR, "libcusna_amd.so")
ARCH = "gfx950"

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-fvisibility=hidden", "-Wno-unused-result",
]

# (source, extra defines, object name)
UNITS = [("capi.hip", [], "capi.o"), ("saom_stats.hip", [], "saom_stats.o"),
         ("saom_launch.hip", [], "saom_launch.o"), ("mom.hip", [], "mom.o"), ("mlp.hip", [], "mlp.o"), ("rtc.hip", [], "rtc.o"),
         ("multi.hip", [], "multi.o")] + [
    ("saom_period.hip", [f"-DSAOM_WMAX={w}"], f"saom_period_w{w}.o") for w in (1, 2, 4, 8, 16, 32, 64, 128)] + [
    ("saom_multi.hip", [f"-DSAOM_WMAX={w}"], f"saom_multi_w{w}.o") for w in (1, 2, 4, 8, 16, 32, 64, 128)]


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise FileNotFoundError("hipcc not found")


def _compile(src: str, defs: list, obj: str, verbose: bool) -> None:
    cmd = [hipcc(), *HIPCC_FLAGS, *defs, "-c", os.path.join(CSRC, src), "-o", obj]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{r.stderr[-4000:]}")


# Sources of the period kernel embedded in the library for runtime
# specialisation (rtc.hip): headers inlined in dependency order, local
# includes dropped, a preamble supplying what hipRTC's implicit runtime header
# names differently.
RTC_SOURCES = ["saom_common.h", "saom_kernels.h", "saom_device.h", "saom_period.hip"]
RTC_PREAMBLE = """#define SAOM_RTC 1
typedef __hip_internal::uint8_t uint8_t;
typedef __hip_internal::uint32_t uint32_t;
typedef __hip_internal::uint64_t uint64_t;
typedef __hip_internal::int32_t int32_t;
typedef __hip_internal::int64_t int64_t;
#ifndef INFINITY
#define INFINITY __builtin_inff()
#endif
"""


# The multi-network kernel's source: the period source's helpers (no period
# kernel) followed by saom_multi.hip.
RTC_MULTI_SOURCES = RTC_SOURCES + ["saom_multi.hip"]


def _rtc_body(files: list, preamble: str) -> str:
    import re

    parts = [preamble]
    for f in files:
        with open(os.path.join(CSRC, f)) as fh:
            src = fh.read()
        src = re.sub(r"^#pragma once\s*$", "", src, flags=re.M)
        src = re.sub(r"^#include .*$", "", src, flags=re.M)
        parts.append(f"// ---- {f}\n{src}")
    body = "\n".join(parts)
    if ")CUSNASRC" in body:
        raise RuntimeError("kernel source contains the raw-string delimiter")
    return body


def write_rtc_source() -> str:
    """Generate csrc/rtc_source.inc (raw string literals of the kernel sources)."""
    out = os.path.join(CSRC, "rtc_source.inc")
    text = ("// generated by cusna_amd/build.py from " + ", ".join(RTC_MULTI_SOURCES) + "\n"
            'static const char kRtcPeriodSource[] = R"CUSNASRC('
            + _rtc_body(RTC_SOURCES, RTC_PREAMBLE) + ')CUSNASRC";\n'
            'static const char kRtcMultiSource[] = R"CUSNASRC('
            + _rtc_body(RTC_MULTI_SOURCES, RTC_PREAMBLE + "#define SAOM_HELPERS_ONLY 1\n")
            + ')CUSNASRC";\n')
    old = open(out).read() if os.path.exists(out) else None
    if old != text:
        with open(out, "w") as fh:
            fh.write(text)
    return out


def build(verbose: bool = False, jobs: int | None = None) -> str:
    """Compile every HIP unit for gfx950 and link the shared library."""
    write_rtc_source()
    os.makedirs(OUT_DIR, exist_ok=True)
    objdir = os.path.join(OUT_DIR, "obj")
    os.makedirs(objdir, exist_ok=True)
    jobs = jobs or min(len(UNITS), max(1, (os.cpu_count() or 4)))
    with cf.ThreadPoolExecutor(jobs) as ex:
        futs = [ex.submit(_compile, s, d, os.path.join(objdir, o), verbose) for s, d, o in UNITS]
        for f in futs:
            f.result()
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp,
           *[os.path.join(objdir, o) for _, _, o in UNITS], "-lhiprtc"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-4000:]}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
