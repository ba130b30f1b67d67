"""CPU: the R package's link layout of INTEGRATION.md, built for real.

The reference's CPU engine (src/libcusna_saom_cpu.cpp) is compiled the way the
package's GPU build compiles it (src/Makevars:9-15: -DLIBCUSNA_BUILD
-DCUSNA_WITH_CUDA, without -DCUSNA_CPU_ONLY) together with a small C++ caller
standing in for the glue's calls (src/cusna_glue.cpp, src/cusna_saom_glue.cpp),
into one package object linked against libcusna_amd.so. A main program opens
that object with RTLD_LOCAL, as R opens a package, and the caller reports
which object each libcusna symbol it binds resolves to:

  libcusna_amd.so   cusna_saom_simulate_period, cusna_saom_simulate_multi,
                    cusna_pack_bits, cusna_philox4x32, cusna_has_cuda,
                    cusna_abi_version, cusna_version_string, cusna_last_error
  package object    cusna_saom_simulate_period_cpu, cusna_saom_simulate_multi_cpu,
                    and the statistic carriers cusna_network_stats,
                    cusna_behavior_stats, cusna_gof_distribution,
                    cusna_cross_stats (the package's own definitions bind
                    first inside it)

and that cusna_last_error, which the CPU-only layout would define in the
package as "", returns the GPU library's message after a refused call.
Skipped where the reference sources are absent.
"""
import os
import subprocess

import pytest

from conftest import ROOT

REFSRC = "/root/reference/src"
AMD_LIB = os.path.join(ROOT, "cusna_amd", "_lib", "libcusna_amd.so")

AMD_SYMBOLS = ["cusna_saom_simulate_period", "cusna_saom_simulate_multi", "cusna_pack_bits",
               "cusna_philox4x32", "cusna_has_cuda", "cusna_abi_version",
               "cusna_version_string", "cusna_last_error"]
PKG_SYMBOLS = ["cusna_saom_simulate_period_cpu", "cusna_saom_simulate_multi_cpu",
               "cusna_network_stats", "cusna_behavior_stats", "cusna_gof_distribution",
               "cusna_cross_stats"]

CALLER = r"""
#include <dlfcn.h>
#include <stdio.h>
#include <string.h>
#include "libcusna.h"

static void where(const char* name, const void* addr) {
  Dl_info info;
  if (dladdr(addr, &info) && info.dli_fname) printf("sym %s %s\n", name, info.dli_fname);
  else printf("sym %s ?\n", name);
}

extern "C" __attribute__((visibility("default"))) int glue_probe(void) {
#define W(f) where(#f, (const void*)&f)
  W(cusna_saom_simulate_period); W(cusna_saom_simulate_multi); W(cusna_pack_bits);
  W(cusna_philox4x32); W(cusna_has_cuda); W(cusna_abi_version); W(cusna_version_string);
  W(cusna_last_error); W(cusna_saom_simulate_period_cpu); W(cusna_saom_simulate_multi_cpu);
  W(cusna_network_stats); W(cusna_behavior_stats); W(cusna_gof_distribution);
  W(cusna_cross_stats);
  printf("abi %d\n", (int)cusna_abi_version());
  printf("version %s\n", cusna_version_string());
  // a call the GPU library refuses before touching a device: unknown effect id
  uint32_t rows[3] = {0u, 0u, 0u};
  int32_t tid[1] = {99};
  float theta[1] = {0.0f}, lam[1] = {1.0f};
  cusna_saom_problem p;
  memset(&p, 0, sizeof(p));
  p.struct_version = CUSNA_ABI_VERSION;
  p.n = 3; p.W = 1; p.k = 1; p.M = 1;
  p.x0_out = rows; p.type_ids = tid; p.theta = theta; p.lam = lam;
  cusna_saom_result r;
  memset(&r, 0, sizeof(r));
  r.struct_version = CUSNA_ABI_VERSION;
  const int st = (int)cusna_saom_simulate_period(&p, &r);
  printf("status %d\nerror %s\n", st, cusna_last_error());
  return 0;
}
"""

MAIN = r"""
#include <dlfcn.h>
#include <stdio.h>
int main(int argc, char** argv) {
  void* h = dlopen(argv[1], RTLD_NOW | RTLD_LOCAL);
  if (!h) { fprintf(stderr, "%s\n", dlerror()); return 2; }
  int (*probe)(void) = (int (*)(void))dlsym(h, "glue_probe");
  if (!probe) { fprintf(stderr, "no glue_probe\n"); return 3; }
  return probe();
}
"""


def test_package_link_layout(tmp_path):
    if not os.path.exists(os.path.join(REFSRC, "libcusna_saom_cpu.cpp")):
        pytest.skip("reference sources not present")
    if not os.path.exists(AMD_LIB):
        pytest.skip("libcusna_amd.so not built")
    libdir = os.path.dirname(AMD_LIB)
    caller = tmp_path / "glue_probe.cpp"
    caller.write_text(CALLER)
    main = tmp_path / "main.c"
    main.write_text(MAIN)
    pkg = tmp_path / "cusna_pkg.so"
    subprocess.run(["g++", "-std=c++17", "-O1", "-fPIC", "-shared", "-fopenmp",
                    "-DLIBCUSNA_BUILD", "-DCUSNA_WITH_CUDA", f"-I{REFSRC}",
                    os.path.join(REFSRC, "libcusna_saom_cpu.cpp"), str(caller),
                    f"-L{libdir}", "-lcusna_amd", f"-Wl,-rpath,{libdir}", "-ldl",
                    "-o", str(pkg)], check=True, capture_output=True)
    exe = tmp_path / "main"
    subprocess.run(["gcc", "-O1", str(main), "-ldl", "-o", str(exe)], check=True,
                   capture_output=True)
    out = subprocess.run([str(exe), str(pkg)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    lines = out.stdout.splitlines()
    where = {ln.split()[1]: ln.split()[2] for ln in lines if ln.startswith("sym ")}
    for s in AMD_SYMBOLS:
        assert os.path.basename(where[s]) == "libcusna_amd.so", (s, where[s])
    for s in PKG_SYMBOLS:
        assert os.path.basename(where[s]) == "cusna_pkg.so", (s, where[s])
    info = dict(ln.split(" ", 1) for ln in lines if not ln.startswith("sym "))
    assert info["abi"] == "1"
    assert "gfx950" in info["version"] or "MI355" in info["version"], info["version"]
    assert info["status"] == "5"  # CUSNA_ERR_UNSUPPORTED
    assert "unknown effect type id" in info["error"]
