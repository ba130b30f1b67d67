// rtc.hip -- runtime specialisation of the period kernel with hipRTC.
//
// The library embeds the period kernel's source (rtc_source.inc, generated by
// the build). For a launch it can compile a kernel specialised to the
// problem's effect list (types, creation/endowment types, behaviour-dynamic
// flags), word bound, block shape and variant: the effect loops unroll over
// compile-time types and, when k x alternatives-per-lane is small, the change
// statistics stay in registers between the choice and score passes. The
// arithmetic and its order are those of the generic kernel, so results are
// identical. Compiled code objects are cached in memory per device and on
// disk (CUSNA_RTC_CACHE, else $XDG_CACHE_HOME/cusna_amd, else
// $HOME/.cache/cusna_amd); any failure falls back to the generic kernel.
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>
#include <sys/stat.h>
#include <unistd.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "saom_kernels.h"

#include "rtc_source.inc"

namespace saom {
namespace {

struct RtcEntry {
  hipModule_t mod = nullptr;
  hipFunction_t fn = nullptr;
  bool failed = false;
};

std::mutex g_rtc_mu;
std::map<std::string, RtcEntry> g_rtc;  // key: device + options

uint64_t fnv1a(const std::string& a, uint64_t h = 1469598103934665603ull) {
  for (unsigned char c : a) {
    h ^= c;
    h *= 1099511628211ull;
  }
  return h;
}

std::string cache_dir() {
  const char* e = std::getenv("CUSNA_RTC_CACHE");
  std::string d;
  if (e && *e) {
    d = e;
  } else if ((e = std::getenv("XDG_CACHE_HOME")) && *e) {
    d = std::string(e) + "/cusna_amd";
  } else if ((e = std::getenv("HOME")) && *e) {
    d = std::string(e) + "/.cache/cusna_amd";
  } else {
    return std::string();
  }
  // create each component (mkdir -p); failures just disable the disk cache
  for (size_t pos = 1; pos <= d.size(); ++pos)
    if (pos == d.size() || d[pos] == '/') (void)mkdir(d.substr(0, pos).c_str(), 0755);
  return access(d.c_str(), W_OK) == 0 ? d : std::string();
}

bool read_file(const std::string& path, std::vector<char>& out) {
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) return false;
  std::fseek(f, 0, SEEK_END);
  const long sz = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  out.resize(sz > 0 ? (size_t)sz : 0);
  const bool ok = sz > 0 && std::fread(out.data(), 1, out.size(), f) == out.size();
  std::fclose(f);
  return ok;
}

void write_file(const std::string& path, const std::vector<char>& data) {
  const std::string tmp = path + ".tmp" + std::to_string((long)getpid());
  FILE* f = std::fopen(tmp.c_str(), "wb");
  if (!f) return;
  const bool ok = std::fwrite(data.data(), 1, data.size(), f) == data.size();
  std::fclose(f);
  if (ok) (void)std::rename(tmp.c_str(), path.c_str());
  else (void)std::remove(tmp.c_str());
}

std::string list(const int* v, int k) {
  std::string s;
  for (int e = 0; e < k; ++e) s += (e ? "," : "") + std::to_string(v[e]);
  return s;
}

bool compile(const char* source, const std::vector<std::string>& opts,
             std::vector<char>& code) {
  hiprtcProgram prog;
  if (hiprtcCreateProgram(&prog, source, "saom_rtc.hip", 0, nullptr, nullptr) != HIPRTC_SUCCESS)
    return false;
  std::vector<const char*> o;
  for (const auto& x : opts) o.push_back(x.c_str());
  bool ok = hiprtcCompileProgram(prog, (int)o.size(), o.data()) == HIPRTC_SUCCESS;
  size_t cs = 0;
  if (ok) ok = hiprtcGetCodeSize(prog, &cs) == HIPRTC_SUCCESS && cs > 0;
  if (ok) {
    code.resize(cs);
    ok = hiprtcGetCode(prog, code.data()) == HIPRTC_SUCCESS;
  }
  (void)hiprtcDestroyProgram(&prog);
  return ok;
}

// Compiled kernel `name` of `source` for these options, from the in-memory
// cache, the disk cache or a fresh compile (null on any failure).
void* rtc_get(const char* source, const char* name, const std::vector<std::string>& opts) {
  int dev = 0;
  (void)hipGetDevice(&dev);
  std::string key = std::to_string(dev) + " " + name;
  for (const auto& x : opts) key += " " + x;
  std::lock_guard<std::mutex> lk(g_rtc_mu);
  RtcEntry& ent = g_rtc[key];
  if (ent.fn || ent.failed) return (void*)ent.fn;
  ent.failed = true;  // until loaded
  std::vector<char> code;
  std::string path;
  const std::string dir = cache_dir();
  if (!dir.empty()) {
    char fname[64];
    std::snprintf(fname, sizeof(fname), "/%s_%016llx.co", name,
                  (unsigned long long)fnv1a(key.substr(key.find(' ')), fnv1a(source)));
    path = dir + fname;
  }
  if (path.empty() || !read_file(path, code)) {
    if (!compile(source, opts, code)) return nullptr;
    if (!path.empty()) write_file(path, code);
  }
  if (hipModuleLoadData(&ent.mod, code.data()) != hipSuccess) return nullptr;
  if (hipModuleGetFunction(&ent.fn, ent.mod, name) != hipSuccess) {
    ent.fn = nullptr;
    return nullptr;
  }
  ent.failed = false;
  return (void*)ent.fn;
}

// Raise a module function's dynamic-LDS limit to the CU's 160 KB once.
void raise_lds(void* fn) {
  static std::mutex mu;
  static std::map<void*, bool> raised;
  std::lock_guard<std::mutex> lk(mu);
  if (!raised[fn]) {
    raised[fn] = true;
    (void)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    (void)hipGetLastError();
  }
}

}  // namespace

void* rtc_period_kernel(const SimParams& P, LaunchShape s, bool replay, bool simple, bool fixed_n,
                        int waves_per_eu, bool lean) {
  const int R = s.wmax / s.nw;
  std::vector<std::string> opts = {
      "--offload-arch=gfx950", "-O3", "-std=c++17", "-ffp-contract=off",
      "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
      "-DSAOM_WMAX=" + std::to_string(s.wmax), "-DSAOM_NW=" + std::to_string(s.nw),
      std::string("-DSAOM_REPLAY=") + (replay ? "1" : "0"),
      std::string("-DSAOM_SIMPLE=") + (simple ? "1" : "0"),
      "-DSAOM_FIXED_K=" + std::to_string(P.k),
      "-DSAOM_FIXED_TYPES=" + list(P.tid, P.k), "-DSAOM_FIXED_ETYPES=" + list(P.ety, P.k),
      "-DSAOM_FIXED_DYN=" + list(P.dyn, P.k),
      std::string("-DSAOM_CACHE_D=") + (P.k * R <= 32 ? "1" : "0"),
      std::string("-DSAOM_HAS_RATE=") + ((P.flags & F_RATE) ? "1" : "0"),
      "-DSAOM_FIXED_KB=" + std::to_string(P.kb), "-DSAOM_FIXED_FLAGS=" + std::to_string(P.flags)};
  if (fixed_n) opts.push_back("-DSAOM_FIXED_N=" + std::to_string(P.n));
  else opts.push_back("-DSAOM_FIXED_W64=" + std::to_string(P.W64));
  if (waves_per_eu > 0) opts.push_back("-DSAOM_WAVES_PER_EU=" + std::to_string(waves_per_eu));
  // CUSNA_RTC_LEAN=1: every specialisation lean (a check that the lean build
  // gives the same results)
  const char* le = std::getenv("CUSNA_RTC_LEAN");
  if (lean || (le && le[0] == '1')) opts.push_back("-DSAOM_LEAN=1");
  return rtc_get(kRtcPeriodSource, "saom_period_rtc", opts);
}

void* rtc_multi_kernel(const MultiParams& Q, int nw, int wmax) {
  const SimParams& P = Q.base;
  const int R = wmax / nw;
  std::vector<int> netref(P.k);
  for (int e = 0; e < P.k; ++e) netref[e] = Q.netref[e];
  std::vector<std::string> opts = {
      "--offload-arch=gfx950", "-O3", "-std=c++17", "-ffp-contract=off",
      "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
      "-DSAOM_WMAX=" + std::to_string(wmax), "-DSAOM_NW=" + std::to_string(nw),
      "-DSAOM_FIXED_K=" + std::to_string(P.k),
      "-DSAOM_FIXED_TYPES=" + list(P.tid, P.k), "-DSAOM_FIXED_ETYPES=" + list(P.ety, P.k),
      "-DSAOM_FIXED_DYN=" + list(P.dyn, P.k),
      "-DSAOM_FIXED_NETOF=" + list(Q.netof, P.k), "-DSAOM_FIXED_NETREF=" + list(netref.data(), P.k),
      std::string("-DSAOM_CACHE_D=") + (P.k * R <= 32 ? "1" : "0")};
  return rtc_get(kRtcMultiSource, "saom_multi_rtc", opts);
}

int launch_multi_rtc(void* fn, const MultiParams* Pd, int grid, size_t lds, int nw,
                     void* stream) {
  raise_lds(fn);
  const MultiParams* pd = Pd;
  void* args[] = {&pd};
  return (int)hipModuleLaunchKernel((hipFunction_t)fn, (unsigned)grid, 1, 1, 64u * (unsigned)nw,
                                    1, 1, (unsigned)lds, (hipStream_t)stream, args, nullptr);
}

bool uses_scratch(void* fn) {
  int bytes = 0;
  if (hipFuncGetAttribute(&bytes, HIP_FUNC_ATTRIBUTE_LOCAL_SIZE_BYTES, (hipFunction_t)fn) !=
      hipSuccess) {
    (void)hipGetLastError();
    return true;
  }
  return bytes > 0;
}

int kernel_vgprs(void* fn) {
  int regs = 0;
  if (hipFuncGetAttribute(&regs, HIP_FUNC_ATTRIBUTE_NUM_REGS, (hipFunction_t)fn) != hipSuccess) {
    (void)hipGetLastError();
    return 1 << 30;
  }
  return regs;
}

bool specialize_policy(long long chains_x_actors) {
  const char* e = std::getenv("CUSNA_SPECIALIZE");
  if (e && e[0] == '0') return false;
  if (e && e[0] == '1') return true;
  return chains_x_actors >= kSpecializeWork;
}

bool same_effects(const SimParams& a, const SimParams& b) {
  if (a.k != b.k) return false;
  // the specialised kernel compiles in the behaviour-effect count and the
  // feature flags (rate effects and their count included): a launch is
  // specialised only for groups that agree on them
  if (a.kb != b.kb || a.flags != b.flags) return false;
  for (int e = 0; e < a.k; ++e)
    if (a.tid[e] != b.tid[e] || a.ety[e] != b.ety[e] || a.dyn[e] != b.dyn[e]) return false;
  return true;
}

int launch_period_rtc(void* fn, const PeriodLaunch& L, int nw, void* stream) {
  raise_lds(fn);
  const SimParams* pd = L.Pd;
  const uint32_t* items = L.items;
  uint32_t k0 = L.key0, k1 = L.key1;
  void* args[] = {&pd, &items, &k0, &k1};
  return (int)hipModuleLaunchKernel((hipFunction_t)fn, (unsigned)L.grid, 1, 1, 64u * (unsigned)nw, 1, 1,
                                    (unsigned)L.lds, (hipStream_t)stream, args, nullptr);
}

}  // namespace saom
