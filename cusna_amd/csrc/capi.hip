// capi.hip -- the libcusna C ABI on top of the gfx950 kernels.
//
// Host side of the drop-in library: argument validation with the reference's
// status codes (cran/cusna src/libcusna_saom_cpu.cpp:317-320), conversion of
// the ABI's u32 bit rows and float32/float64 descriptors to the kernels'
// u64 rows and float64 descriptors (the promotions of :334-353), device
// buffers, launches on a per-thread HIP stream, and copies back. Every call is
// synchronous: outputs are in the caller's buffers on return.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/libcusna.h"
#include "saom_kernels.h"
#include "plan.h"

using namespace cusna_impl;

namespace cusna_impl {

thread_local std::string g_last_error;

void set_error(const std::string& s) { g_last_error = s; }

// One HIP stream per host thread: concurrent callers never share a stream.
struct ThreadStream {
  hipStream_t s = nullptr;
  int dev = -1;
  ~ThreadStream() {
    if (s) (void)hipStreamDestroy(s);
  }
};
thread_local ThreadStream g_stream;

int device_count() {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

hipStream_t thread_stream() {
  int dev = 0;
  (void)hipGetDevice(&dev);
  if (!g_stream.s || g_stream.dev != dev) {
    if (g_stream.s) (void)hipStreamDestroy(g_stream.s);
    g_stream.s = nullptr;
    if (hipStreamCreateWithFlags(&g_stream.s, hipStreamNonBlocking) != hipSuccess) return nullptr;
    g_stream.dev = dev;
  }
  return g_stream.s;
}

// ABI bit rows (n*W u32) -> kernel rows (n*W64 u64); NULL -> off-diagonal
// (the reference's default masks, libcusna_saom_cpu.cpp:326-331) or zero.
std::vector<uint64_t> rows64(const uint32_t* rows, int n, int W, int W64, int dflt) {
  std::vector<uint64_t> out((size_t)n * W64, 0ull);
  for (int i = 0; i < n; ++i) {
    uint64_t* o = &out[(size_t)i * W64];
    if (rows) {
      for (int w = 0; w < W; ++w) o[w >> 1] |= (uint64_t)rows[(size_t)i * W + w] << (32 * (w & 1));
      // bits past n are not ties
      if (n & 63) o[W64 - 1] &= (1ull << (n & 63)) - 1;
    } else if (dflt == 1) {
      for (int j = 0; j < n; ++j)
        if (j != i) o[j >> 6] |= 1ull << (j & 63);
    }
  }
  return out;
}

std::vector<uint64_t> dense_rows64(const uint8_t* a, int n, int W64) {
  std::vector<uint64_t> out((size_t)n * W64, 0ull);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j)
      if (a[(size_t)i * n + j]) out[(size_t)i * W64 + (j >> 6)] |= 1ull << (j & 63);
  return out;
}

int wmax_for(int W64) {
  int w = 1;
  while (w < W64) w <<= 1;
  return w;
}

constexpr size_t kLdsLimit = 160 * 1024;

// Effects whose change statistic needs a given per-alternative count.
int effect_flags(const int32_t* t, int k, int kb, const int32_t* dyn) {
  int f = 0;
  for (int e = 0; e < k; ++e) {
    switch (t[e]) {
      case 2: case 24: f |= saom::F_TP; break;
      case 5: case 29: f |= saom::F_TP | saom::F_TPALL; break;
      case 6: f |= saom::F_TP; break;
      case 19: f |= saom::F_RDEG; break;
      case 30: f |= saom::F_SOALL; break;
      default: break;
    }
    if (dyn && dyn[e]) f |= saom::F_Z;
  }
  if (kb > 0) f |= saom::F_Z;
  return f;
}

// Block shapes chosen by cusna_period_plan_autotune, by (word bound, words,
// effects, effect flags); plan_create uses them when the caller leaves
// `threads` at 0.
struct ShapeKey {
  int wmax, W64, k, flags;
  bool operator<(const ShapeKey& o) const {
    if (wmax != o.wmax) return wmax < o.wmax;
    if (W64 != o.W64) return W64 < o.W64;
    if (k != o.k) return k < o.k;
    return flags < o.flags;
  }
};
std::mutex g_shape_mu;
std::map<ShapeKey, int> g_shape_cache;

int cached_nw(const ShapeKey& key) {
  std::lock_guard<std::mutex> lk(g_shape_mu);
  auto it = g_shape_cache.find(key);
  return it == g_shape_cache.end() ? 0 : it->second;
}

int pick_nw(int wmax, int threads) {
  if (wmax > 16) return 16;  // streamed networks: one shape
  if (threads > 0) {
    int nw = threads / 64;
    if (nw < 1) nw = 1;
    int p = 1;
    while (p * 2 <= nw) p *= 2;
    nw = p;
    if (nw > wmax) nw = wmax;
    if (wmax == 16 && nw < 2) nw = 2;
    return nw;
  }
  // default shape: about four alternatives per lane
  int nw = wmax / 4;
  if (nw < 1) nw = 1;
  if (wmax == 16 && nw < 2) nw = 2;
  return nw;
}

cusna_status validate(const cusna_saom_problem* pb, const cusna_saom_result* r) {
  if (!pb || !r) return CUSNA_ERR_NULL_ARG;
  if (pb->struct_version != CUSNA_ABI_VERSION || r->struct_version != CUSNA_ABI_VERSION)
    return CUSNA_ERR_SHAPE;
  const int n = pb->n, W = pb->W, k = pb->k, M = pb->M;
  if (n < 1 || n > CUSNA_MAX_N || W != (n + 31) / 32 || k < 1 || M < 1) return CUSNA_ERR_SHAPE;
  if (k > CUSNA_MAXEFF || pb->kb < 0 || pb->kb > CUSNA_MAXBEFF) return CUSNA_ERR_SHAPE;
  if (!pb->x0_out || !pb->type_ids || !pb->theta) return CUSNA_ERR_NULL_ARG;
  if (pb->nrate < 0 || pb->nrate > CUSNA_MAXRATE) return CUSNA_ERR_SHAPE;
  if (pb->nrate > 0) {  // rate effects (libcusna_saom_cpu.cpp:170-193)
    if (!pb->rkind || !pb->ralpha) return CUSNA_ERR_NULL_ARG;
    for (int q = 0; q < pb->nrate; ++q) {
      if (pb->rkind[q] < 0 || pb->rkind[q] > 2) return CUSNA_ERR_UNSUPPORTED;
      if (pb->rkind[q] == 0 && !pb->ru) return CUSNA_ERR_NULL_ARG;
    }
  }
  if (pb->kb > 0 && (!pb->lamb || !pb->thetab || !pb->bty || !pb->z0)) return CUSNA_ERR_NULL_ARG;
  for (int e = 0; e < k; ++e) {
    const int t = pb->type_ids[e];
    if (t < 0 || t > 32) {
      set_error("unknown effect type id");
      return CUSNA_ERR_UNSUPPORTED;
    }
    if (t == 32) {
      const int c0 = pb->int0 ? pb->int0[e] : -1, c1 = pb->int1 ? pb->int1[e] : -1;
      const int c2 = pb->int2 ? pb->int2[e] : -1;
      if (c0 < 0 || c0 >= e || c1 < 0 || c1 >= e || c2 >= e) return CUSNA_ERR_SHAPE;
      if (pb->type_ids[c0] == 32 || pb->type_ids[c1] == 32 || (c2 >= 0 && pb->type_ids[c2] == 32))
        return CUSNA_ERR_SHAPE;
    }
  }
  if (pb->act && (pb->nact < 1 || pb->nact > n)) return CUSNA_ERR_SHAPE;
  if (pb->act)
    for (int q = 0; q < pb->nact; ++q)
      if (pb->act[q] < 0 || pb->act[q] >= n) return CUSNA_ERR_SHAPE;
  if (pb->kb > 0)
    for (int e = 0; e < pb->kb; ++e)
      if (pb->bty[e] < 0 || pb->bty[e] > 4) return CUSNA_ERR_UNSUPPORTED;
  return CUSNA_OK;
}

}  // namespace cusna_impl

namespace cusna_impl {

cusna_status plan_create(const cusna_saom_problem* pb, bool want_dense, const double* log_u0,
                         const double* loglog_u, int32_t smax, hipStream_t stream,
                         cusna_period_plan** out, const DeviceArena* arena) {
  cusna_saom_result dummy{};
  dummy.struct_version = CUSNA_ABI_VERSION;
  cusna_status st = validate(pb, &dummy);
  if (st != CUSNA_OK) return st;
  if (device_count() < 1) {
    set_error("no HIP device");
    return CUSNA_ERR_NO_CUDA;
  }
  const int n = pb->n, W = pb->W, k = pb->k, M = pb->M, kb = pb->kb;
  const int W64 = (n + 63) / 64;
  const int wmax = wmax_for(W64);
  if (wmax > 128) {
    set_error("n above 8192");
    return CUSNA_ERR_SHAPE;
  }
  const bool stream_rows = wmax > 16;  // network rows in device memory
  const bool replay = log_u0 != nullptr;
  std::unique_ptr<cusna_period_plan> pl(new cusna_period_plan());
  pl->n = n;
  pl->k = k;
  pl->kb = kb;
  pl->M = M;
  pl->W64 = W64;
  pl->wmax = wmax;
  pl->replay = replay;
  pl->want_dense = want_dense;
  pl->symflag = pb->symflag;
  saom::SimParams& P = pl->P;
  std::memset(&P, 0, sizeof(P));
  P.n = n;
  P.k = k;
  P.kb = kb;
  P.M = M;
  P.W64 = W64;
  P.RS = n;  // word-major LDS rows: word w of actor h at w * n + h
  P.nact = pb->act ? pb->nact : n;
  P.has_comp = pb->has_comp;
  P.symflag = pb->symflag;
  P.maxdeg = pb->maxdegree;
  P.cond = pb->conditional;
  P.target = pb->target_dist;
  P.zmin = pb->zmin;
  P.zmax = pb->zmax;
  P.smax = replay ? smax : 0;
  P.key0 = (uint32_t)(pb->seed & 0xFFFFFFFFu);
  P.key1 = (uint32_t)(pb->seed >> 32);
  P.zbar = pb->zbar;
  P.brange = pb->brange;
  P.bsim = pb->bsimmean;
  P.flags = effect_flags(pb->type_ids, k, kb, pb->dyn_flags);
  P.nrate = pb->nrate;
  if (P.nrate > 0) {
    // rate effects; the count rides in bits 16-18 for the LDS layout
    P.flags |= saom::F_RATE | (P.nrate << 16);
    for (int q = 0; q < P.nrate; ++q) {
      P.rkind[q] = pb->rkind[q];
      P.ralpha[q] = (double)pb->ralpha[q];
    }
  }
  int nd = 0;
  for (int e = 0; e < k; ++e) {
    P.tid[e] = pb->type_ids[e];
    P.ety[e] = pb->etypes ? pb->etypes[e] : 0;
    P.dyn[e] = pb->dyn_flags ? pb->dyn_flags[e] : 0;
    P.dci[e] = pb->dyc_idx ? pb->dyc_idx[e] : -1;
    P.c0[e] = pb->int0 ? pb->int0[e] : -1;
    P.c1[e] = pb->int1 ? pb->int1[e] : -1;
    P.c2[e] = pb->int2 ? pb->int2[e] : -1;
    P.p1[e] = pb->p1_64 ? pb->p1_64[e] : (pb->p1 ? (double)pb->p1[e] : 0.0);
    P.p2[e] = pb->p2_64 ? pb->p2_64[e] : (pb->p2 ? (double)pb->p2[e] : 0.0);
    if (P.dci[e] + 1 > nd) nd = P.dci[e] + 1;
  }
  for (int e = 0; e < kb; ++e) P.bty[e] = pb->bty[e];
  // a fixed arena (no allocation) rules out loading a specialised module
  pl->specialize = !replay && !(arena && !arena->grow) && saom::specialize_policy((long long)M * n);
  pl->nw = pick_nw(wmax, pb->threads);
  if (pb->threads <= 0) {
    const int tuned = cached_nw(ShapeKey{wmax, W64, k, P.flags});
    if (tuned > 0) pl->nw = tuned;
    pl->free_shape = tuned <= 0;
  }
  if (stream_rows) {
    // partial residency: the leading words of every row in LDS, as many as
    // fit beside the rest of the chain state (budgeted for the largest
    // block, 16 waves, so that any launch shape fits)
    const size_t base = saom::sim_lds_bytes(P, 16, wmax);
    if (base < kLdsLimit)
      P.lw = (int)std::min<size_t>((size_t)W64, (kLdsLimit - base) / ((size_t)n * 8));
  }
  if (saom::sim_lds_bytes(P, pl->nw, wmax) > kLdsLimit) {
    set_error("chain state exceeds the 160 KB LDS of a CU");
    return CUSNA_ERR_UNSUPPORTED;
  }
  const size_t stat_words = saom::net_stats_scratch_words(n, W64, saom::stats_need_cols(P.tid, k));

  // host-side staging
  const std::vector<uint64_t> x0 = rows64(pb->x0_out, n, W, W64, 0);
  const std::vector<uint64_t> xs = rows64(pb->sim_out ? pb->sim_out : pb->x0_out, n, W, W64, 0);
  const std::vector<uint64_t> kd = rows64(pb->kdist, n, W, W64, 1);
  const std::vector<uint64_t> pm = rows64(pb->perm, n, W, W64, 1);
  const std::vector<uint64_t> keep = rows64(pb->keep_out, n, W, W64, 1);
  const std::vector<uint64_t> force = rows64(pb->force_out, n, W, W64, 0);
  std::vector<uint64_t> inact;
  if (pb->has_comp && pb->inactw) {
    inact.assign(W64, 0ull);
    for (int w = 0; w < W; ++w) inact[w >> 1] |= (uint64_t)pb->inactw[w] << (32 * (w & 1));
    if (n & 63) inact[W64 - 1] &= (1ull << (n & 63)) - 1;
  }
  std::vector<double> vmat((size_t)k * n), wtab((size_t)k * (n + 1)), vb(n);
  for (size_t q = 0; q < vmat.size(); ++q)
    vmat[q] = pb->vmat64 ? pb->vmat64[q] : (pb->vmat ? (double)pb->vmat[q] : 0.0);
  for (size_t q = 0; q < wtab.size(); ++q)
    wtab[q] = pb->wtab64 ? pb->wtab64[q] : (pb->wtab ? (double)pb->wtab[q] : 0.0);
  for (int q = 0; q < n; ++q) vb[q] = pb->vb64 ? pb->vb64[q] : (pb->vb ? (double)pb->vb[q] : 0.0);
  std::vector<double> dyc;
  if (pb->dyc && nd > 0) {
    dyc.resize((size_t)nd * n * n);
    for (size_t q = 0; q < dyc.size(); ++q) dyc[q] = pb->dyc[q];
  }
  std::vector<double> ru;
  if (P.nrate > 0 && pb->ru) {
    ru.resize((size_t)P.nrate * n);
    for (size_t q = 0; q < ru.size(); ++q) ru[q] = pb->ru[q];
  }
  std::vector<int32_t> z0(n, 0), bkd(n, 1);
  for (int q = 0; q < n; ++q) {
    if (pb->z0) z0[q] = pb->z0[q];
    if (pb->bkdist) bkd[q] = pb->bkdist[q];
  }

  // device buffers: the inputs first, then the outputs the host reads, then
  // the dense end networks and the working buffers, so that a staged call
  // moves its inputs and its outputs with one copy each
  DeviceBlock& D = pl->D;
  const size_t rowsb = (size_t)n * W64 * 8;
  const size_t o_x0 = D.add(rowsb), o_xs = D.add(rowsb), o_kd = D.add(rowsb), o_pm = D.add(rowsb);
  pl->o_keep = D.add(rowsb);
  pl->o_force = D.add(rowsb);
  const size_t o_inact = D.add(W64 * 8);
  pl->o_th = D.add((size_t)M * k * 8);
  pl->o_lam = D.add((size_t)M * 8);
  pl->o_thb = D.add((size_t)M * (kb > 0 ? kb : 1) * 8);
  pl->o_lamb = D.add((size_t)M * 8);
  const size_t o_vmat = D.add(vmat.size() * 8), o_wtab = D.add(wtab.size() * 8);
  pl->o_vb = D.add((size_t)n * 8);
  const size_t o_dyc = D.add(dyc.size() * 8);
  const size_t o_z0 = D.add((size_t)n * 4), o_bkd = D.add((size_t)n * 4), o_act = D.add((size_t)n * 4);
  pl->o_vmiss = D.add((size_t)k * n);
  pl->o_od0 = D.add((size_t)n * 4);
  pl->o_id0 = D.add((size_t)n * 4);
  const size_t o_ru = D.add(ru.size() * 8);
  pl->o_params = D.add(sizeof(saom::SimParams));
  size_t o_lu0 = 0, o_llu = 0;
  if (replay) {
    o_lu0 = D.add((size_t)M * smax * 8);
    o_llu = D.add((size_t)M * smax * n * 8);
  }
  pl->in_bytes = D.bytes();
  const size_t o_ntim = D.add((size_t)M * 8), o_sc = D.add((size_t)M * (k + 1) * 8);
  const size_t o_bsc = D.add((size_t)M * (kb + 1) * 8);
  const size_t o_dist = D.add((size_t)M * 4), o_dstat = D.add((size_t)M * 4);
  const size_t o_steps = D.add((size_t)M * 4);
  const size_t o_bdist = D.add((size_t)M * 4), o_zend = D.add(kb > 0 ? (size_t)M * n * 4 : 0);
  pl->o_stats = D.add((size_t)M * k * 8);
  pl->o_rsc = D.add((size_t)M * P.nrate * 8);
  pl->o_adist = D.add(P.nrate > 0 ? (size_t)M * n * 4 : 0);
  pl->out_begin = pl->in_bytes;
  pl->out_end = D.bytes();
  if (want_dense) pl->o_dense = D.add((size_t)M * n * n);
  const size_t o_end = D.add((size_t)M * rowsb);
  const size_t o_work = D.add(stream_rows ? (size_t)M * rowsb : 0);
  pl->o_sscr = D.add((size_t)M * stat_words * 8);
  pl->has_sscr = stat_words > 0;
  // the staged one-shot call also fills the per-chain parameters from the
  // problem (the conversions of plan_set_params)
  const bool staged = arena && arena->staging;
  if (arena) {
    void* base = arena->grow ? arena->grow(D.bytes())
                             : (D.bytes() <= arena->capacity ? arena->base : nullptr);
    if (!base || !D.place(base, arena->grow ? D.bytes() : arena->capacity)) {
      if (arena->grow) {
        set_error("device memory of " + std::to_string(D.bytes()) + " bytes not available");
        return CUSNA_ERR_CUDA;
      }
      set_error("the call needs " + std::to_string(D.bytes()) + " bytes of device memory, above the " +
                std::to_string(arena->capacity) + " bytes it may use");
      return CUSNA_ERR_SHAPE;
    }
    pl->s = stream;
    if (staged) {
      pl->staging = static_cast<char*>(arena->staging(std::max(pl->in_bytes,
                                                               pl->out_end - pl->out_begin)));
      if (!pl->staging) {
        set_error("pinned host memory not available");
        return CUSNA_ERR_CUDA;
      }
    }
  } else {
    HIP_TRY(D.alloc());
    if (stream) {
      pl->s = stream;
    } else {
      HIP_TRY(hipStreamCreateWithFlags(&pl->s, hipStreamNonBlocking));
      pl->own_stream = true;
    }
    HIP_TRY(hipEventCreate(&pl->ev0));
    HIP_TRY(hipEventCreate(&pl->ev1));
  }
  hipStream_t s = pl->s;
  // an upload: into the pinned staging buffer (copied with the others at the
  // end) or one asynchronous copy of its own
  auto up = [&](size_t off, const void* src, size_t bytes) -> hipError_t {
    if (!bytes) return hipSuccess;
    if (staged) {
      std::memcpy(pl->staging + off, src, bytes);
      return hipSuccess;
    }
    return hipMemcpyAsync(D.at<char>(off), src, bytes, hipMemcpyHostToDevice, s);
  };
  HIP_TRY(up(o_x0, x0.data(), rowsb));
  HIP_TRY(up(o_xs, xs.data(), rowsb));
  HIP_TRY(up(o_kd, kd.data(), rowsb));
  HIP_TRY(up(o_pm, pm.data(), rowsb));
  HIP_TRY(up(pl->o_keep, keep.data(), rowsb));
  HIP_TRY(up(pl->o_force, force.data(), rowsb));
  if (!inact.empty()) HIP_TRY(up(o_inact, inact.data(), W64 * 8));
  HIP_TRY(up(o_vmat, vmat.data(), vmat.size() * 8));
  HIP_TRY(up(o_wtab, wtab.data(), wtab.size() * 8));
  HIP_TRY(up(pl->o_vb, vb.data(), (size_t)n * 8));
  if (!dyc.empty()) HIP_TRY(up(o_dyc, dyc.data(), dyc.size() * 8));
  HIP_TRY(up(o_ru, ru.data(), ru.size() * 8));
  HIP_TRY(up(o_z0, z0.data(), (size_t)n * 4));
  HIP_TRY(up(o_bkd, bkd.data(), (size_t)n * 4));
  if (pb->act) HIP_TRY(up(o_act, pb->act, (size_t)P.nact * 4));
  if (pb->vmiss) HIP_TRY(up(pl->o_vmiss, pb->vmiss, (size_t)k * n));
  if (pb->od0) HIP_TRY(up(pl->o_od0, pb->od0, (size_t)n * 4));
  if (pb->id0) HIP_TRY(up(pl->o_id0, pb->id0, (size_t)n * 4));
  pl->has_vmiss = pb->vmiss != nullptr;
  pl->has_od0 = pb->od0 != nullptr;
  pl->has_id0 = pb->id0 != nullptr;
  if (replay) {
    HIP_TRY(up(o_lu0, log_u0, (size_t)M * smax * 8));
    HIP_TRY(up(o_llu, loglog_u, (size_t)M * smax * n * 8));
  }
  if (staged) {
    double* th = reinterpret_cast<double*>(pl->staging + pl->o_th);
    double* lm = reinterpret_cast<double*>(pl->staging + pl->o_lam);
    double* thb = reinterpret_cast<double*>(pl->staging + pl->o_thb);
    double* lmb = reinterpret_cast<double*>(pl->staging + pl->o_lamb);
    for (size_t q = 0; q < (size_t)M * k; ++q) th[q] = pb->theta[q];
    for (int c = 0; c < M; ++c) {
      lm[c] = pb->lam ? (double)pb->lam[c] : 1.0;
      lmb[c] = kb > 0 ? (double)pb->lamb[c] : 0.0;
    }
    for (size_t q = 0; q < (size_t)M * (kb > 0 ? kb : 1); ++q)
      thb[q] = kb > 0 ? (double)pb->thetab[q] : 0.0;
    pl->params_set = true;
  }
  P.x0 = D.at<uint64_t>(o_x0);
  P.xs = D.at<uint64_t>(o_xs);
  P.kd = D.at<uint64_t>(o_kd);
  P.pm = D.at<uint64_t>(o_pm);
  P.inactw = inact.empty() ? nullptr : D.at<uint64_t>(o_inact);
  P.theta = D.at<double>(pl->o_th);
  P.lam = D.at<double>(pl->o_lam);
  P.th_stride = k;
  P.thb_stride = kb;
  P.lam_stride = 1;
  P.keys = nullptr;
  P.thetab = D.at<double>(pl->o_thb);
  P.lamb = D.at<double>(pl->o_lamb);
  P.vmat = D.at<double>(o_vmat);
  P.wtab = D.at<double>(o_wtab);
  P.dyc = dyc.empty() ? nullptr : D.at<double>(o_dyc);
  P.z0 = D.at<int32_t>(o_z0);
  P.bkd = D.at<int32_t>(o_bkd);
  P.act = pb->act ? D.at<int32_t>(o_act) : nullptr;
  P.lu0 = replay ? D.at<double>(o_lu0) : nullptr;
  P.llu = replay ? D.at<double>(o_llu) : nullptr;
  P.end_rows = D.at<uint64_t>(o_end);
  P.ntim = D.at<double>(o_ntim);
  P.scores = D.at<double>(o_sc);
  P.bscores = D.at<double>(o_bsc);
  P.dist = D.at<int32_t>(o_dist);
  P.dist_stat = D.at<int32_t>(o_dstat);
  P.steps = D.at<int32_t>(o_steps);
  P.bdist = D.at<int32_t>(o_bdist);
  P.z_end = kb > 0 ? D.at<int32_t>(o_zend) : nullptr;
  // written by the reference only when nrate > 0 (:406-409), with the rate scores
  P.actor_dist = P.nrate > 0 ? D.at<int32_t>(pl->o_adist) : nullptr;
  P.ru = ru.empty() ? nullptr : D.at<double>(o_ru);
  P.rscores = P.nrate > 0 ? D.at<double>(pl->o_rsc) : nullptr;
  P.work = stream_rows ? D.at<uint64_t>(o_work) : nullptr;
  HIP_TRY(up(pl->o_params, &P, sizeof(P)));
  if (staged) {
    // every input in one copy (the staging buffer outlives the call)
    HIP_TRY(hipMemcpyAsync(D.at<char>(0), pl->staging, pl->in_bytes, hipMemcpyHostToDevice, s));
  } else {
    HIP_TRY(hipStreamSynchronize(s));  // pageable sources: the copies have consumed them
  }
  *out = pl.release();
  return CUSNA_OK;
}

cusna_status plan_set_params(cusna_period_plan* pl, const float* theta, const float* lam,
                             const float* thetab, const float* lamb) {
  const int M = pl->M, k = pl->k, kb = pl->kb;
  std::vector<double> th((size_t)M * k), lm(M, 1.0), thb((size_t)M * (kb > 0 ? kb : 1), 0.0),
      lmb(M, 0.0);
  for (size_t q = 0; q < th.size(); ++q) th[q] = theta[q];
  if (lam)
    for (int c = 0; c < M; ++c) lm[c] = lam[c];
  if (kb > 0) {
    for (size_t q = 0; q < (size_t)M * kb; ++q) thb[q] = thetab[q];
    for (int c = 0; c < M; ++c) lmb[c] = lamb[c];
  }
  HIP_TRY(hipMemcpyAsync(pl->D.at<char>(pl->o_th), th.data(), th.size() * 8,
                         hipMemcpyHostToDevice, pl->s));
  HIP_TRY(hipMemcpyAsync(pl->D.at<char>(pl->o_lam), lm.data(), (size_t)M * 8,
                         hipMemcpyHostToDevice, pl->s));
  HIP_TRY(hipMemcpyAsync(pl->D.at<char>(pl->o_thb), thb.data(), thb.size() * 8,
                         hipMemcpyHostToDevice, pl->s));
  HIP_TRY(hipMemcpyAsync(pl->D.at<char>(pl->o_lamb), lmb.data(), (size_t)M * 8,
                         hipMemcpyHostToDevice, pl->s));
  // pageable host sources: make sure the copies have consumed them
  HIP_TRY(hipStreamSynchronize(pl->s));
  return CUSNA_OK;
}

cusna_status plan_run(cusna_period_plan* pl, uint64_t seed, bool want_stats, int M_run,
                      hipStream_t stream, bool timed) {
  saom::SimParams P = pl->P;  // host copy: grid size and LDS of this launch
  P.key0 = (uint32_t)(seed & 0xFFFFFFFFu);
  P.key1 = (uint32_t)(seed >> 32);
  if (M_run > 0) P.M = M_run;
  hipStream_t s = stream ? stream : pl->s;
  timed = timed && pl->ev0 && pl->ev1;  // a plan placed in an arena has no events
  if (timed) HIP_TRY(hipEventRecord(pl->ev0, s));
  int le = saom::launch_period(P, pl->D.at<saom::SimParams>(pl->o_params),
                               saom::LaunchShape{pl->nw, pl->wmax}, pl->replay, P.key0, P.key1,
                               s, pl->specialize, &pl->last_rtc, pl->free_shape,
                               &pl->last_threads);
  if (le != 0) {
    set_error(std::string("period kernel launch: ") + hipGetErrorString((hipError_t)le));
    return CUSNA_ERR_CUDA;
  }
  if (timed) HIP_TRY(hipEventRecord(pl->ev1, s));
  pl->stats_valid = false;
  if (want_stats) return plan_stats(pl, P.M, s);
  return CUSNA_OK;
}

cusna_status plan_stats(cusna_period_plan* pl, int M_run, hipStream_t s) {
  const saom::SimParams& P = pl->P;
  {
    saom::StatParams SP;
    std::memset(&SP, 0, sizeof(SP));
    SP.n = pl->n;
    SP.k = pl->k;
    SP.M = M_run;
    SP.W64 = pl->W64;
    SP.symflag = pl->symflag;
    for (int e = 0; e < pl->k; ++e) {
      SP.tid[e] = P.tid[e];
      SP.dyn[e] = P.dyn[e];
      SP.p1[e] = P.p1[e];
      SP.p2[e] = P.p2[e];
    }
    SP.rows = P.end_rows;
    SP.keep = pl->D.at<uint64_t>(pl->o_keep);
    SP.force = pl->D.at<uint64_t>(pl->o_force);
    SP.vmat = P.vmat;
    SP.vb = pl->D.at<double>(pl->o_vb);
    SP.wtab = P.wtab;
    SP.vmiss = pl->has_vmiss ? pl->D.at<uint8_t>(pl->o_vmiss) : nullptr;
    // start degrees pass through as the reference passes them (:420)
    SP.od0 = pl->has_od0 ? pl->D.at<int32_t>(pl->o_od0) : nullptr;
    SP.id0 = pl->has_id0 ? pl->D.at<int32_t>(pl->o_id0) : nullptr;
    SP.out = pl->D.at<double>(pl->o_stats);
    SP.scratch = pl->has_sscr ? pl->D.at<uint64_t>(pl->o_sscr) : nullptr;
    const int le = saom::launch_net_stats(SP, s);
    if (le != 0) {
      set_error(std::string("stats kernel launch: ") + hipGetErrorString((hipError_t)le));
      return CUSNA_ERR_CUDA;
    }
    pl->stats_valid = true;
  }
  return CUSNA_OK;
}

cusna_status plan_fetch(cusna_period_plan* pl, cusna_saom_result* r) {
  if (!r) return CUSNA_ERR_NULL_ARG;
  if (r->struct_version != CUSNA_ABI_VERSION) return CUSNA_ERR_SHAPE;
  const int n = pl->n, k = pl->k, kb = pl->kb, M = pl->M;
  const saom::SimParams& P = pl->P;
  hipStream_t s = pl->s;
  const bool dense = r->want_end_nets && r->a_ends;
  if (dense) {
    if (!pl->want_dense) {
      set_error("plan was created without end-network output");
      return CUSNA_ERR_UNSUPPORTED;
    }
    if (saom::launch_expand_dense(P.end_rows, n, pl->W64, M, pl->D.at<uint8_t>(pl->o_dense), s))
      return CUSNA_ERR_CUDA;
  }
  auto down = [&](void* dst, const void* src, size_t bytes) -> hipError_t {
    if (!dst || !bytes) return hipSuccess;
    return hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s);
  };
  if (pl->staging) {
    // a staged call: the outputs the host reads in one copy, then unpacked
    const char* dev0 = pl->D.at<char>(pl->out_begin);
    HIP_TRY(hipMemcpyAsync(pl->staging, dev0, pl->out_end - pl->out_begin,
                           hipMemcpyDeviceToHost, s));
    if (dense) HIP_TRY(down(r->a_ends, pl->D.at<char>(pl->o_dense), (size_t)M * n * n));
    HIP_TRY(hipStreamSynchronize(s));
    auto host = [&](const void* dev) -> const char* {
      return pl->staging + (static_cast<const char*>(dev) - dev0);
    };
    auto take = [&](void* dst, const void* dev, size_t bytes) {
      if (dst && bytes) std::memcpy(dst, host(dev), bytes);
    };
    auto take_f32 = [&](float* dst, const void* dev, size_t count) {
      if (!dst) return;
      const double* v = reinterpret_cast<const double*>(host(dev));
      for (size_t q = 0; q < count; ++q) dst[q] = (float)v[q];
    };
    if (pl->stats_valid) take(r->stats, pl->D.at<char>(pl->o_stats), (size_t)M * k * 8);
    take(r->ntim, P.ntim, (size_t)M * 8);
    take_f32(r->scores, P.scores, (size_t)M * (k + 1));
    take(r->dist, P.dist, (size_t)M * 4);
    take(r->dist_stat, P.dist_stat, (size_t)M * 4);
    take(r->steps, P.steps, (size_t)M * 4);
    if (kb > 0) {
      take(r->bdist, P.bdist, (size_t)M * 4);
      take_f32(r->bscores, P.bscores, (size_t)M * (kb + 1));
      take(r->z_end, P.z_end, (size_t)M * n * 4);
    }
    if (P.nrate > 0) {
      take_f32(r->rscores, P.rscores, (size_t)M * P.nrate);
      take(r->actor_dist, P.actor_dist, (size_t)M * n * 4);
    }
    return CUSNA_OK;
  }
  std::vector<double> sc((size_t)M * (k + 1)), bsc((size_t)M * (kb + 1));
  if (r->stats && pl->stats_valid)
    HIP_TRY(down(r->stats, pl->D.at<char>(pl->o_stats), (size_t)M * k * 8));
  HIP_TRY(down(r->ntim, P.ntim, (size_t)M * 8));
  HIP_TRY(down(sc.data(), P.scores, sc.size() * 8));
  HIP_TRY(down(r->dist, P.dist, (size_t)M * 4));
  HIP_TRY(down(r->dist_stat, P.dist_stat, (size_t)M * 4));
  HIP_TRY(down(r->steps, P.steps, (size_t)M * 4));
  if (kb > 0) {
    HIP_TRY(down(r->bdist, P.bdist, (size_t)M * 4));
    HIP_TRY(down(bsc.data(), P.bscores, bsc.size() * 8));
    HIP_TRY(down(r->z_end, P.z_end, (size_t)M * n * 4));
  }
  if (dense) HIP_TRY(down(r->a_ends, pl->D.at<char>(pl->o_dense), (size_t)M * n * n));
  std::vector<double> rsc((size_t)M * P.nrate);
  if (P.nrate > 0) {
    if (r->rscores) HIP_TRY(down(rsc.data(), P.rscores, rsc.size() * 8));
    HIP_TRY(down(r->actor_dist, P.actor_dist, (size_t)M * n * 4));
  }
  HIP_TRY(hipStreamSynchronize(s));
  if (P.nrate > 0 && r->rscores)
    for (size_t q = 0; q < rsc.size(); ++q) r->rscores[q] = (float)rsc[q];
  if (r->scores)
    for (size_t q = 0; q < sc.size(); ++q) r->scores[q] = (float)sc[q];
  if (kb > 0 && r->bscores)
    for (size_t q = 0; q < bsc.size(); ++q) r->bscores[q] = (float)bsc[q];
  return CUSNA_OK;
}

// Device and pinned host memory of the one-shot period call, cached per host
// thread (grown on demand, kept for the thread's later calls), so that a call
// makes no allocation once the thread has made one of its size.
struct CallBuffers {
  int dev = -1;
  void* d = nullptr;
  size_t dcap = 0;
  void* h = nullptr;
  size_t hcap = 0;
  ~CallBuffers() {
    if (h) (void)hipHostFree(h);
    if (d) (void)hipFree(d);
  }
};
thread_local CallBuffers g_call;

CallBuffers& call_buffers() {
  int dev = 0;
  (void)hipGetDevice(&dev);
  if (g_call.dev != dev) {  // another device: its buffers are of no use here
    if (g_call.d) (void)hipFree(g_call.d);
    g_call.d = nullptr;
    g_call.dcap = 0;
    g_call.dev = dev;
  }
  return g_call;
}

void* call_device(size_t bytes) {
  CallBuffers& b = call_buffers();
  if (bytes > b.dcap) {
    if (b.d) (void)hipFree(b.d);
    b.d = nullptr;
    b.dcap = 0;
    const size_t want = std::max(bytes, b.dcap + b.dcap / 2);
    if (hipMalloc(&b.d, want) != hipSuccess) {
      (void)hipGetLastError();
      return nullptr;
    }
    b.dcap = want;
  }
  return b.d;
}

void* call_host(size_t bytes) {
  CallBuffers& b = g_call;
  if (bytes > b.hcap) {
    if (b.h) (void)hipHostFree(b.h);
    b.h = nullptr;
    b.hcap = 0;
    const size_t want = std::max(bytes, b.hcap + b.hcap / 2);
    if (hipHostMalloc(&b.h, want, hipHostMallocDefault) != hipSuccess) {
      (void)hipGetLastError();
      return nullptr;
    }
    b.hcap = want;
  }
  return b.h;
}

cusna_status simulate(const cusna_saom_problem* pb, cusna_saom_result* r, const double* log_u0,
                      const double* loglog_u, int32_t smax) {
  cusna_status st = validate(pb, r);
  if (st != CUSNA_OK) return st;
  cusna_period_plan* raw = nullptr;
  // one-shot calls run on the thread's cached buffers, inputs and outputs
  // staged through pinned memory (the replay variant's recorded noise is
  // large and passes through its own copies)
  DeviceArena cached;
  cached.grow = call_device;
  cached.staging = call_host;
  st = plan_create(pb, r->want_end_nets && r->a_ends, log_u0, loglog_u, smax, thread_stream(),
                   &raw, log_u0 ? nullptr : &cached);
  if (st != CUSNA_OK) return st;
  std::unique_ptr<cusna_period_plan> pl(raw);
  if (!pl->params_set) {
    st = plan_set_params(pl.get(), pb->theta, pb->lam, pb->thetab, pb->lamb);
    if (st != CUSNA_OK) return st;
  }
  st = plan_run(pl.get(), pb->seed, r->stats != nullptr, -1, nullptr, true);
  if (st != CUSNA_OK) return st;
  return plan_fetch(pl.get(), r);
}

// ---- one period without device allocation
// The device memory of cusna_saom_simulate_period_noalloc: one array per GPU,
// which each GPU holds from the load of the library's code object until the
// process ends. Host code never passes the array's name to a kernel or a copy
// (on the host it is a stand-in, not a device address): the address comes
// from hipGetSymbolAddress, once per device, with that device current.
__device__ __align__(256) unsigned char scratch[CUSNA_NOALLOC_SCRATCH_BYTES];

constexpr int kNoAllocDevices = 64;
struct NoAllocDevice {
  std::mutex mu;  // calls on one device run one at a time
  void* ptr = nullptr;
};
NoAllocDevice g_noalloc[kNoAllocDevices];

cusna_status simulate_noalloc(const cusna_saom_problem* pb, cusna_saom_result* r) {
  cusna_status st = validate(pb, r);
  if (st != CUSNA_OK) return st;
  if (device_count() < 1) {
    set_error("no HIP device");
    return CUSNA_ERR_NO_CUDA;
  }
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= kNoAllocDevices) {
    set_error("cusna_saom_simulate_period_noalloc: device index above 63");
    return CUSNA_ERR_UNSUPPORTED;
  }
  NoAllocDevice& S = g_noalloc[dev];
  std::lock_guard<std::mutex> lk(S.mu);
  if (!S.ptr) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      set_error(std::string("cusna_saom_simulate_period_noalloc: hipGetSymbolAddress: ") +
                hipGetErrorString(e));
      return CUSNA_ERR_CUDA;
    }
    S.ptr = p;
  }
  const DeviceArena arena{S.ptr, (size_t)CUSNA_NOALLOC_SCRATCH_BYTES};
  cusna_period_plan* raw = nullptr;
  // the null stream: no stream is created; calls on one device are serialised
  st = plan_create(pb, r->want_end_nets && r->a_ends, nullptr, nullptr, 0, nullptr, &raw, &arena);
  if (st != CUSNA_OK) {
    if (st == CUSNA_ERR_SHAPE)
      set_error("cusna_saom_simulate_period_noalloc: " + g_last_error);
    return st;
  }
  std::unique_ptr<cusna_period_plan> pl(raw);
  st = plan_set_params(pl.get(), pb->theta, pb->lam, pb->thetab, pb->lamb);
  if (st != CUSNA_OK) return st;
  st = plan_run(pl.get(), pb->seed, r->stats != nullptr, -1, nullptr, false);
  if (st != CUSNA_OK) return st;
  return plan_fetch(pl.get(), r);
}

// Several independent problems ("groups": networks, periods or parameter
// sets) in as few period-kernel launches as there are distinct launch shapes.
// Within a launch the items run longest-first: groups sorted by an estimate of
// their per-chain work, so that the hardware's in-order block dispatch keeps
// the CUs busy until the end.
// Streams (default priority) and events of one multi-group call, released
// when the call returns.
constexpr int kGroupStreams = 4;
struct GroupStreams {
  hipStream_t st[kGroupStreams] = {};
  hipEvent_t done[kGroupStreams] = {};
  std::vector<hipEvent_t> ready;  // one per bucket: its inputs are on the device
  int n = 0;
  hipError_t create(int count) {
    hipError_t e = hipSuccess;
    for (; e == hipSuccess && n < count; ++n) {
      e = hipStreamCreateWithFlags(&st[n], hipStreamNonBlocking);
      if (e == hipSuccess) e = hipEventCreateWithFlags(&done[n], hipEventDisableTiming);
    }
    return e;
  }
  hipError_t record_ready(hipStream_t s, hipEvent_t* out) {
    hipEvent_t e = nullptr;
    hipError_t r = hipEventCreateWithFlags(&e, hipEventDisableTiming);
    if (r != hipSuccess) return r;
    ready.push_back(e);
    *out = e;
    return hipEventRecord(e, s);
  }
  ~GroupStreams() {
    for (int q = 0; q < kGroupStreams; ++q) {
      if (st[q]) (void)hipStreamSynchronize(st[q]);
      if (done[q]) (void)hipEventDestroy(done[q]);
      if (st[q]) (void)hipStreamDestroy(st[q]);
    }
    for (hipEvent_t e : ready) (void)hipEventDestroy(e);
  }
};

// The launch shape plan_create gives pb's plan (waves per chain, word
// bound, words), computed without building the plan.
struct PlanShape {
  int nw, wmax, W64;
};
PlanShape plan_shape(const cusna_saom_problem* pb) {
  PlanShape sh;
  sh.W64 = (pb->n + 63) / 64;
  sh.wmax = wmax_for(sh.W64);
  sh.nw = pick_nw(sh.wmax, pb->threads);
  if (pb->threads <= 0) {
    int flags = effect_flags(pb->type_ids, pb->k, pb->kb, pb->dyn_flags);
    if (pb->nrate > 0) flags |= saom::F_RATE | (pb->nrate << 16);
    const int tuned = cached_nw(ShapeKey{sh.wmax, sh.W64, pb->k, flags});
    if (tuned > 0) sh.nw = tuned;
  }
  return sh;
}

cusna_status simulate_groups(const cusna_saom_problem* const* pbs, int G,
                             cusna_saom_result* const* rs) {
  hipStream_t s = thread_stream();
  for (int g = 0; g < G; ++g) {
    if (!pbs[g] || !rs[g]) return CUSNA_ERR_NULL_ARG;
    cusna_status st = validate(pbs[g], rs[g]);
    if (st != CUSNA_OK) return st;
    if (pbs[g]->M > (1 << 20)) {
      set_error("cusna_saom_simulate_groups: at most 2^20 chains per group");
      return CUSNA_ERR_SHAPE;
    }
  }
  // buckets of one launch shape
  // (one word count per bucket, so a specialised launch can use the exact
  // word bound of its networks)
  struct Bucket {
    int nw, wmax, W64;
    bool simple;
    std::vector<int> groups;
  };
  std::vector<Bucket> buckets;
  std::vector<PlanShape> shape(G);
  for (int g = 0; g < G; ++g) {
    shape[g] = plan_shape(pbs[g]);
    const PlanShape& sh = shape[g];
    const bool simple = saom::simple_effects(pbs[g]->type_ids, pbs[g]->k);
    Bucket* b = nullptr;
    for (auto& x : buckets)
      if (x.nw == sh.nw && x.wmax == sh.wmax && x.W64 == sh.W64 && x.simple == simple) b = &x;
    if (!b) {
      buckets.push_back(Bucket{sh.nw, sh.wmax, sh.W64, simple, {}});
      b = &buckets.back();
    }
    b->groups.push_back(g);
  }
  // expected ministeps x alternatives x words per chain
  auto cost = [&](int g) {
    const cusna_saom_problem* pb = pbs[g];
    const double lam = pb->lam ? (double)pb->lam[0] : 1.0;
    const double nact = pb->act ? pb->nact : pb->n;
    const double steps = pb->conditional ? 2.0 * pb->target_dist + 1.0 : lam * nact + 1.0;
    // a chain whose rows live (partly) in device memory runs about 1.3 x an
    // LDS-resident chain of the same size (measured on the 256-group batch)
    const double streamed = shape[g].wmax > 16 ? 1.3 : 1.0;
    return streamed * steps * pb->n * (shape[g].W64 + pb->k + 8.0);
  };
  std::vector<double> work_b(buckets.size(), 0.0);
  for (size_t bi = 0; bi < buckets.size(); ++bi) {
    Bucket& b = buckets[bi];
    std::stable_sort(b.groups.begin(), b.groups.end(),
                     [&](int a, int c) { return cost(a) > cost(c); });
    for (int g : b.groups) work_b[bi] += cost(g) * pbs[g]->M;
  }
  std::vector<size_t> order(buckets.size());
  for (size_t bi = 0; bi < order.size(); ++bi) order[bi] = bi;
  std::stable_sort(order.begin(), order.end(),
                   [&](size_t a, size_t c) { return work_b[a] > work_b[c]; });

  // One parameter slot per group in launch order, so that a bucket's
  // parameters and work items are contiguous; items are (slot << 20 | chain).
  size_t n_items = 0;
  for (int g = 0; g < G; ++g) n_items += (size_t)pbs[g]->M;
  DeviceBlock D;
  const size_t o_grp = D.add((size_t)G * sizeof(saom::SimParams)), o_items = D.add(n_items * 4);
  HIP_TRY(D.alloc());
  std::vector<saom::SimParams> grp(G);
  std::vector<uint32_t> items;
  items.reserve(n_items);
  std::vector<std::unique_ptr<cusna_period_plan>> plans(G);

  // The buckets run concurrently on up to four streams (largest first, each
  // onto the least-loaded stream), so one bucket's tail and a CU's spare LDS
  // and wave slots take work from another; chains are independent, results
  // do not depend on the order. A bucket's plans are built (and their inputs
  // uploaded on the calling thread's stream) just before its launch, so that
  // the set-up of the later buckets overlaps the earlier buckets' periods.
  GroupStreams gs;
  const int nst = (int)std::min<size_t>(kGroupStreams, buckets.size());
  HIP_TRY(gs.create(nst));
  std::vector<double> queued(nst, 0.0);
  int slot = 0;
  for (size_t o = 0; o < order.size(); ++o) {
    const size_t bi = order[o];
    Bucket& b = buckets[bi];
    const int slot0 = slot;
    const size_t first = items.size();
    for (int g : b.groups) {
      cusna_period_plan* raw = nullptr;
      cusna_status st =
          plan_create(pbs[g], rs[g]->want_end_nets && rs[g]->a_ends, nullptr, nullptr, 0, s, &raw);
      if (st != CUSNA_OK) return st;
      plans[g].reset(raw);
      st = plan_set_params(raw, pbs[g]->theta, pbs[g]->lam, pbs[g]->thetab, pbs[g]->lamb);
      if (st != CUSNA_OK) return st;
      if (raw->nw != b.nw || raw->wmax != b.wmax || raw->W64 != b.W64) {
        set_error("cusna_saom_simulate_groups: plan shape differs from its bucket");
        return CUSNA_ERR_SHAPE;
      }
      grp[slot] = raw->P;
      for (int c = 0; c < raw->M; ++c) items.push_back(((uint32_t)slot << 20) | (uint32_t)c);
      ++slot;
    }
    const int ng = slot - slot0;
    HIP_TRY(hipMemcpyAsync(D.at<saom::SimParams>(o_grp) + slot0, grp.data() + slot0,
                           (size_t)ng * sizeof(saom::SimParams), hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(D.at<uint32_t>(o_items) + first, items.data() + first,
                           (items.size() - first) * 4, hipMemcpyHostToDevice, s));
    hipEvent_t ready = nullptr;
    HIP_TRY(gs.record_ready(s, &ready));

    size_t lds = 0, lds_exact = 0, lds_alt = 0;
    bool uniform = true, free_shape = true;
    long long work = 0;
    const cusna_period_plan* lead = plans[b.groups[0]].get();
    for (int g : b.groups) {
      uniform = uniform && saom::same_effects(plans[g]->P, lead->P);
      free_shape = free_shape && plans[g]->free_shape;
      work += (long long)plans[g]->M * plans[g]->n;
      lds = std::max(lds, saom::sim_lds_bytes(plans[g]->P, b.nw, b.wmax));
    }
    const int kf = uniform && free_shape ? lead->P.k : 0;
    const saom::LaunchShape gen{b.nw, b.wmax};
    const saom::LaunchShape ex = saom::exact_shape(b.W64, gen, kf, lds);
    const saom::LaunchShape alt = saom::exact_shape(b.W64, gen, kf);
    for (int g : b.groups) {
      lds_exact = std::max(lds_exact, saom::sim_lds_bytes(plans[g]->P, ex.nw, ex.wmax));
      lds_alt = std::max(lds_alt, saom::sim_lds_bytes(plans[g]->P, alt.nw, alt.wmax));
    }
    saom::PeriodLaunch L;
    L.Pd = D.at<saom::SimParams>(o_grp);
    L.items = D.at<uint32_t>(o_items) + first;
    L.grid = (int)(items.size() - first);
    L.lds = lds;
    L.key0 = L.key1 = 0u;
    L.Ph = &lead->P;
    L.specialize = uniform && saom::specialize_policy(work);
    for (int g : b.groups) L.same_n = L.same_n && plans[g]->n == lead->n;
    L.exact = ex;
    L.lds_exact = lds_exact;
    L.exact_alt = alt;
    L.lds_exact_alt = lds_alt;

    // longest bucket first onto the stream with the least work queued
    // (estimated as the chains' expected ministeps x alternatives x words),
    // so that no stream is left with a long tail of its own
    const int q = (int)(std::min_element(queued.begin(), queued.end()) - queued.begin());
    queued[q] += work_b[bi];
    HIP_TRY(hipStreamWaitEvent(gs.st[q], ready, 0));
    const int le = saom::launch_period_ex(L, saom::LaunchShape{b.nw, b.wmax}, false, b.simple,
                                          gs.st[q]);
    if (le != 0) {
      set_error(std::string("period kernel launch: ") + hipGetErrorString((hipError_t)le));
      return CUSNA_ERR_CUDA;
    }
    // the bucket's statistics follow it on its stream, so that they overlap
    // the other buckets' periods
    for (int g : b.groups) {
      if (rs[g]->stats) {
        cusna_status st = plan_stats(plans[g].get(), plans[g]->M, gs.st[q]);
        if (st != CUSNA_OK) return st;
      }
    }
  }
  for (int q = 0; q < nst; ++q) {
    HIP_TRY(hipEventRecord(gs.done[q], gs.st[q]));
    HIP_TRY(hipStreamWaitEvent(s, gs.done[q], 0));
  }
  for (int g = 0; g < G; ++g) {
    cusna_status st = plan_fetch(plans[g].get(), rs[g]);
    if (st != CUSNA_OK) return st;
  }
  HIP_TRY(hipStreamSynchronize(s));
  return CUSNA_OK;
}

// Time the period kernel of `pl` in every compiled block shape on its first
// `probe` chains (one untimed and one timed launch each) and keep the fastest.
cusna_status autotune(cusna_period_plan* pl, int probe, int* best_threads) {
  if (pl->replay) return CUSNA_ERR_UNSUPPORTED;
  if (probe <= 0) probe = std::min(pl->M, 2048);
  probe = std::min(probe, pl->M);
  const int keep = pl->nw;
  const bool keep_free = pl->free_shape;
  pl->free_shape = false;  // time exactly the shapes tried
  int best = keep;
  double best_ms = 1e300;
  for (int nw = 1; nw <= 16; nw *= 2) {
    if (nw > pl->wmax || (pl->wmax == 16 && nw < 2)) continue;
    if (saom::sim_lds_bytes(pl->P, nw, pl->wmax) > kLdsLimit) continue;
    pl->nw = nw;
    float ms = 0.0f;
    for (int rep = 0; rep < 2; ++rep) {
      cusna_status st = plan_run(pl, 0x5eedull + (uint64_t)rep, false, probe, nullptr, true);
      if (st != CUSNA_OK) {
        pl->nw = keep;
        pl->free_shape = keep_free;
        return st;
      }
      HIP_TRY(hipEventSynchronize(pl->ev1));
      HIP_TRY(hipEventElapsedTime(&ms, pl->ev0, pl->ev1));
    }
    if (ms < best_ms) {
      best_ms = ms;
      best = nw;
    }
  }
  pl->nw = best;
  pl->last_threads = 0;  // the shape query reports the choice until the next run
  {
    std::lock_guard<std::mutex> lk(g_shape_mu);
    g_shape_cache[ShapeKey{pl->wmax, pl->W64, pl->k, pl->P.flags}] = best;
  }
  if (best_threads) *best_threads = 64 * best;
  return CUSNA_OK;
}

cusna_status net_stats_packed(const uint64_t* rows64_host, int n, int M, int k,
                              const int32_t* type_ids, const double* p1, const double* p2,
                              const double* vmat, const uint8_t* vmiss, const int32_t* dyn_flags,
                              const double* vb, const int32_t* od0, const int32_t* id0,
                              const double* wtab, int symflag, const uint64_t* keep,
                              const uint64_t* force, double* out) {
  if (device_count() < 1) {
    set_error("no HIP device");
    return CUSNA_ERR_NO_CUDA;
  }
  if (k > CUSNA_MAXEFF) return CUSNA_ERR_SHAPE;
  const int W64 = (n + 63) / 64;
  const size_t stat_words =
      saom::net_stats_scratch_words(n, W64, saom::stats_need_cols(type_ids, k));
  saom::StatParams SP;
  std::memset(&SP, 0, sizeof(SP));
  SP.n = n;
  SP.k = k;
  SP.M = M;
  SP.W64 = W64;
  SP.symflag = symflag;
  for (int e = 0; e < k; ++e) {
    SP.tid[e] = type_ids[e];
    SP.dyn[e] = dyn_flags ? dyn_flags[e] : 0;
    SP.p1[e] = p1 ? p1[e] : 0.0;
    SP.p2[e] = p2 ? p2[e] : 0.0;
  }
  DeviceBlock D;
  const size_t rowsb = (size_t)n * W64 * 8;
  const size_t o_rows = D.add((size_t)M * rowsb);
  const size_t o_keep = D.add(keep ? rowsb : 0), o_force = D.add(force ? rowsb : 0);
  const size_t o_vmat = D.add((size_t)k * n * 8), o_vb = D.add((size_t)n * 8);
  const size_t o_wtab = D.add((size_t)k * (n + 1) * 8), o_vmiss = D.add((size_t)k * n);
  const size_t o_od0 = D.add((size_t)n * 4), o_id0 = D.add((size_t)n * 4);
  const size_t o_out = D.add((size_t)M * k * 8);
  const size_t o_scr = D.add((size_t)M * stat_words * 8);
  HIP_TRY(D.alloc());
  hipStream_t s = thread_stream();
  std::vector<double> zk((size_t)k * (n + 1), 0.0);
  HIP_TRY(hipMemcpyAsync(D.at<char>(o_rows), rows64_host, (size_t)M * rowsb,
                         hipMemcpyHostToDevice, s));
  if (keep) HIP_TRY(hipMemcpyAsync(D.at<char>(o_keep), keep, rowsb, hipMemcpyHostToDevice, s));
  if (force) HIP_TRY(hipMemcpyAsync(D.at<char>(o_force), force, rowsb, hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(D.at<char>(o_vmat), vmat ? vmat : zk.data(), (size_t)k * n * 8,
                         hipMemcpyHostToDevice, s));
  if (vb) HIP_TRY(hipMemcpyAsync(D.at<char>(o_vb), vb, (size_t)n * 8, hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(D.at<char>(o_wtab), wtab ? wtab : zk.data(), (size_t)k * (n + 1) * 8,
                         hipMemcpyHostToDevice, s));
  if (vmiss) HIP_TRY(hipMemcpyAsync(D.at<char>(o_vmiss), vmiss, (size_t)k * n,
                                    hipMemcpyHostToDevice, s));
  if (od0) HIP_TRY(hipMemcpyAsync(D.at<char>(o_od0), od0, (size_t)n * 4, hipMemcpyHostToDevice, s));
  if (id0) HIP_TRY(hipMemcpyAsync(D.at<char>(o_id0), id0, (size_t)n * 4, hipMemcpyHostToDevice, s));
  SP.rows = D.at<uint64_t>(o_rows);
  SP.keep = keep ? D.at<uint64_t>(o_keep) : nullptr;
  SP.force = force ? D.at<uint64_t>(o_force) : nullptr;
  SP.vmat = D.at<double>(o_vmat);
  SP.vb = vb ? D.at<double>(o_vb) : nullptr;
  SP.wtab = D.at<double>(o_wtab);
  SP.vmiss = vmiss ? D.at<uint8_t>(o_vmiss) : nullptr;
  SP.od0 = od0 ? D.at<int32_t>(o_od0) : nullptr;
  SP.id0 = id0 ? D.at<int32_t>(o_id0) : nullptr;
  SP.out = D.at<double>(o_out);
  SP.scratch = stat_words ? D.at<uint64_t>(o_scr) : nullptr;
  const int le = saom::launch_net_stats(SP, s);
  if (le != 0) {
    set_error(std::string("stats kernel launch: ") + hipGetErrorString((hipError_t)le));
    return CUSNA_ERR_CUDA;
  }
  HIP_TRY(hipMemcpyAsync(out, D.at<char>(o_out), (size_t)M * k * 8, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  return CUSNA_OK;
}

// GoF distributions of M packed networks already in device memory (rows)
// or uploaded from host bit rows; masks optional (device).
cusna_status gof_device(const uint64_t* rows_d, int n, int M, const uint64_t* keep_d,
                        const uint64_t* force_d, int kind, const int32_t* levls, int nlev,
                        double* out, hipStream_t s) {
  if (kind < 0 || kind > 2) return CUSNA_ERR_UNSUPPORTED;
  const int W64 = (n + 63) / 64;
  int kmax = 0;
  if (kind == 2) {
    for (int L = 0; L < nlev; ++L) {
      if (levls[L] < 0) {
        set_error("cusna_gof_distribution: negative geodesic level");
        return CUSNA_ERR_SHAPE;
      }
      kmax = std::max(kmax, levls[L]);
    }
    kmax = std::min(kmax, n);  // no shortest path is longer than n - 1
  }
  // geodesic rows from device memory where the network exceeds the LDS
  // (CUSNA_GOF_STREAM=1 takes that form at every size)
  const char* force_stream = std::getenv("CUSNA_GOF_STREAM");
  const bool stream = kind == 2 && (saom::gof_geodesic_lds(n, W64, kmax) == 0 ||
                                    (force_stream && force_stream[0] == '1'));
  const int cols = nlev + (kind == 2 ? 1 : 0);
  DeviceBlock D;
  const size_t o_lv = D.add((size_t)nlev * 4), o_out = D.add((size_t)M * cols * 8);
  const size_t o_cnt = stream ? D.add((size_t)M * (kmax + 1) * 8) : 0;
  HIP_TRY(D.alloc());
  HIP_TRY(hipMemcpyAsync(D.at<char>(o_lv), levls, (size_t)nlev * 4, hipMemcpyHostToDevice, s));
  saom::GofParams G;
  G.n = n;
  G.M = M;
  G.W64 = W64;
  G.kind = kind;
  G.nlev = nlev;
  G.kmax = kmax;
  G.rows = rows_d;
  G.keep = keep_d;
  G.force = force_d;
  G.levls = D.at<int32_t>(o_lv);
  G.out = D.at<double>(o_out);
  G.counters = stream ? D.at<unsigned long long>(o_cnt) : nullptr;
  G.stream = stream ? 1 : 0;
  const int le = saom::launch_gof(G, s);
  if (le != 0) {
    set_error(std::string("gof kernel launch: ") + hipGetErrorString((hipError_t)le));
    return CUSNA_ERR_CUDA;
  }
  HIP_TRY(hipMemcpyAsync(out, D.at<char>(o_out), (size_t)M * cols * 8, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  return CUSNA_OK;
}

}  // namespace cusna_impl

extern "C" {

CUSNA_API cusna_status cusna_gof_distribution(const uint8_t* a_batch, int32_t n, int32_t M,
                                              int32_t kind, const int32_t* levls, int32_t nlev,
                                              double* out) {
  g_last_error.clear();
  // argument checks and codes as libcusna_saom_cpu.cpp:719
  if (!a_batch || !levls || !out || n < 1 || M < 1 || nlev < 1) return CUSNA_ERR_NULL_ARG;
  if (kind < 0 || kind > 2) return CUSNA_ERR_UNSUPPORTED;
  if (n > CUSNA_MAX_N) return CUSNA_ERR_SHAPE;
  if (device_count() < 1) return CUSNA_ERR_NO_CUDA;
  const int W64 = (n + 63) / 64;
  std::vector<uint64_t> rows((size_t)M * n * W64, 0ull);
  for (int c = 0; c < M; ++c) {
    const std::vector<uint64_t> r = dense_rows64(a_batch + (size_t)c * n * n, n, W64);
    std::copy(r.begin(), r.end(), rows.begin() + (size_t)c * n * W64);
  }
  hipStream_t s = thread_stream();
  DeviceBlock D;
  const size_t o_rows = D.add(rows.size() * 8);
  HIP_TRY(D.alloc());
  HIP_TRY(hipMemcpyAsync(D.at<char>(o_rows), rows.data(), rows.size() * 8, hipMemcpyHostToDevice,
                         s));
  return gof_device(D.at<uint64_t>(o_rows), n, M, nullptr, nullptr, kind, levls, nlev, out, s);
}

CUSNA_API cusna_status cusna_period_plan_gof(cusna_period_plan* plan, int32_t masked, int32_t kind,
                                             const int32_t* levls, int32_t nlev, double* out) {
  g_last_error.clear();
  if (!plan || !levls || !out) return CUSNA_ERR_NULL_ARG;
  if (nlev < 1) return CUSNA_ERR_NULL_ARG;
  HIP_TRY(hipStreamSynchronize(plan->s));
  return gof_device(plan->P.end_rows, plan->n, plan->M,
                    masked ? plan->D.at<uint64_t>(plan->o_keep) : nullptr,
                    masked ? plan->D.at<uint64_t>(plan->o_force) : nullptr, kind, levls, nlev,
                    out, plan->s);
}

CUSNA_API int32_t cusna_abi_version(void) { return CUSNA_ABI_VERSION; }

CUSNA_API const char* cusna_version_string(void) {
  return "cusna-amd 0.1 (gfx950 HIP kernels)";
}

CUSNA_API int32_t cusna_has_cuda(void) { return device_count() > 0 ? 1 : 0; }

CUSNA_API const char* cusna_last_error(void) { return g_last_error.c_str(); }

CUSNA_API cusna_status cusna_saom_simulate_period(const cusna_saom_problem* problem,
                                                  cusna_saom_result* result) {
  g_last_error.clear();
  return simulate(problem, result, nullptr, nullptr, 0);
}

CUSNA_API cusna_status cusna_saom_simulate_period_noalloc(const cusna_saom_problem* problem,
                                                          cusna_saom_result* result) {
  g_last_error.clear();
  return simulate_noalloc(problem, result);
}

CUSNA_API cusna_status cusna_saom_simulate_groups(const cusna_saom_problem* const* problems,
                                                  int32_t ngroups,
                                                  cusna_saom_result* const* results) {
  g_last_error.clear();
  if (!problems || !results) return CUSNA_ERR_NULL_ARG;
  if (ngroups < 1 || ngroups > 4096) return CUSNA_ERR_SHAPE;
  if (device_count() < 1) return CUSNA_ERR_NO_CUDA;
  return simulate_groups(problems, ngroups, results);
}

CUSNA_API cusna_status cusna_period_plan_specialize(cusna_period_plan* plan, int32_t mode) {
  if (!plan) return CUSNA_ERR_NULL_ARG;
  if (plan->replay && mode > 0) return CUSNA_ERR_UNSUPPORTED;
  plan->specialize = mode > 0 ? true
                     : mode == 0 ? false
                                 : (!plan->replay &&
                                    saom::specialize_policy((long long)plan->M * plan->n));
  return CUSNA_OK;
}

CUSNA_API int32_t cusna_period_plan_specialized(const cusna_period_plan* plan) {
  return plan && plan->last_rtc ? 1 : 0;
}

CUSNA_API cusna_status cusna_period_plan_autotune(cusna_period_plan* plan, int32_t probe_chains,
                                                  int32_t* best_threads) {
  if (!plan) return CUSNA_ERR_NULL_ARG;
  return autotune(plan, probe_chains, best_threads);
}

CUSNA_API cusna_status cusna_saom_simulate_period_replay(const cusna_saom_problem* problem,
                                                         cusna_saom_result* result,
                                                         const double* log_u0,
                                                         const double* loglog_u, int32_t smax) {
  g_last_error.clear();
  if (!log_u0 || !loglog_u) return CUSNA_ERR_NULL_ARG;
  if (smax < 1) return CUSNA_ERR_SHAPE;
  return simulate(problem, result, log_u0, loglog_u, smax);
}

CUSNA_API cusna_status cusna_pack_bits(const uint8_t* dense, int32_t n, uint32_t* out_rows,
                                       uint32_t* in_rows, int32_t* W_out) {
  if (!dense || !out_rows || !in_rows) return CUSNA_ERR_NULL_ARG;
  if (n < 1 || n > CUSNA_MAX_N) return CUSNA_ERR_SHAPE;
  const int W = (n + 31) / 32;
  std::memset(out_rows, 0, sizeof(uint32_t) * (size_t)n * W);
  std::memset(in_rows, 0, sizeof(uint32_t) * (size_t)n * W);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j)
      if (dense[(size_t)i * n + j]) {
        out_rows[(size_t)i * W + (j >> 5)] |= 1u << (j & 31);
        in_rows[(size_t)j * W + (i >> 5)] |= 1u << (i & 31);
      }
  if (W_out) *W_out = W;
  return CUSNA_OK;
}

CUSNA_API cusna_status cusna_philox4x32(const uint32_t counter[4], const uint32_t key[2],
                                        uint32_t out[4]) {
  if (!counter || !key || !out) return CUSNA_ERR_NULL_ARG;
  if (device_count() < 1) return CUSNA_ERR_NO_CUDA;
  DeviceBlock D;
  const size_t o_c = D.add(16), o_k = D.add(8), o_o = D.add(16);
  HIP_TRY(D.alloc());
  hipStream_t s = thread_stream();
  HIP_TRY(hipMemcpyAsync(D.at<char>(o_c), counter, 16, hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(D.at<char>(o_k), key, 8, hipMemcpyHostToDevice, s));
  if (saom::launch_philox(D.at<uint32_t>(o_c), D.at<uint32_t>(o_k), D.at<uint32_t>(o_o), s))
    return CUSNA_ERR_CUDA;
  HIP_TRY(hipMemcpyAsync(out, D.at<char>(o_o), 16, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  return CUSNA_OK;
}

CUSNA_API cusna_status cusna_gumbel_bound_check(uint32_t x0, uint32_t stride, uint32_t count,
                                               double* max_abs_err) {
  if (!max_abs_err) return CUSNA_ERR_NULL_ARG;
  if (device_count() < 1) return CUSNA_ERR_NO_CUDA;
  DeviceBlock D;
  const size_t o = D.add(8);
  HIP_TRY(D.alloc());
  hipStream_t s = thread_stream();
  HIP_TRY(hipMemsetAsync(D.at<char>(o), 0, 8, s));
  if (saom::launch_gumbel_check(x0, stride, count, D.at<unsigned long long>(o), s))
    return CUSNA_ERR_CUDA;
  HIP_TRY(hipMemcpyAsync(max_abs_err, D.at<char>(o), 8, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  return CUSNA_OK;
}

CUSNA_API cusna_status cusna_exp_weight_eval(const double* x, int32_t count, double* out) {
  if (!x || !out) return CUSNA_ERR_NULL_ARG;
  if (count < 1) return CUSNA_ERR_SHAPE;
  if (device_count() < 1) return CUSNA_ERR_NO_CUDA;
  DeviceBlock D;
  const size_t ox = D.add((size_t)count * 8), oo = D.add((size_t)count * 8);
  HIP_TRY(D.alloc());
  hipStream_t s = thread_stream();
  HIP_TRY(hipMemcpyAsync(D.at<char>(ox), x, (size_t)count * 8, hipMemcpyHostToDevice, s));
  if (saom::launch_exp_weight(D.at<double>(ox), count, D.at<double>(oo), s)) return CUSNA_ERR_CUDA;
  HIP_TRY(hipMemcpyAsync(out, D.at<char>(oo), (size_t)count * 8, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  return CUSNA_OK;
}

CUSNA_API cusna_status cusna_network_stats(const uint8_t* a, int32_t n, int32_t k,
                                           const int32_t* type_ids, const double* p1,
                                           const double* p2, const double* vmat,
                                           const uint8_t* vmiss, const int32_t* dyn_flags,
                                           const double* vb, const int32_t* od0,
                                           const int32_t* id0, const double* wtab,
                                           int32_t symflag, double* out) {
  g_last_error.clear();
  // argument checks of the reference (:625), including its status code
  if (!a || !type_ids || !out || n < 1 || n > CUSNA_MAX_N || k < 1) return CUSNA_ERR_NULL_ARG;
  const int W64 = (n + 63) / 64;
  const std::vector<uint64_t> rows = dense_rows64(a, n, W64);
  return net_stats_packed(rows.data(), n, 1, k, type_ids, p1, p2, vmat, vmiss, dyn_flags, vb, od0,
                          id0, wtab, symflag, nullptr, nullptr, out);
}

CUSNA_API cusna_status cusna_network_stats_packed(
    const uint32_t* rows, int32_t n, int32_t M, int32_t k, const int32_t* type_ids,
    const double* p1, const double* p2, const double* vmat, const uint8_t* vmiss,
    const int32_t* dyn_flags, const double* vb, const int32_t* od0, const int32_t* id0,
    const double* wtab, int32_t symflag, const uint32_t* keep, const uint32_t* force,
    double* out) {
  g_last_error.clear();
  if (!rows || !type_ids || !out) return CUSNA_ERR_NULL_ARG;
  if (n < 1 || n > CUSNA_MAX_N || k < 1 || M < 1) return CUSNA_ERR_SHAPE;
  const int W = (n + 31) / 32, W64 = (n + 63) / 64;
  std::vector<uint64_t> r64((size_t)M * n * W64);
  for (int c = 0; c < M; ++c) {
    const std::vector<uint64_t> one = rows64(rows + (size_t)c * n * W, n, W, W64, 0);
    std::copy(one.begin(), one.end(), r64.begin() + (size_t)c * n * W64);
  }
  std::vector<uint64_t> k64, f64;
  if (keep) k64 = rows64(keep, n, W, W64, 0);
  if (force) f64 = rows64(force, n, W, W64, 0);
  return net_stats_packed(r64.data(), n, M, k, type_ids, p1, p2, vmat, vmiss, dyn_flags, vb, od0,
                          id0, wtab, symflag, keep ? k64.data() : nullptr,
                          force ? f64.data() : nullptr, out);
}

CUSNA_API cusna_status cusna_behavior_stats(const int32_t* z_batch, int32_t n, int32_t M,
                                            int32_t kb, const int32_t* bty,
                                            const uint8_t* xnet, const int32_t* bkdist,
                                            double zbar, double brange, double bsimmean,
                                            double* out) {
  g_last_error.clear();
  // argument checks of the reference (:773-774), including its status code
  if (!z_batch || !bty || !xnet || !out || n < 1 || n > CUSNA_MAX_N || M < 1 || kb < 1)
    return CUSNA_ERR_NULL_ARG;
  if (kb > CUSNA_MAXBEFF) return CUSNA_ERR_SHAPE;
  if (device_count() < 1) {
    set_error("no HIP device");
    return CUSNA_ERR_NO_CUDA;
  }
  const int W64 = (n + 63) / 64;
  const std::vector<uint64_t> xr = dense_rows64(xnet, n, W64);
  saom::BehStatParams BP;
  std::memset(&BP, 0, sizeof(BP));
  BP.n = n;
  BP.M = M;
  BP.kb = kb;
  BP.W64 = W64;
  for (int e = 0; e < kb; ++e) BP.bty[e] = bty[e];
  BP.zbar = zbar;
  BP.brange = brange;
  BP.bsim = bsimmean;
  DeviceBlock D;
  const size_t o_z = D.add((size_t)M * n * 4), o_x = D.add(xr.size() * 8);
  const size_t o_bk = D.add((size_t)n * 4), o_out = D.add((size_t)M * kb * 8);
  const size_t o_scr = D.add((size_t)M * saom::beh_stats_scratch_doubles(n) * 8);
  HIP_TRY(D.alloc());
  hipStream_t s = thread_stream();
  HIP_TRY(hipMemcpyAsync(D.at<char>(o_z), z_batch, (size_t)M * n * 4, hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(D.at<char>(o_x), xr.data(), xr.size() * 8, hipMemcpyHostToDevice, s));
  if (bkdist)
    HIP_TRY(hipMemcpyAsync(D.at<char>(o_bk), bkdist, (size_t)n * 4, hipMemcpyHostToDevice, s));
  BP.z = D.at<int32_t>(o_z);
  BP.xnet = D.at<uint64_t>(o_x);
  BP.bkdist = bkdist ? D.at<int32_t>(o_bk) : nullptr;
  BP.out = D.at<double>(o_out);
  BP.scratch = saom::beh_stats_scratch_doubles(n) ? D.at<double>(o_scr) : nullptr;
  if (saom::launch_beh_stats(BP, s)) return CUSNA_ERR_CUDA;
  HIP_TRY(hipMemcpyAsync(out, D.at<char>(o_out), (size_t)M * kb * 8, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  return CUSNA_OK;
}

CUSNA_API cusna_status cusna_period_plan_create(const cusna_saom_problem* problem,
                                                int32_t want_end_nets, cusna_period_plan** plan) {
  g_last_error.clear();
  if (!problem || !plan) return CUSNA_ERR_NULL_ARG;
  cusna_status st = plan_create(problem, want_end_nets != 0, nullptr, nullptr, 0, nullptr, plan);
  if (st != CUSNA_OK) return st;
  st = plan_set_params(*plan, problem->theta, problem->lam, problem->thetab, problem->lamb);
  if (st != CUSNA_OK) {
    delete *plan;
    *plan = nullptr;
  }
  return st;
}

CUSNA_API cusna_status cusna_period_plan_set_params(cusna_period_plan* plan, const float* theta,
                                                    const float* lam, const float* thetab,
                                                    const float* lamb) {
  if (!plan || !theta) return CUSNA_ERR_NULL_ARG;
  if (plan->kb > 0 && (!thetab || !lamb)) return CUSNA_ERR_NULL_ARG;
  return plan_set_params(plan, theta, lam, thetab, lamb);
}

CUSNA_API cusna_status cusna_period_plan_run(cusna_period_plan* plan, uint64_t seed,
                                             int32_t want_stats) {
  if (!plan) return CUSNA_ERR_NULL_ARG;
  return plan_run(plan, seed, want_stats != 0, -1, nullptr, true);
}

CUSNA_API cusna_status cusna_period_plan_sync(cusna_period_plan* plan) {
  if (!plan) return CUSNA_ERR_NULL_ARG;
  HIP_TRY(hipStreamSynchronize(plan->s));
  return CUSNA_OK;
}

CUSNA_API cusna_status cusna_period_plan_fetch(cusna_period_plan* plan, cusna_saom_result* result) {
  if (!plan) return CUSNA_ERR_NULL_ARG;
  return plan_fetch(plan, result);
}

CUSNA_API double cusna_period_plan_kernel_ms(cusna_period_plan* plan) {
  if (!plan) return -1.0;
  float ms = -1.0f;
  if (hipEventSynchronize(plan->ev1) != hipSuccess) return -1.0;
  if (hipEventElapsedTime(&ms, plan->ev0, plan->ev1) != hipSuccess) return -1.0;
  return (double)ms;
}

CUSNA_API int32_t cusna_period_plan_shape(const cusna_period_plan* plan) {
  if (!plan) return 0;
  return plan->last_threads > 0 ? plan->last_threads : 64 * plan->nw;
}

CUSNA_API void cusna_period_plan_destroy(cusna_period_plan* plan) { delete plan; }

}  // extern "C"
