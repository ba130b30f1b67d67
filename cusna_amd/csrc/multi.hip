// multi.hip -- C ABI of the multi-network period simulator and the
// cross-network statistics (host side).
//
// cusna_saom_simulate_multi replaces the reference's multi-network entry
// point (cran/cusna src/libcusna.h:284-285) with the semantics of its CPU
// mirror cusna_saom_simulate_multi_cpu (libcusna_saom_cpu.cpp:548-604):
// argument checks and status codes, float32 descriptors promoted to float64,
// per-chain theta/lam rows, perm defaulting to the off-diagonal, act to the
// identity. A device-resident plan (cusna_multi_plan_*) keeps one problem in
// HBM across runs; the one-shot call is create, run, fetch, destroy.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "../../include/libcusna.h"
#include "plan.h"
#include "saom_kernels.h"

using namespace cusna_impl;

// Statistics of one network's end states (set by cusna_multi_plan_set_stats):
// its non-cross effects through the network statistics kernel, its cross
// effects through the cross-statistics kernel, all on the masked network.
struct MultiStats {
  std::vector<int> cols_nc, cols_cross;  // columns (global effect index)
  saom::StatParams SP{};
  std::vector<saom::CrossParams> CP;
  DeviceBlock B;
  size_t o_out_nc = 0, o_out_cross = 0;
};

struct cusna_multi_plan {
  int n = 0, k = 0, D = 0, M = 0, W64 = 0, wmax = 0, nw = 0, nd = 0;
  std::vector<std::unique_ptr<MultiStats>> stats;
  bool want_dense = false;
  bool last_rtc = false;  // the last run launched the specialised kernel
  hipStream_t s = nullptr;
  DeviceBlock Dv;
  saom::MultiParams P;  // host copy (device pointers inside)
  saom::MultiParams* Pd = nullptr;
  size_t o_th = 0, o_lam = 0, o_end = 0, o_sc = 0, o_rsc = 0, o_ntim = 0, o_steps = 0,
         o_dist = 0, o_dense = 0, o_x0 = 0, o_pm = 0, o_kd = 0;
};

namespace cusna_impl {
namespace {

cusna_status validate_multi(const cusna_saom_multi_problem* pb) {
  // libcusna_saom_cpu.cpp:550-554
  if (!pb) return CUSNA_ERR_NULL_ARG;
  if (pb->struct_version != CUSNA_ABI_VERSION) return CUSNA_ERR_SHAPE;
  const int n = pb->n, W = pb->W, k = pb->k, M = pb->M, D = pb->D;
  if (n < 1 || n > CUSNA_MAX_N || W != (n + 31) / 32 || k < 1 || k > CUSNA_MAXEFF || M < 1 ||
      D < 2 || D > CUSNA_MAXNET)
    return CUSNA_ERR_SHAPE;
  if (!pb->x0_out || !pb->lam || !pb->theta || !pb->type_ids || !pb->etypes || !pb->netof ||
      !pb->symf)
    return CUSNA_ERR_NULL_ARG;
  for (int e = 0; e < k; ++e) {
    const int t = pb->type_ids[e];
    if (pb->netof[e] < 0 || pb->netof[e] >= D) return CUSNA_ERR_SHAPE;
    if (t >= 33 && t <= 37 && !pb->netref) return CUSNA_ERR_NULL_ARG;
    if (t >= 33 && t <= 37 && (pb->netref[e] < 0 || pb->netref[e] >= D)) return CUSNA_ERR_SHAPE;
    if (pb->dyc_idx && pb->dyc_idx[e] >= 0 && !pb->dyc) return CUSNA_ERR_NULL_ARG;
  }
  if (pb->act && (pb->nact < 1 || pb->nact > n)) return CUSNA_ERR_SHAPE;
  if (pb->act)
    for (int q = 0; q < pb->nact; ++q)
      if (pb->act[q] < 0 || pb->act[q] >= n) return CUSNA_ERR_SHAPE;
  return CUSNA_OK;
}

// Block shape: `threads` / 64 waves per chain rounded down to 1, 2 or 4 (at
// most the word bound); 0 = the default below. Streamed word bounds run 16
// waves.
int multi_nw(int wmax, int threads) {
  if (wmax > 16) return 16;
  if (threads > 0) {
    int nw = 1;
    while (nw * 2 <= threads / 64 && nw * 2 <= std::min(4, wmax)) nw *= 2;
    return nw;
  }
  // default: two words per lane, at most four waves (D networks' rows leave
  // room for few chains per CU, so the waves of one chain hide the latency;
  // two networks of 500 actors: 368 / 262 / 218 ms per 4096-chain period at
  // one / two / four waves per chain; eight waves of one word: 216 against
  // 178 ms with the float32-bound choice)
  return std::max(1, std::min(4, wmax / 2));
}

}  // namespace

cusna_status multi_plan_create(const cusna_saom_multi_problem* pb, bool want_dense,
                               const uint32_t* kdist, hipStream_t stream,
                               cusna_multi_plan** out) {
  cusna_status st = validate_multi(pb);
  if (st != CUSNA_OK) return st;
  if (device_count() < 1) {
    set_error("no HIP device");
    return CUSNA_ERR_NO_CUDA;
  }
  const int n = pb->n, W = pb->W, k = pb->k, M = pb->M, D = pb->D;
  const int W64 = (n + 63) / 64;
  const int wmax = wmax_for(W64);
  std::unique_ptr<cusna_multi_plan> pl(new cusna_multi_plan());
  pl->n = n;
  pl->k = k;
  pl->D = D;
  pl->M = M;
  pl->W64 = W64;
  pl->wmax = wmax;
  pl->nw = multi_nw(wmax, pb->threads);
  pl->want_dense = want_dense;
  pl->s = stream;
  saom::MultiParams& Q = pl->P;
  std::memset(&Q, 0, sizeof(Q));
  saom::SimParams& P = Q.base;
  P.n = n;
  P.k = k;
  P.M = M;
  P.W64 = W64;
  P.RS = n;
  P.nact = pb->act ? pb->nact : n;
  P.key0 = (uint32_t)(pb->seed & 0xFFFFFFFFu);
  P.key1 = (uint32_t)(pb->seed >> 32);
  P.flags = effect_flags(pb->type_ids, k, 0, nullptr) & ~(saom::F_RDEG | saom::F_Z);
  int nd = 0;
  for (int e = 0; e < k; ++e) {
    P.tid[e] = pb->type_ids[e];
    P.ety[e] = pb->etypes[e];
    P.dyn[e] = 0;
    P.dci[e] = pb->dyc_idx ? pb->dyc_idx[e] : -1;
    P.c0[e] = P.c1[e] = P.c2[e] = -1;
    P.p1[e] = pb->p1 ? (double)pb->p1[e] : 0.0;
    P.p2[e] = pb->p2 ? (double)pb->p2[e] : 0.0;
    if (P.dci[e] + 1 > nd) nd = P.dci[e] + 1;
    Q.netof[e] = pb->netof[e];
    Q.netref[e] = pb->netref ? pb->netref[e] : 0;
  }
  for (int d = 0; d < D; ++d) Q.symf[d] = pb->symf[d];
  Q.D = D;
  Q.M = M;
  pl->nd = pb->dyc ? nd : 0;
  // working rows in LDS when the D networks fit beside the rest of the state
  Q.rows_global = 0;
  if (saom::multi_lds_bytes(Q, pl->nw, wmax) > 160 * 1024) Q.rows_global = 1;
  if (saom::multi_lds_bytes(Q, pl->nw, wmax) > 160 * 1024) {
    set_error("multi-network chain state exceeds the 160 KB LDS of a CU");
    return CUSNA_ERR_UNSUPPORTED;
  }
  const bool rg = Q.rows_global || wmax > 16;
  const size_t WS = wmax > 16 ? (size_t)W64 : (size_t)wmax;

  // host staging
  const size_t rowsb = (size_t)n * W64 * 8;
  std::vector<uint64_t> x0((size_t)D * n * W64), pm((size_t)D * n * W64), kd;
  for (int d = 0; d < D; ++d) {
    const std::vector<uint64_t> a = rows64(pb->x0_out + (size_t)d * n * W, n, W, W64, 0);
    const std::vector<uint64_t> p =
        rows64(pb->perm ? pb->perm + (size_t)d * n * W : nullptr, n, W, W64, 1);
    std::copy(a.begin(), a.end(), x0.begin() + (size_t)d * n * W64);
    std::copy(p.begin(), p.end(), pm.begin() + (size_t)d * n * W64);
  }
  if (kdist) {
    kd.resize((size_t)D * n * W64);
    for (int d = 0; d < D; ++d) {
      const std::vector<uint64_t> a = rows64(kdist + (size_t)d * n * W, n, W, W64, 0);
      std::copy(a.begin(), a.end(), kd.begin() + (size_t)d * n * W64);
    }
  }
  std::vector<double> vmat((size_t)k * n, 0.0), wtab((size_t)k * (n + 1), 0.0), dyc;
  if (pb->vmat)
    for (size_t q = 0; q < vmat.size(); ++q) vmat[q] = pb->vmat[q];
  if (pb->wtab)
    for (size_t q = 0; q < wtab.size(); ++q) wtab[q] = pb->wtab[q];
  if (pl->nd > 0) {
    dyc.resize((size_t)pl->nd * n * n);
    for (size_t q = 0; q < dyc.size(); ++q) dyc[q] = pb->dyc[q];
  }
  std::vector<int32_t> act;
  if (pb->act) act.assign(pb->act, pb->act + n);

  DeviceBlock& B = pl->Dv;
  const size_t o_P = B.add(sizeof(saom::MultiParams));
  pl->o_x0 = B.add((size_t)D * rowsb);
  pl->o_pm = B.add((size_t)D * rowsb);
  pl->o_kd = B.add(kd.empty() ? 0 : (size_t)D * rowsb);
  const size_t o_vmat = B.add(vmat.size() * 8), o_wtab = B.add(wtab.size() * 8);
  const size_t o_dyc = B.add(dyc.size() * 8), o_act = B.add(act.size() * 4);
  pl->o_th = B.add((size_t)M * k * 8);
  pl->o_lam = B.add((size_t)M * D * 8);
  const size_t o_work = B.add(rg ? (size_t)M * D * WS * n * 8 : 0);
  pl->o_end = B.add((size_t)D * M * rowsb);
  pl->o_sc = B.add((size_t)M * k * 8);
  pl->o_rsc = B.add((size_t)M * D * 8);
  pl->o_ntim = B.add((size_t)M * 8);
  pl->o_steps = B.add((size_t)M * 4);
  pl->o_dist = B.add((size_t)D * M * 4);
  pl->o_dense = B.add(want_dense ? (size_t)M * n * n : 0);
  HIP_TRY(B.alloc());
  auto up = [&](size_t off, const void* src, size_t bytes) -> hipError_t {
    if (!bytes) return hipSuccess;
    return hipMemcpyAsync(B.at<char>(off), src, bytes, hipMemcpyHostToDevice, stream);
  };
  HIP_TRY(up(pl->o_x0, x0.data(), x0.size() * 8));
  HIP_TRY(up(pl->o_pm, pm.data(), pm.size() * 8));
  HIP_TRY(up(pl->o_kd, kd.data(), kd.size() * 8));
  HIP_TRY(up(o_vmat, vmat.data(), vmat.size() * 8));
  HIP_TRY(up(o_wtab, wtab.data(), wtab.size() * 8));
  HIP_TRY(up(o_dyc, dyc.data(), dyc.size() * 8));
  HIP_TRY(up(o_act, act.data(), act.size() * 4));
  P.vmat = B.at<double>(o_vmat);
  P.wtab = B.at<double>(o_wtab);
  P.dyc = dyc.empty() ? nullptr : B.at<double>(o_dyc);
  P.act = act.empty() ? nullptr : B.at<int32_t>(o_act);
  Q.x0 = B.at<uint64_t>(pl->o_x0);
  Q.pm = B.at<uint64_t>(pl->o_pm);
  Q.kd = kd.empty() ? nullptr : B.at<uint64_t>(pl->o_kd);
  Q.theta = B.at<double>(pl->o_th);
  Q.lam = B.at<double>(pl->o_lam);
  Q.work = rg ? B.at<uint64_t>(o_work) : nullptr;
  Q.end_rows = B.at<uint64_t>(pl->o_end);
  Q.scores = B.at<double>(pl->o_sc);
  Q.rscores = B.at<double>(pl->o_rsc);
  Q.ntim = B.at<double>(pl->o_ntim);
  Q.steps = B.at<int32_t>(pl->o_steps);
  Q.dist_stat = Q.kd ? B.at<int32_t>(pl->o_dist) : nullptr;
  pl->Pd = B.at<saom::MultiParams>(o_P);
  HIP_TRY(up(o_P, &Q, sizeof(Q)));
  // the host staging vectors die with this scope
  HIP_TRY(hipStreamSynchronize(stream));
  *out = pl.release();
  return CUSNA_OK;
}

cusna_status multi_plan_set_params(cusna_multi_plan* pl, const float* theta, const float* lam) {
  const int M = pl->M, k = pl->k, D = pl->D;
  std::vector<double> th((size_t)M * k), lm((size_t)M * D);
  for (size_t q = 0; q < th.size(); ++q) th[q] = theta[q];
  for (size_t q = 0; q < lm.size(); ++q) lm[q] = lam[q];
  HIP_TRY(hipMemcpyAsync(pl->Dv.at<char>(pl->o_th), th.data(), th.size() * 8,
                         hipMemcpyHostToDevice, pl->s));
  HIP_TRY(hipMemcpyAsync(pl->Dv.at<char>(pl->o_lam), lm.data(), lm.size() * 8,
                         hipMemcpyHostToDevice, pl->s));
  // the staging vectors die here: finish the copies first
  HIP_TRY(hipStreamSynchronize(pl->s));
  return CUSNA_OK;
}

cusna_status multi_plan_run(cusna_multi_plan* pl, uint64_t seed) {
  saom::MultiParams& Q = pl->P;
  Q.base.key0 = (uint32_t)(seed & 0xFFFFFFFFu);
  Q.base.key1 = (uint32_t)(seed >> 32);
  HIP_TRY(hipMemcpyAsync(pl->Pd, &Q, sizeof(Q), hipMemcpyHostToDevice, pl->s));
  // the kernel specialised to the effect list when the launch is large
  // (CUSNA_SPECIALIZE=0/1 overrides); the generic kernel otherwise or when
  // the specialised one is unavailable
  pl->last_rtc = false;
  if (saom::specialize_policy((long long)pl->M * pl->n)) {
    void* fn = saom::rtc_multi_kernel(Q, pl->nw, pl->wmax);
    if (fn && saom::launch_multi_rtc(fn, pl->Pd, pl->M, saom::multi_lds_bytes(Q, pl->nw, pl->wmax),
                                     pl->nw, pl->s) == 0) {
      pl->last_rtc = true;
    } else {
      (void)hipGetLastError();
    }
  }
  const int e = pl->last_rtc ? 0 : saom::launch_multi(Q, pl->Pd, pl->nw, pl->wmax, pl->s);
  if (e != 0) {
    set_error(std::string("multi-network period launch: ") + hipGetErrorString((hipError_t)e));
    return CUSNA_ERR_CUDA;
  }
  // the parameter copy above reads the plan's host struct: keep it alive
  HIP_TRY(hipStreamSynchronize(pl->s));
  return CUSNA_OK;
}

cusna_status multi_plan_fetch(cusna_multi_plan* pl, cusna_saom_multi_result* r) {
  if (!r) return CUSNA_ERR_NULL_ARG;
  if (r->struct_version != CUSNA_ABI_VERSION) return CUSNA_ERR_SHAPE;
  const int M = pl->M, k = pl->k, D = pl->D, n = pl->n;
  hipStream_t s = pl->s;
  std::vector<double> sc, rsc;
  if (r->scores) sc.resize((size_t)M * k);
  if (r->rscores) rsc.resize((size_t)M * D);
  auto down = [&](void* dst, size_t off, size_t bytes) -> hipError_t {
    return hipMemcpyAsync(dst, pl->Dv.at<char>(off), bytes, hipMemcpyDeviceToHost, s);
  };
  if (r->scores) HIP_TRY(down(sc.data(), pl->o_sc, sc.size() * 8));
  if (r->rscores) HIP_TRY(down(rsc.data(), pl->o_rsc, rsc.size() * 8));
  if (r->ntim) HIP_TRY(down(r->ntim, pl->o_ntim, (size_t)M * 8));
  if (r->steps) HIP_TRY(down(r->steps, pl->o_steps, (size_t)M * 4));
  if (r->a_ends) {
    if (!pl->want_dense) {
      set_error("plan created without dense end networks");
      return CUSNA_ERR_SHAPE;
    }
    // network d of chain c lands at a_ends[(c*D + d)*n*n]
    const size_t nn = (size_t)n * n;
    for (int d = 0; d < D; ++d) {
      const int le = saom::launch_expand_dense(
          pl->P.end_rows + (size_t)d * M * n * pl->W64, n, pl->W64, M,
          pl->Dv.at<uint8_t>(pl->o_dense), s);
      if (le != 0) {
        set_error("expand kernel launch failed");
        return CUSNA_ERR_CUDA;
      }
      HIP_TRY(hipMemcpy2DAsync(r->a_ends + (size_t)d * nn, (size_t)D * nn,
                               pl->Dv.at<uint8_t>(pl->o_dense), nn, nn, M,
                               hipMemcpyDeviceToHost, s));
    }
  }
  HIP_TRY(hipStreamSynchronize(s));
  if (r->scores)
    for (size_t q = 0; q < sc.size(); ++q) r->scores[q] = (float)sc[q];
  if (r->rscores)
    for (size_t q = 0; q < rsc.size(); ++q) r->rscores[q] = (float)rsc[q];
  return CUSNA_OK;
}

}  // namespace cusna_impl

extern "C" {

CUSNA_API cusna_status cusna_saom_simulate_multi(const cusna_saom_multi_problem* pb,
                                                 cusna_saom_multi_result* r) {
  if (!pb || !r) return CUSNA_ERR_NULL_ARG;
  if (r->struct_version != CUSNA_ABI_VERSION) return CUSNA_ERR_SHAPE;
  cusna_multi_plan* raw = nullptr;
  cusna_status st = multi_plan_create(pb, r->a_ends != nullptr, nullptr, thread_stream(), &raw);
  if (st != CUSNA_OK) return st;
  std::unique_ptr<cusna_multi_plan> pl(raw);
  st = multi_plan_set_params(pl.get(), pb->theta, pb->lam);
  if (st == CUSNA_OK) st = multi_plan_run(pl.get(), pb->seed);
  if (st == CUSNA_OK) st = multi_plan_fetch(pl.get(), r);
  return st;
}

CUSNA_API cusna_status cusna_multi_plan_create(const cusna_saom_multi_problem* pb,
                                               int32_t want_end_nets, const uint32_t* kdist,
                                               cusna_multi_plan** plan) {
  if (!plan) return CUSNA_ERR_NULL_ARG;
  *plan = nullptr;
  hipStream_t s = nullptr;
  if (device_count() >= 1 && hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess)
    return CUSNA_ERR_CUDA;
  cusna_status st = multi_plan_create(pb, want_end_nets != 0, kdist, s, plan);
  if (st != CUSNA_OK) {
    if (s) (void)hipStreamDestroy(s);
    return st;
  }
  st = multi_plan_set_params(*plan, pb->theta, pb->lam);
  if (st != CUSNA_OK) {
    delete *plan;
    *plan = nullptr;
    (void)hipStreamDestroy(s);
  }
  return st;
}

CUSNA_API cusna_status cusna_multi_plan_set_params(cusna_multi_plan* plan, const float* theta,
                                                   const float* lam) {
  if (!plan || !theta || !lam) return CUSNA_ERR_NULL_ARG;
  return multi_plan_set_params(plan, theta, lam);
}

CUSNA_API cusna_status cusna_multi_plan_run(cusna_multi_plan* plan, uint64_t seed) {
  if (!plan) return CUSNA_ERR_NULL_ARG;
  return multi_plan_run(plan, seed);
}

CUSNA_API cusna_status cusna_multi_plan_fetch(cusna_multi_plan* plan,
                                              cusna_saom_multi_result* r) {
  if (!plan) return CUSNA_ERR_NULL_ARG;
  return multi_plan_fetch(plan, r);
}

CUSNA_API cusna_status cusna_multi_plan_dist(cusna_multi_plan* plan, int32_t* dist_stat) {
  if (!plan || !dist_stat) return CUSNA_ERR_NULL_ARG;
  if (!plan->P.kd) {
    set_error("plan created without kdist");
    return CUSNA_ERR_SHAPE;
  }
  HIP_TRY(hipMemcpyAsync(dist_stat, plan->Dv.at<char>(plan->o_dist),
                         (size_t)plan->D * plan->M * 4, hipMemcpyDeviceToHost, plan->s));
  HIP_TRY(hipStreamSynchronize(plan->s));
  return CUSNA_OK;
}

CUSNA_API cusna_status cusna_cross_stats(const uint8_t* a_batch, const uint8_t* wp, int32_t n,
                                         int32_t M, int32_t t, double parm, double* out) {
  // argument checks and codes as libcusna_saom_cpu.cpp:822-823, 852
  if (!a_batch || !wp || !out || n < 1 || n > CUSNA_MAX_N || M < 1) return CUSNA_ERR_NULL_ARG;
  if (t < 33 || t > 37) return CUSNA_ERR_UNSUPPORTED;
  if (device_count() < 1) return CUSNA_ERR_NO_CUDA;
  const int W64 = (n + 63) / 64;
  const size_t nw = (size_t)n * W64;
  std::vector<uint64_t> rows((size_t)M * nw);
  for (int c = 0; c < M; ++c) {
    const std::vector<uint64_t> r = dense_rows64(a_batch + (size_t)c * n * n, n, W64);
    std::copy(r.begin(), r.end(), rows.begin() + (size_t)c * nw);
  }
  std::vector<uint8_t> wpt((size_t)n * n);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) wpt[(size_t)j * n + i] = wp[(size_t)i * n + j];
  const std::vector<uint64_t> w = dense_rows64(wp, n, W64), wt = dense_rows64(wpt.data(), n, W64);
  hipStream_t s = thread_stream();
  DeviceBlock B;
  const size_t o_rows = B.add(rows.size() * 8), o_w = B.add(nw * 8), o_wt = B.add(nw * 8),
               o_out = B.add((size_t)M * 8);
  HIP_TRY(B.alloc());
  HIP_TRY(hipMemcpyAsync(B.at<char>(o_rows), rows.data(), rows.size() * 8,
                         hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(B.at<char>(o_w), w.data(), nw * 8, hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(B.at<char>(o_wt), wt.data(), nw * 8, hipMemcpyHostToDevice, s));
  saom::CrossParams C{};
  C.n = n;
  C.M = M;
  C.W64 = W64;
  C.t = t;
  C.parm = parm;
  C.rows = B.at<uint64_t>(o_rows);
  C.wp = B.at<uint64_t>(o_w);
  C.wpT = B.at<uint64_t>(o_wt);
  C.out = B.at<double>(o_out);
  C.out_stride = 1;
  const int le = saom::launch_cross_stats(C, s);
  if (le != 0) {
    set_error("cross statistics launch failed");
    return CUSNA_ERR_CUDA;
  }
  HIP_TRY(hipMemcpyAsync(out, B.at<char>(o_out), (size_t)M * 8, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  return CUSNA_OK;
}

CUSNA_API cusna_status cusna_multi_plan_set_stats(cusna_multi_plan* plan, int32_t d,
                                                  const cusna_multi_stat_spec* sp) {
  if (!plan || !sp) return CUSNA_ERR_NULL_ARG;
  if (sp->struct_version != CUSNA_ABI_VERSION || d < 0 || d >= plan->D) return CUSNA_ERR_SHAPE;
  const int n = plan->n, W = (n + 31) / 32, W64 = plan->W64, M = plan->M;
  const saom::MultiParams& Q = plan->P;
  std::vector<int> cols;
  for (int e = 0; e < plan->k; ++e)
    if (Q.netof[e] == d) cols.push_back(e);
  if (sp->k != (int)cols.size()) {
    set_error("statistic spec: k differs from the network's effect count");
    return CUSNA_ERR_SHAPE;
  }
  if (sp->k > 0 && (!sp->type_ids || !sp->p1 || !sp->p2 || !sp->vmat)) return CUSNA_ERR_NULL_ARG;
  std::unique_ptr<MultiStats> st(new MultiStats());
  std::vector<int> nc, cr;  // positions within the spec
  for (int q = 0; q < sp->k; ++q) {
    const int t = sp->type_ids[q];
    if (t != Q.base.tid[cols[q]]) {
      set_error("statistic spec: effect ids differ from the plan's");
      return CUSNA_ERR_SHAPE;
    }
    if (t >= 33 && t <= 37) {
      if (!sp->cross_wp) return CUSNA_ERR_NULL_ARG;
      cr.push_back(q);
      st->cols_cross.push_back(cols[q]);
    } else {
      nc.push_back(q);
      st->cols_nc.push_back(cols[q]);
    }
  }
  const int knc = (int)nc.size(), kcr = (int)cr.size();
  const size_t rowsb = (size_t)n * W64 * 8;
  std::vector<int32_t> tnc(knc);
  for (int q = 0; q < knc; ++q) tnc[q] = sp->type_ids[nc[q]];
  const size_t words = knc ? saom::net_stats_scratch_words(n, W64, saom::stats_need_cols(
                                                                       tnc.data(), knc))
                           : 0;
  // host staging of the non-cross descriptors (sub-list) and the masks
  std::vector<double> vmat((size_t)knc * n), wtab((size_t)knc * (n + 1), 0.0), vb(n, 0.0);
  std::vector<uint8_t> vmiss((size_t)knc * n, 0);
  for (int q = 0; q < knc; ++q) {
    std::copy(sp->vmat + (size_t)nc[q] * n, sp->vmat + (size_t)(nc[q] + 1) * n,
              vmat.begin() + (size_t)q * n);
    if (sp->wtab)
      std::copy(sp->wtab + (size_t)nc[q] * (n + 1), sp->wtab + (size_t)(nc[q] + 1) * (n + 1),
                wtab.begin() + (size_t)q * (n + 1));
    if (sp->vmiss)
      std::copy(sp->vmiss + (size_t)nc[q] * n, sp->vmiss + (size_t)(nc[q] + 1) * n,
                vmiss.begin() + (size_t)q * n);
  }
  const std::vector<uint64_t> keep = rows64(sp->keep, n, W, W64, 1);
  const std::vector<uint64_t> force = rows64(sp->force, n, W, W64, 0);
  std::vector<uint64_t> wp((size_t)kcr * n * W64), wpt((size_t)kcr * n * W64);
  for (int q = 0; q < kcr; ++q) {
    const uint32_t* src = sp->cross_wp + (size_t)cr[q] * n * W;
    const std::vector<uint64_t> r = rows64(src, n, W, W64, 0);
    std::copy(r.begin(), r.end(), wp.begin() + (size_t)q * n * W64);
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j)
        if ((r[(size_t)i * W64 + (j >> 6)] >> (j & 63)) & 1)
          wpt[(size_t)q * n * W64 + (size_t)j * W64 + (i >> 6)] |= 1ull << (i & 63);
  }
  DeviceBlock& B = st->B;
  const size_t o_keep = B.add(rowsb), o_force = B.add(rowsb);
  const size_t o_vmat = B.add(vmat.size() * 8), o_wtab = B.add(wtab.size() * 8),
               o_vb = B.add(vb.size() * 8), o_vmiss = B.add(vmiss.size());
  const size_t o_od0 = B.add(sp->od0 ? (size_t)n * 4 : 0), o_id0 = B.add(sp->id0 ? (size_t)n * 4 : 0);
  const size_t o_wp = B.add(wp.size() * 8), o_wpt = B.add(wpt.size() * 8);
  st->o_out_nc = B.add((size_t)M * knc * 8);
  st->o_out_cross = B.add((size_t)M * kcr * 8);
  const size_t o_scr = B.add((size_t)M * words * 8);
  HIP_TRY(B.alloc());
  hipStream_t s = plan->s;
  auto up = [&](size_t off, const void* src, size_t bytes) -> hipError_t {
    if (!bytes) return hipSuccess;
    return hipMemcpyAsync(B.at<char>(off), src, bytes, hipMemcpyHostToDevice, s);
  };
  HIP_TRY(up(o_keep, keep.data(), rowsb));
  HIP_TRY(up(o_force, force.data(), rowsb));
  HIP_TRY(up(o_vmat, vmat.data(), vmat.size() * 8));
  HIP_TRY(up(o_wtab, wtab.data(), wtab.size() * 8));
  HIP_TRY(up(o_vb, vb.data(), vb.size() * 8));
  HIP_TRY(up(o_vmiss, vmiss.data(), vmiss.size()));
  if (sp->od0) HIP_TRY(up(o_od0, sp->od0, (size_t)n * 4));
  if (sp->id0) HIP_TRY(up(o_id0, sp->id0, (size_t)n * 4));
  HIP_TRY(up(o_wp, wp.data(), wp.size() * 8));
  HIP_TRY(up(o_wpt, wpt.data(), wpt.size() * 8));
  HIP_TRY(hipStreamSynchronize(s));
  const uint64_t* rows = Q.end_rows + (size_t)d * M * n * W64;
  saom::StatParams& SP = st->SP;
  std::memset(&SP, 0, sizeof(SP));
  SP.n = n;
  SP.k = knc;
  SP.M = M;
  SP.W64 = W64;
  SP.symflag = sp->symflag;
  for (int q = 0; q < knc; ++q) {
    SP.tid[q] = tnc[q];
    SP.dyn[q] = 0;
    SP.p1[q] = sp->p1[nc[q]];
    SP.p2[q] = sp->p2[nc[q]];
  }
  SP.rows = rows;
  SP.keep = B.at<uint64_t>(o_keep);
  SP.force = B.at<uint64_t>(o_force);
  SP.vmat = B.at<double>(o_vmat);
  SP.vb = B.at<double>(o_vb);
  SP.wtab = B.at<double>(o_wtab);
  SP.vmiss = sp->vmiss ? B.at<uint8_t>(o_vmiss) : nullptr;
  SP.od0 = sp->od0 ? B.at<int32_t>(o_od0) : nullptr;
  SP.id0 = sp->id0 ? B.at<int32_t>(o_id0) : nullptr;
  SP.out = B.at<double>(st->o_out_nc);
  SP.scratch = words ? B.at<uint64_t>(o_scr) : nullptr;
  for (int q = 0; q < kcr; ++q) {
    saom::CrossParams C{};
    C.n = n;
    C.M = M;
    C.W64 = W64;
    C.t = sp->type_ids[cr[q]];
    C.parm = sp->p1[cr[q]];
    C.rows = rows;
    C.keep = SP.keep;
    C.force = SP.force;
    C.wp = B.at<uint64_t>(o_wp) + (size_t)q * n * W64;
    C.wpT = B.at<uint64_t>(o_wpt) + (size_t)q * n * W64;
    C.out = B.at<double>(st->o_out_cross) + q;
    C.out_stride = kcr;
    st->CP.push_back(C);
  }
  if (plan->stats.size() != (size_t)plan->D) plan->stats.resize(plan->D);
  plan->stats[d] = std::move(st);
  return CUSNA_OK;
}

CUSNA_API cusna_status cusna_multi_plan_stats(cusna_multi_plan* plan, double* out) {
  if (!plan || !out) return CUSNA_ERR_NULL_ARG;
  const int M = plan->M, k = plan->k;
  if (plan->stats.size() != (size_t)plan->D) {
    set_error("statistics not set for every network");
    return CUSNA_ERR_SHAPE;
  }
  for (const auto& st : plan->stats)
    if (!st) {
      set_error("statistics not set for every network");
      return CUSNA_ERR_SHAPE;
    }
  hipStream_t s = plan->s;
  for (const auto& st : plan->stats) {
    if (st->SP.k > 0) {
      const int le = saom::launch_net_stats(st->SP, s);
      if (le != 0) {
        set_error("network statistics launch failed");
        return CUSNA_ERR_CUDA;
      }
    }
    for (const auto& C : st->CP) {
      const int le = saom::launch_cross_stats(C, s);
      if (le != 0) {
        set_error("cross statistics launch failed");
        return CUSNA_ERR_CUDA;
      }
    }
  }
  std::vector<std::vector<double>> host(plan->D * 2);
  for (int d = 0; d < plan->D; ++d) {
    const MultiStats& st = *plan->stats[d];
    host[2 * d].resize((size_t)M * st.cols_nc.size());
    host[2 * d + 1].resize((size_t)M * st.cols_cross.size());
    if (!host[2 * d].empty())
      HIP_TRY(hipMemcpyAsync(host[2 * d].data(), st.B.at<char>(st.o_out_nc),
                             host[2 * d].size() * 8, hipMemcpyDeviceToHost, s));
    if (!host[2 * d + 1].empty())
      HIP_TRY(hipMemcpyAsync(host[2 * d + 1].data(), st.B.at<char>(st.o_out_cross),
                             host[2 * d + 1].size() * 8, hipMemcpyDeviceToHost, s));
  }
  HIP_TRY(hipStreamSynchronize(s));
  for (int d = 0; d < plan->D; ++d) {
    const MultiStats& st = *plan->stats[d];
    const size_t a = st.cols_nc.size(), b = st.cols_cross.size();
    for (int c = 0; c < M; ++c) {
      for (size_t q = 0; q < a; ++q) out[(size_t)c * k + st.cols_nc[q]] = host[2 * d][c * a + q];
      for (size_t q = 0; q < b; ++q)
        out[(size_t)c * k + st.cols_cross[q]] = host[2 * d + 1][c * b + q];
    }
  }
  return CUSNA_OK;
}

CUSNA_API int32_t cusna_multi_plan_specialized(const cusna_multi_plan* plan) {
  return plan && plan->last_rtc ? 1 : 0;
}

CUSNA_API cusna_status cusna_multi_plan_destroy(cusna_multi_plan* plan) {
  if (!plan) return CUSNA_OK;
  hipStream_t s = plan->s;
  if (s) (void)hipStreamSynchronize(s);
  delete plan;
  if (s) (void)hipStreamDestroy(s);
  return CUSNA_OK;
}

}  // extern "C"
