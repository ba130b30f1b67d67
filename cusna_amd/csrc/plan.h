// plan.h -- internal: device-resident period plans shared by the C ABI
// (capi.hip) and the method-of-moments engine (mom.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstring>
#include <string>
#include <vector>

#include "../../include/libcusna.h"
#include "saom_kernels.h"

namespace cusna_impl {

// Device memory of one call: one hipMalloc, carved in 256-byte aligned
// pieces, freed on scope exit; or the same pieces placed in device memory the
// caller provides (place), which is then neither allocated nor freed.
class DeviceBlock {
 public:
  size_t add(size_t bytes) {
    const size_t off = total_;
    total_ += (bytes + 255) & ~(size_t)255;
    return off;
  }
  hipError_t alloc() {
    if (total_ == 0) total_ = 256;
    return hipMalloc(&base_, total_);
  }
  // false when the pieces need more than `capacity` bytes
  bool place(void* base, size_t capacity) {
    if (total_ == 0) total_ = 256;
    if (total_ > capacity) return false;
    base_ = base;
    owned_ = false;
    return true;
  }
  size_t bytes() const { return total_; }
  template <class T>
  T* at(size_t off) const {
    return reinterpret_cast<T*>(static_cast<char*>(base_) + off);
  }
  ~DeviceBlock() {
    if (base_ && owned_) (void)hipFree(base_);
  }

 private:
  void* base_ = nullptr;
  size_t total_ = 0;
  bool owned_ = true;
};

// Device memory a call places its buffers in instead of allocating them:
// fixed memory of `capacity` bytes at `base`, or `grow(bytes)` returning at
// least `bytes` (cached per thread, null when unavailable). `staging(bytes)`
// (optional) returns pinned host memory of at least `bytes` through which
// the call moves its inputs and its outputs with one copy each.
struct DeviceArena {
  void* base = nullptr;
  size_t capacity = 0;
  void* (*grow)(size_t bytes) = nullptr;
  void* (*staging)(size_t bytes) = nullptr;
};

void set_error(const std::string& s);
int device_count();
hipStream_t thread_stream();

// ABI bit rows (n*W u32) -> kernel rows (n*W64 u64); NULL -> off-diagonal
// (dflt 1) or zero (dflt 0).
std::vector<uint64_t> rows64(const uint32_t* rows, int n, int W, int W64, int dflt);
std::vector<uint64_t> dense_rows64(const uint8_t* a, int n, int W64);
// Smallest power of two >= W64 (the compiled word bound).
int wmax_for(int W64);
// Feature flags (saom::F_*) an effect list needs.
int effect_flags(const int32_t* t, int k, int kb, const int32_t* dyn);

}  // namespace cusna_impl

#define HIP_TRY(expr)                                                                    \
  do {                                                                                   \
    hipError_t e_ = (expr);                                                              \
    if (e_ != hipSuccess) {                                                              \
      cusna_impl::set_error(std::string(#expr) + ": " + hipGetErrorString(e_));                      \
      return CUSNA_ERR_CUDA;                                                             \
    }                                                                                    \
  } while (0)

// A period plan owns the device copy of one problem: chain-invariant inputs
// are uploaded once, per-chain parameters can be replaced between runs, and
// outputs stay in HBM until fetched. The one-shot ABI call is create, run,
// fetch, destroy.
struct cusna_period_plan {
  int n = 0, k = 0, kb = 0, M = 0, W64 = 0, wmax = 0, nw = 0, nd = 0;
  bool replay = false, want_dense = false, own_stream = false;
  int symflag = 0;
  bool has_vmiss = false, has_od0 = false, has_id0 = false;
  saom::SimParams P{};
  cusna_impl::DeviceBlock D;
  hipStream_t s = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  size_t o_keep = 0, o_force = 0, o_vb = 0, o_vmiss = 0, o_od0 = 0, o_id0 = 0, o_stats = 0;
  size_t o_dense = 0, o_th = 0, o_lam = 0, o_thb = 0, o_lamb = 0, o_params = 0;
  size_t o_sscr = 0;       // statistics scratch (networks above the LDS)
  size_t o_rsc = 0, o_adist = 0;  // rate-effect scores, per-actor distances
  bool has_sscr = false;
  bool specialize = false;  // launch the runtime-specialised period kernel
  bool free_shape = false;  // block shape left to the library (no threads, none tuned)
  bool last_rtc = false;    // the last run did
  int last_threads = 0;     // threads per chain of the last run's launch (0: none yet)
  // device layout: inputs in [0, in_bytes), the outputs the host reads in
  // [out_begin, out_end); staging: pinned host memory of a staged call
  size_t in_bytes = 0, out_begin = 0, out_end = 0;
  char* staging = nullptr;
  bool params_set = false;  // the per-chain parameters came with the inputs
  bool stats_valid = false;
  ~cusna_period_plan() {
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (own_stream && s) (void)hipStreamDestroy(s);
  }
};


namespace cusna_impl {

// arena: place the plan's device buffers there (no allocation, no timing
// events, no runtime-specialised kernel, whose module load allocates) and
// run on `stream` as given, the null stream included; CUSNA_ERR_SHAPE when
// they do not fit.
cusna_status plan_create(const cusna_saom_problem* pb, bool want_dense, const double* log_u0,
                         const double* loglog_u, int32_t smax, hipStream_t stream,
                         cusna_period_plan** out, const DeviceArena* arena = nullptr);
cusna_status plan_set_params(cusna_period_plan* pl, const float* theta, const float* lam,
                             const float* thetab, const float* lamb);
// Launch one period (and the statistics kernel) for the first M_run chains
// (M_run <= 0: all) on `stream` (nullptr: the plan's stream); `timed` records
// the HIP events around the period kernel.
cusna_status plan_run(cusna_period_plan* pl, uint64_t seed, bool want_stats, int M_run,
                      hipStream_t stream, bool timed);
// The statistics kernel over the first M_run chains' end networks.
cusna_status plan_stats(cusna_period_plan* pl, int M_run, hipStream_t stream);
cusna_status plan_fetch(cusna_period_plan* pl, cusna_saom_result* r);

}  // namespace cusna_impl
