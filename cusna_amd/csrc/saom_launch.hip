// saom_launch.hip -- dispatch of the period kernel over its compiled shapes.
#include <hip/hip_runtime.h>

#include "saom_kernels.h"

namespace saom {

int launch_period_w1(const SimParams& P, int nw, bool replay, void* stream);
int launch_period_w2(const SimParams& P, int nw, bool replay, void* stream);
int launch_period_w4(const SimParams& P, int nw, bool replay, void* stream);
int launch_period_w8(const SimParams& P, int nw, bool replay, void* stream);
int launch_period_w16(const SimParams& P, int nw, bool replay, void* stream);
size_t sim_lds_bytes_w1(const SimParams& P, int nw);

size_t sim_lds_bytes(const SimParams& P, int nw) { return sim_lds_bytes_w1(P, nw); }

int launch_period(const SimParams& P, LaunchShape s, bool replay, void* stream) {
  switch (s.wmax) {
    case 1: return launch_period_w1(P, s.nw, replay, stream);
    case 2: return launch_period_w2(P, s.nw, replay, stream);
    case 4: return launch_period_w4(P, s.nw, replay, stream);
    case 8: return launch_period_w8(P, s.nw, replay, stream);
    case 16: return launch_period_w16(P, s.nw, replay, stream);
    default: return (int)hipErrorInvalidValue;
  }
}

}  // namespace saom
