// Half-tile instantiations of the depth-sweep kernel for 1024- and
// 2048-lane rows (half the waves per workgroup of the full tile; for rows of
// other FFT lengths, 1024-lane only), in a translation unit of their own so
// that they compile beside kernels.hip.
#include <hip/hip_runtime.h>

#include "kernels.h"
#include "sweep_kernel.hpp"

namespace wemfwi {

hipError_t launch_oneway_half(int N, int k, const SweepRow* rows, int nrows, const DepthTables* models,
                              StepGeom g, int nz, int step, bool adj, const float2* tw, hipStream_t s) {
  if (nrows <= 0 || nz < 2) return hipSuccess;
  if (g.fg.n != N) {
    if (N == 1024) return launch_oneway_gen<1024, true>(rows, nrows, models, g, nz, step, adj, tw, s);
    return hipErrorInvalidValue;
  }
  switch (N) {
    case 1024: return launch_oneway_plan<1024, true>(k, rows, nrows, models, g, nz, step, adj, tw, s);
    case 2048: return launch_oneway_plan<2048, true>(k, rows, nrows, models, g, nz, step, adj, tw, s);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace wemfwi
