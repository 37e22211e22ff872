// Generated instantiation unit: FP form, 1 row(s) per lane.
#include "kernels_impl.h"

namespace ppi {
hipError_t launch_fp_r1(const EvalArgs& A, int F, int S, bool P, bool N, bool rows, hipStream_t st) {
    if (F == 32 && S == 16) return launch::launch_fp_t<1, 32, 16>(A, P, N, rows, st);
    if (F == 0 && S == 16) return launch::launch_fp_t<1, 0, 16>(A, P, N, rows, st);
    return hipErrorInvalidValue;
}
}  // namespace ppi
