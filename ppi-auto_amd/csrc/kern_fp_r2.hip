// Generated instantiation unit: FP form, 2 row(s) per lane.
#include "kernels_impl.h"

namespace ppi {
hipError_t launch_fp_r2(const EvalArgs& A, int F, int S, bool P, bool N, bool rows, hipStream_t st) {
    if (F == 8 && S == 9) return launch::launch_fp_t<2, 8, 9>(A, P, N, rows, st);
    if (F == 16 && S == 9) return launch::launch_fp_t<2, 16, 9>(A, P, N, rows, st);
    if (F == 32 && S == 9) return launch::launch_fp_t<2, 32, 9>(A, P, N, rows, st);
    if (F == 0 && S == 9) return launch::launch_fp_t<2, 0, 9>(A, P, N, rows, st);
    if (F == 32 && S == 16) return launch::launch_fp_t<2, 32, 16>(A, P, N, rows, st);
    return hipErrorInvalidValue;
}
}  // namespace ppi
