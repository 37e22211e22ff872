// Generated instantiation unit: FP form, 4 row(s) per lane.
#include "kernels_impl.h"

namespace ppi {
hipError_t launch_fp_r4(const EvalArgs& A, int F, int S, bool P, bool N, bool rows, hipStream_t st) {
    if (F == 8 && S == 9) return launch::launch_fp_t<4, 8, 9>(A, P, N, rows, st);
    if (F == 16 && S == 9) return launch::launch_fp_t<4, 16, 9>(A, P, N, rows, st);
    if (F == 32 && S == 9) return launch::launch_fp_t<4, 32, 9>(A, P, N, rows, st);
    if (F == 0 && S == 9) return launch::launch_fp_t<4, 0, 9>(A, P, N, rows, st);
    if (F == 8 && S == 4) return launch::launch_fp_t<4, 8, 4>(A, P, N, rows, st);
    if (F == 16 && S == 4) return launch::launch_fp_t<4, 16, 4>(A, P, N, rows, st);
    if (F == 32 && S == 4) return launch::launch_fp_t<4, 32, 4>(A, P, N, rows, st);
    if (F == 0 && S == 4) return launch::launch_fp_t<4, 0, 4>(A, P, N, rows, st);
    return hipErrorInvalidValue;
}
}  // namespace ppi
