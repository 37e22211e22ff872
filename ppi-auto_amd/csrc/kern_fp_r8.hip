// Generated instantiation unit: FP form, 8 row(s) per lane.
#include "kernels_impl.h"

namespace ppi {
hipError_t launch_fp_r8(const EvalArgs& A, int F, int S, bool P, bool N, bool rows, hipStream_t st) {
    if (F == 8 && S == 9) return launch::launch_fp_t<8, 8, 9>(A, P, N, rows, st);
    if (F == 16 && S == 9) return launch::launch_fp_t<8, 16, 9>(A, P, N, rows, st);
    if (F == 8 && S == 4) return launch::launch_fp_t<8, 8, 4>(A, P, N, rows, st);
    if (F == 16 && S == 4) return launch::launch_fp_t<8, 16, 4>(A, P, N, rows, st);
    return hipErrorInvalidValue;
}
}  // namespace ppi
