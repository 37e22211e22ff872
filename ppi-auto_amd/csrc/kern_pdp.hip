// Generated instantiation unit: PDP form.
#include "kernels_impl.h"

namespace ppi {
hipError_t launch_pdp(const EvalArgs& A, int R, int F, int S, bool P, bool N, hipStream_t st) {
    if (R == 1 && F == 32 && S == 16) return launch::launch_pdp_t<1, 32, 16>(A, P, N, st);
    if (R == 1 && F == 0 && S == 16) return launch::launch_pdp_t<1, 0, 16>(A, P, N, st);
    if (R == 2 && F == 8 && S == 9) return launch::launch_pdp_t<2, 8, 9>(A, P, N, st);
    if (R == 2 && F == 16 && S == 9) return launch::launch_pdp_t<2, 16, 9>(A, P, N, st);
    if (R == 2 && F == 32 && S == 9) return launch::launch_pdp_t<2, 32, 9>(A, P, N, st);
    if (R == 2 && F == 0 && S == 9) return launch::launch_pdp_t<2, 0, 9>(A, P, N, st);
    if (R == 2 && F == 32 && S == 16) return launch::launch_pdp_t<2, 32, 16>(A, P, N, st);
    return hipErrorInvalidValue;
}

bool pdp_shape_supported(int R, int F, int S) {
    return (R == 1 && F == 32 && S == 16) ||
           (R == 1 && F == 0 && S == 16) ||
           (R == 2 && F == 8 && S == 9) ||
           (R == 2 && F == 16 && S == 9) ||
           (R == 2 && F == 32 && S == 9) ||
           (R == 2 && F == 0 && S == 9) ||
           (R == 2 && F == 32 && S == 16);
}
}  // namespace ppi
