// Instantiation unit: reverse-mode adjoint kernels and their epilogues.
#define PPI_ADJOINT_MISC
#include "adjoint_impl.h"

namespace ppi {
// (rows per lane, features in VGPRs (0: read from the dataset), stack slots, adjoint slots,
// partial store (0: VGPR, 1: LDS)); two rows per lane: kern_adjoint_r2.hip.
#define PPI_ADJ_SHAPES(X) \
    X(1, 12, 12, 12, 0) X(1, 0, 12, 12, 0) X(1, 0, 16, 12, 0) X(1, 12, 12, 12, 1) X(1, 0, 12, 12, 1) X(1, 0, 16, 12, 1)

bool adjoint_shape_supported(int R, int F, int S, int AS, int PS) {
    if (R == 2) return adjoint_shape_supported_r2(F, S, AS, PS);
#define PPI_X(r, f, s, as, ps) if (R == r && F == f && S == s && AS == as && PS == ps) return true;
    PPI_ADJ_SHAPES(PPI_X)
#undef PPI_X
    return false;
}

int adjoint_store_partials(int PS, int R) {
    if (R == 2) return PS == 1 ? dev::StoreCfg<1, 2>::B : dev::StoreCfg<0, 2>::B;
    return PS == 1 ? dev::StoreCfg<1, 1>::B : dev::StoreCfg<0, 1>::B;
}

hipError_t launch_adjoint(const AdjArgs& A, int grid, int R, int F, int S, int AS, int PS, bool P, bool N,
                          hipStream_t st) {
    if (R == 2) return launch_adjoint_r2(A, grid, F, S, AS, PS, P, N, st);
#define PPI_X(r, f, s, as, ps)                                 \
    if (R == r && F == f && S == s && AS == as && PS == ps) \
        return launch::launch_adjoint_t<r, f, s, as, ps>(A, grid, P, N, st);
    PPI_ADJ_SHAPES(PPI_X)
#undef PPI_X
    return hipErrorInvalidValue;
}

int adjoint_occupancy(int R, int F, int S, int AS, int PS, bool P, bool N) {
    if (R == 2) return adjoint_occupancy_r2(F, S, AS, PS, P, N);
#define PPI_X(r, f, s, as, ps) \
    if (R == r && F == f && S == s && AS == as && PS == ps) return launch::adjoint_occupancy_t<r, f, s, as, ps>(P, N);
    PPI_ADJ_SHAPES(PPI_X)
#undef PPI_X
    return 0;
}

hipError_t launch_adjoint_finalize(const AdjFinal& F, hipStream_t st) {
    hipLaunchKernelGGL(dev::adj_fit_kernel, dim3((unsigned)((F.nprog + 255) / 256)), dim3(256), 0, st, F);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || F.nconst == 0) return e;
    hipLaunchKernelGGL(dev::adj_grad_kernel, dim3((unsigned)((F.nconst + 255) / 256)), dim3(256), 0, st, F);
    return hipGetLastError();
}

hipError_t launch_adjoint_descend(const AdjDescend& D, hipStream_t st) {
    hipLaunchKernelGGL(dev::adj_descend_kernel, dim3((unsigned)((D.nprog + 255) / 256)), dim3(256), 0, st, D);
    return hipGetLastError();
}
}  // namespace ppi
