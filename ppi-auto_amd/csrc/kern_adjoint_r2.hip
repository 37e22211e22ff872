// Instantiation unit: reverse-mode adjoint kernels at two rows per lane.
#include "adjoint_impl.h"

namespace ppi {
// (features in VGPRs (0: read from the dataset), stack slots, adjoint slots, partial store (0: VGPR, 1: LDS)).
// Five slots in one register tuple of 10 per stack: programs of up to five stack and adjoint slots.
// Features are read from the dataset: holding 12 per row in VGPRs (24 registers) costs a wave per
// SIMD (151 VGPRs, 3 waves) and measured 15% slower on the 1M-row adjoint workload (565 vs 494 ms),
// although the tuner's short sample ranked it 4% faster.
#define PPI_ADJ_SHAPES_R2(X) X(0, 5, 5, 0) X(0, 5, 5, 1)

bool adjoint_shape_supported_r2(int F, int S, int AS, int PS) {
#define PPI_X(f, s, as, ps) if (F == f && S == s && AS == as && PS == ps) return true;
    PPI_ADJ_SHAPES_R2(PPI_X)
#undef PPI_X
    return false;
}

hipError_t launch_adjoint_r2(const AdjArgs& A, int grid, int F, int S, int AS, int PS, bool P, bool N, hipStream_t st) {
#define PPI_X(f, s, as, ps)                         \
    if (F == f && S == s && AS == as && PS == ps) \
        return launch::launch_adjoint_t<2, f, s, as, ps>(A, grid, P, N, st);
    PPI_ADJ_SHAPES_R2(PPI_X)
#undef PPI_X
    return hipErrorInvalidValue;
}

int adjoint_occupancy_r2(int F, int S, int AS, int PS, bool P, bool N) {
#define PPI_X(f, s, as, ps) \
    if (F == f && S == s && AS == as && PS == ps) return launch::adjoint_occupancy_t<2, f, s, as, ps>(P, N);
    PPI_ADJ_SHAPES_R2(PPI_X)
#undef PPI_X
    return 0;
}
}  // namespace ppi
