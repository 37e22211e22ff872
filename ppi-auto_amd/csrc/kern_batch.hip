// Instantiation unit: batched persistent form.
#include "kernels_impl.h"

namespace ppi {
#define PPI_BATCH_SHAPES(X) X(2, 32, 9) X(4, 32, 9) X(8, 32, 9) X(4, 16, 9) X(8, 16, 9) X(2, 32, 16) X(4, 0, 9) \
    X(4, 32, 4) X(8, 32, 4) X(4, 16, 4) X(8, 16, 4) X(4, 0, 4)

bool batch_shape_supported(int R, int F, int S) {
#define PPI_X(r, f, s) if (R == r && F == f && S == s) return true;
    PPI_BATCH_SHAPES(PPI_X)
#undef PPI_X
    return false;
}

hipError_t launch_batch(const BatchArgs& B, int grid, int R, int F, int S, bool P, bool N, hipStream_t st) {
#define PPI_X(r, f, s) if (R == r && F == f && S == s) return launch::launch_batch_t<r, f, s>(B, grid, P, N, st);
    PPI_BATCH_SHAPES(PPI_X)
#undef PPI_X
    return hipErrorInvalidValue;
}

int batch_occupancy(int R, int F, int S, bool P, bool N) {
#define PPI_X(r, f, s) if (R == r && F == f && S == s) return launch::batch_occupancy_t<r, f, s>(P, N);
    PPI_BATCH_SHAPES(PPI_X)
#undef PPI_X
    return 0;
}
}  // namespace ppi
