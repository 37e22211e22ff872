// Generated instantiation unit: dispatch, PP form, finalize, predict.
#define PPI_KERNELS_MISC 1
#include "kernels_impl.h"

namespace ppi {
hipError_t launch_fp_r1(const EvalArgs& A, int F, int S, bool P, bool N, bool rows, hipStream_t st);
hipError_t launch_fp_r2(const EvalArgs& A, int F, int S, bool P, bool N, bool rows, hipStream_t st);
hipError_t launch_fp_r4(const EvalArgs& A, int F, int S, bool P, bool N, bool rows, hipStream_t st);
hipError_t launch_fp_r8(const EvalArgs& A, int F, int S, bool P, bool N, bool rows, hipStream_t st);

bool shape_supported(int R, int F, int S) {
    return (R == 1 && F == 32 && S == 16) ||
           (R == 1 && F == 0 && S == 16) ||
           (R == 2 && F == 8 && S == 9) ||
           (R == 2 && F == 16 && S == 9) ||
           (R == 2 && F == 32 && S == 9) ||
           (R == 2 && F == 0 && S == 9) ||
           (R == 2 && F == 32 && S == 16) ||
           (R == 4 && F == 8 && S == 9) ||
           (R == 4 && F == 16 && S == 9) ||
           (R == 4 && F == 32 && S == 9) ||
           (R == 4 && F == 0 && S == 9) ||
           (R == 8 && F == 8 && S == 9) ||
           (R == 8 && F == 16 && S == 9) ||
           (R == 4 && F == 8 && S == 4) ||
           (R == 4 && F == 16 && S == 4) ||
           (R == 4 && F == 32 && S == 4) ||
           (R == 4 && F == 0 && S == 4) ||
           (R == 8 && F == 8 && S == 4) ||
           (R == 8 && F == 16 && S == 4);
}

hipError_t launch_fp(const EvalArgs& A, int R, int F, int S, bool P, bool N, hipStream_t st, bool rows) {
    switch (R) {
    case 1: return launch_fp_r1(A, F, S, P, N, rows, st);
    case 2: return launch_fp_r2(A, F, S, P, N, rows, st);
    case 4: return launch_fp_r4(A, F, S, P, N, rows, st);
    case 8: return launch_fp_r8(A, F, S, P, N, rows, st);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_ordered_sum(const EvalArgs& A, int rows_per_block, bool last, hipStream_t st) {
    // One wave per workgroup: 30,000 programs are 469 waves, spread over every CU (256-thread
    // workgroups put them on 118 CUs).
    const dim3 grid((unsigned)((A.nprog + 63) / 64)), block(64);
    if (A.reducemax)
        hipLaunchKernelGGL(dev::ordered_sum_kernel<true>, grid, block, 0, st, A, rows_per_block, last ? 1 : 0);
    else
        hipLaunchKernelGGL(dev::ordered_sum_kernel<false>, grid, block, 0, st, A, rows_per_block, last ? 1 : 0);
    return hipGetLastError();
}

hipError_t launch_finalize(const EvalArgs& A, int nrb, hipStream_t st) {
    // One wave per workgroup: 30,000 programs are 469 waves, over every CU.
    hipLaunchKernelGGL(dev::finalize_kernel, dim3((unsigned)((A.nprog + 63) / 64)), dim3(64), 0, st, A, nrb);
    return hipGetLastError();
}

hipError_t launch_recheck(const EvalArgs& A, bool P, bool N, hipStream_t st) {
    if (A.nprog <= 0) return hipSuccess;
    if (A.pdp_lanes < 1 || A.pdp_lanes > kPdpMaxLanes || A.dp_local < 1 || A.dp_local > kPdpMaxLanes)
        return hipErrorInvalidValue;
    dim3 grid((unsigned)((A.nprog + dev::kRecheckSpan - 1) / dev::kRecheckSpan));
    if (P && N) hipLaunchKernelGGL((dev::recheck_kernel<true, true>), grid, dim3(dev::kBlock), 0, st, A);
    else if (P) hipLaunchKernelGGL((dev::recheck_kernel<true, false>), grid, dim3(dev::kBlock), 0, st, A);
    else if (N) hipLaunchKernelGGL((dev::recheck_kernel<false, true>), grid, dim3(dev::kBlock), 0, st, A);
    else hipLaunchKernelGGL((dev::recheck_kernel<false, false>), grid, dim3(dev::kBlock), 0, st, A);
    return hipGetLastError();
}

// Every pair's pdp_lanes and dp_local come from the configured work-group limit (<= kPdpMaxLanes).
hipError_t launch_batch_recheck(const BatchArgs& B, const int32_t* pair_of, int32_t n_total, bool P, bool N,
                                hipStream_t st) {
    if (n_total <= 0) return hipSuccess;
    dim3 grid((unsigned)((n_total + dev::kRecheckSpan - 1) / dev::kRecheckSpan));
    if (P && N) hipLaunchKernelGGL((dev::batch_recheck_kernel<true, true>), grid, dim3(dev::kBlock), 0, st, B, pair_of, n_total);
    else if (P) hipLaunchKernelGGL((dev::batch_recheck_kernel<true, false>), grid, dim3(dev::kBlock), 0, st, B, pair_of, n_total);
    else if (N) hipLaunchKernelGGL((dev::batch_recheck_kernel<false, true>), grid, dim3(dev::kBlock), 0, st, B, pair_of, n_total);
    else hipLaunchKernelGGL((dev::batch_recheck_kernel<false, false>), grid, dim3(dev::kBlock), 0, st, B, pair_of, n_total);
    return hipGetLastError();
}

hipError_t launch_deep(const DeepArgs& D, int grid, bool P, bool N, hipStream_t st) {
    if (D.base.deep_cap <= 0) return hipSuccess;
    if (D.base.pdp_lanes > kPdpMaxLanes || D.base.dp_local > kPdpMaxLanes || D.max_rows > D.stride || grid < 1)
        return hipErrorInvalidValue;
    const dim3 g((unsigned)grid), b(dev::kBlock);
    if (P && N) hipLaunchKernelGGL((dev::deep_rows_kernel<true, true>), g, b, 0, st, D);
    else if (P) hipLaunchKernelGGL((dev::deep_rows_kernel<true, false>), g, b, 0, st, D);
    else if (N) hipLaunchKernelGGL((dev::deep_rows_kernel<false, true>), g, b, 0, st, D);
    else hipLaunchKernelGGL((dev::deep_rows_kernel<false, false>), g, b, 0, st, D);
    hipLaunchKernelGGL(dev::deep_partition_kernel, dim3((unsigned)D.base.deep_cap), b, 0, st, D);
    return hipGetLastError();
}

hipError_t launch_pp(const EvalArgs& A, bool P, bool N, hipStream_t st) {
    const dim3 grid((unsigned)((A.nprog + dev::kPPBlock - 1) / dev::kPPBlock)), block(dev::kPPBlock);
    if (P && N) hipLaunchKernelGGL((dev::pp_kernel<true, true>), grid, block, 0, st, A);
    else if (P) hipLaunchKernelGGL((dev::pp_kernel<true, false>), grid, block, 0, st, A);
    else if (N) hipLaunchKernelGGL((dev::pp_kernel<false, true>), grid, block, 0, st, A);
    else hipLaunchKernelGGL((dev::pp_kernel<false, false>), grid, block, 0, st, A);
    return hipGetLastError();
}

hipError_t launch_predict(const EvalArgs& A, int F, int S, bool P, bool N, float* out, hipStream_t st) {
    if (F == 32 && S == 16) return launch::launch_predict_t<32, 16>(A, P, N, out, st);
    if (F == 0 && S == 16) return launch::launch_predict_t<0, 16>(A, P, N, out, st);
    return hipErrorInvalidValue;
}

hipError_t launch_fold(const ProgMeta* meta, const Ins* code, int32_t n, float* out, bool P, bool N, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    dim3 grid((unsigned)((n + dev::kWaves - 1) / dev::kWaves));
    if (P && N) hipLaunchKernelGGL((dev::fold_kernel<true, true>), grid, dim3(dev::kBlock), 0, st, meta, code, n, out);
    else if (P) hipLaunchKernelGGL((dev::fold_kernel<true, false>), grid, dim3(dev::kBlock), 0, st, meta, code, n, out);
    else if (N) hipLaunchKernelGGL((dev::fold_kernel<false, true>), grid, dim3(dev::kBlock), 0, st, meta, code, n, out);
    else hipLaunchKernelGGL((dev::fold_kernel<false, false>), grid, dim3(dev::kBlock), 0, st, meta, code, n, out);
    return hipGetLastError();
}

hipError_t launch_derive(float* X, int64_t ldx, const uint32_t* ops, const int32_t* srcs, int32_t first,
                         int32_t count, bool P, bool N, hipStream_t st) {
    if (count <= 0) return hipSuccess;
    dim3 grid((unsigned)((ldx + 255) / 256));
    if (P && N) hipLaunchKernelGGL((dev::derive_kernel<true, true>), grid, dim3(256), 0, st, X, ldx, ops, srcs, first, count);
    else if (P) hipLaunchKernelGGL((dev::derive_kernel<true, false>), grid, dim3(256), 0, st, X, ldx, ops, srcs, first, count);
    else if (N) hipLaunchKernelGGL((dev::derive_kernel<false, true>), grid, dim3(256), 0, st, X, ldx, ops, srcs, first, count);
    else hipLaunchKernelGGL((dev::derive_kernel<false, false>), grid, dim3(256), 0, st, X, ldx, ops, srcs, first, count);
    return hipGetLastError();
}

hipError_t launch_fill_u32(uint32_t* p, uint32_t v, int64_t n, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(dev::fill_u32_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, v, n);
    return hipGetLastError();
}

hipError_t launch_ef_prepare(const EvalArgs& A, hipStream_t st) {
    if (A.nprog <= 0) return hipSuccess;
    hipLaunchKernelGGL(dev::ef_prepare_kernel, dim3((unsigned)((A.nprog + 255) / 256)), dim3(256), 0, st, A);
    return hipGetLastError();
}
}  // namespace ppi

namespace ppi {
hipError_t launch_batch_finalize(const BatchArgs& B, const int32_t* pair_of, int32_t n_total, hipStream_t st) {
    hipLaunchKernelGGL(dev::batch_finalize_kernel, dim3((unsigned)((n_total + 63) / 64)), dim3(64), 0, st, B,
                       pair_of, n_total);
    return hipGetLastError();
}
}  // namespace ppi
