// engine.cc — runtime of the PPI evaluation engine and its C ABI (include/ppi_engine.h).
//
// Replaces the OpenCL host of the reference (src/interpreter/accelerator.cc):
// device selection, dataset staging (once), per-generation program upload,
// kernel launch, result download and ranking.  One engine = one HIP device,
// one stream, one dataset, one resident program set.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstring>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

#include "../../include/ppi_engine.h"
#include "engine_internal.h"
#include "ge.h"
#include "kernels.h"
#include "tape.h"

namespace ppi {

thread_local std::string g_last_error;

void set_error(const std::string& s) { g_last_error = s; }

#define PPI_HIP(call)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (call);                                                                \
        if (e_ != hipSuccess) {                                                                \
            set_error(std::string(#call) + ": " + hipGetErrorString(e_));                      \
            return PPI_E_DEVICE;                                                               \
        }                                                                                      \
    } while (0)

int32_t derived_capacity(int64_t ldx, int32_t ncol) {
    if (ldx <= 0) return 0;
    const int64_t by_mem = ((int64_t)1 << 31) / (ldx * (int64_t)sizeof(float)) - ncol;
    return (int32_t)std::max<int64_t>(0, std::min<int64_t>(4096, by_mem));
}

ppi_status stage_derived_columns(float** X, size_t* cap_floats, int64_t ldx, int32_t ncol, DerivedColumns& table,
                                 int32_t* ready, uint32_t** dOps, int32_t** dSrc, size_t* cap_list, bool P, bool N,
                                 hipStream_t st) {
    if (table.staged_epoch != table.epoch) {  // the table was refilled: every column is new
        *ready = 0;
        table.staged_epoch = table.epoch;
    }
    const int32_t total = (int32_t)table.cols.size();
    const int32_t fresh = total - *ready;
    if (fresh <= 0) return PPI_OK;
    const size_t need = (size_t)(ncol + total) * (size_t)ldx;
    if (need > *cap_floats) {
        const int32_t room = std::min<int32_t>(table.cap, total + 32);
        const size_t want = (size_t)(ncol + room) * (size_t)ldx;
        float* nx = nullptr;
        if (hipMalloc((void**)&nx, want * sizeof(float)) != hipSuccess) {
            set_error("device allocation of derived columns failed");
            return PPI_E_NOMEM;
        }
        const size_t keep = (size_t)(ncol + *ready) * (size_t)ldx;
        hipError_t e = hipMemcpyAsync(nx, *X, keep * sizeof(float), hipMemcpyDeviceToDevice, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) {
            (void)hipFree(nx);
            set_error(std::string("derived columns: ") + hipGetErrorString(e));
            return PPI_E_DEVICE;
        }
        (void)hipFree(*X);
        *X = nx;
        *cap_floats = want;
    }
    std::vector<uint32_t> ops((size_t)total);
    std::vector<int32_t> srcs((size_t)total);
    for (int32_t j = 0; j < total; ++j) {
        ops[(size_t)j] = table.cols[(size_t)j].op;
        srcs[(size_t)j] = table.cols[(size_t)j].src;
    }
    if ((size_t)total > *cap_list) {
        if (*dOps) (void)hipFree(*dOps);
        if (*dSrc) (void)hipFree(*dSrc);
        *dOps = nullptr;
        *dSrc = nullptr;
        *cap_list = 0;
        const size_t want = (size_t)table.cap;
        if (hipMalloc((void**)dOps, want * sizeof(uint32_t)) != hipSuccess ||
            hipMalloc((void**)dSrc, want * sizeof(int32_t)) != hipSuccess) {
            set_error("device allocation of the derived column list failed");
            return PPI_E_NOMEM;
        }
        *cap_list = want;
    }
    PPI_HIP(hipMemcpyAsync(*dOps, ops.data(), ops.size() * sizeof(uint32_t), hipMemcpyHostToDevice, st));
    PPI_HIP(hipMemcpyAsync(*dSrc, srcs.data(), srcs.size() * sizeof(int32_t), hipMemcpyHostToDevice, st));
    PPI_HIP(launch_derive(*X, ldx, *dOps + *ready, *dSrc + *ready, ncol + *ready, fresh, P, N, st));
    PPI_HIP(hipStreamSynchronize(st));  // the host vectors above go out of scope
    *ready = total;
    return PPI_OK;
}

template <typename T>
static ppi_status ensure(T** p, size_t* cap, size_t n) {
    if (n <= *cap && *p) return PPI_OK;
    if (*p) (void)hipFree(*p);
    *p = nullptr;
    *cap = 0;
    size_t want = std::max<size_t>(n, 1);
    hipError_t e = hipMalloc((void**)p, want * sizeof(T));
    if (e != hipSuccess) {
        set_error(std::string("hipMalloc(") + std::to_string(want * sizeof(T)) + " B): " + hipGetErrorString(e));
        *p = nullptr;
        return PPI_E_NOMEM;
    }
    *cap = want;
    return PPI_OK;
}

template <typename T>
static ppi_status ensure_host(T** p, size_t* cap, size_t n) {
    if (n <= *cap && *p) return PPI_OK;
    if (*p) (void)hipHostFree(*p);
    *p = nullptr;
    *cap = 0;
    size_t want = std::max<size_t>(n, 1);
    hipError_t e = hipHostMalloc((void**)p, want * sizeof(T), hipHostMallocDefault);
    if (e != hipSuccess) {
        set_error(std::string("hipHostMalloc: ") + hipGetErrorString(e));
        *p = nullptr;
        return PPI_E_NOMEM;
    }
    *cap = want;
    return PPI_OK;
}

static double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

void FoldBuffers::release() {
    if (meta) (void)hipFree(meta);
    if (code) (void)hipFree(code);
    if (out) (void)hipFree(out);
    meta = nullptr;
    code = nullptr;
    out = nullptr;
    cap_meta = cap_code = cap_out = 0;
}

ppi_status evaluate_folds(Tape* t, FoldBuffers* B, bool P, bool N, hipStream_t st) {
    const FoldSet& F = t->folds;
    const size_t n = F.meta.size();
    if (n == 0) return PPI_OK;
    ppi_status s;
    if ((s = ensure(&B->meta, &B->cap_meta, n))) return s;
    if ((s = ensure(&B->code, &B->cap_code, F.code.size()))) return s;
    if ((s = ensure(&B->out, &B->cap_out, n))) return s;
    B->host.resize(n);
    PPI_HIP(hipMemcpyAsync(B->meta, F.meta.data(), n * sizeof(ProgMeta), hipMemcpyHostToDevice, st));
    PPI_HIP(hipMemcpyAsync(B->code, F.code.data(), F.code.size() * sizeof(Ins), hipMemcpyHostToDevice, st));
    PPI_HIP(launch_fold(B->meta, B->code, (int32_t)n, B->out, P, N, st));
    PPI_HIP(hipMemcpyAsync(B->host.data(), B->out, n * sizeof(float), hipMemcpyDeviceToHost, st));
    PPI_HIP(hipStreamSynchronize(st));
    apply_folds(t, B->host.data());
    return PPI_OK;
}

// util::PickNBest (Util.h:158-173): k-th best = k-th smallest error, ties to the
// larger index; `indices` maps positions to program indices when given.
void pick_n_best(int n, int32_t* best, int num, const float* errors, const int32_t* indices) {
    std::vector<int32_t> ord((size_t)num);
    std::iota(ord.begin(), ord.end(), 0);
    auto better = [&](int32_t a, int32_t b) {
        if (errors[a] < errors[b]) return true;
        if (errors[b] < errors[a]) return false;
        return a > b;
    };
    n = std::min(n, num);
    std::partial_sort(ord.begin(), ord.begin() + n, ord.end(), better);
    for (int k = 0; k < n; ++k) best[k] = indices ? indices[ord[k]] : ord[k];
}

// best_individual (accelerator.cl:270-313) over groups of `local2`, then host
// PickNBest over the group winners (accelerator.cc:866-883).
void best_acc(const float* vec, int P, int local2, int32_t* index, int32_t* best_size) {
    int ng = (P + local2 - 1) / local2;
    std::vector<float> PB((size_t)ng), lb((size_t)local2);
    std::vector<int32_t> PI((size_t)ng), li((size_t)local2);
    int np2 = 1;
    while (np2 < local2) np2 <<= 1;
    for (int g = 0; g < ng; ++g) {
        for (int l = 0; l < local2; ++l) {
            int gl = g * local2 + l;
            if (gl < P) { lb[l] = vec[gl]; li[l] = gl; }
            else { lb[l] = FLT_MAX; li[l] = -1; }
        }
        for (int s = np2 / 2; s > 0; s /= 2)
            for (int l = 0; l < s; ++l)
                if (l + s < local2 && lb[l] > lb[l + s]) { lb[l] = lb[l + s]; li[l] = li[l + s]; }
        PB[g] = lb[0];
        PI[g] = li[0];
    }
    if (*best_size > ng) *best_size = ng;
    pick_n_best(*best_size, index, ng, PB.data(), PI.data());
}

}  // namespace ppi

using namespace ppi;

// Geometry the reference's acc_interpret_init computes for this device (accelerator.cc:272-355).
static void reference_geometry(ppi_engine* E) {
    const int cu = E->cfg.compute_units > 0 ? E->cfg.compute_units : E->dev_cu;
    int mls1 = E->cfg.max_local_size > 0 ? E->cfg.max_local_size
                                          : (int)std::min<size_t>((size_t)E->dev_max_wg, E->dev_lds / 4);
    int mls2 = (int)std::min<size_t>((size_t)E->dev_max_wg, E->dev_lds / 8);
    if (E->nlin > 0) {
        E->dp_local = std::max(1, std::min(mls1, (int)std::ceil(E->nlin / (float)cu)));
        E->dp_groups = (E->nlin + E->dp_local - 1) / E->dp_local;
        E->pdp_local = E->nlin < mls1 ? E->nlin : mls1;
    }
    if (E->nprog > 0) E->best_local = std::max(1, std::min(mls2, (int)std::ceil(E->nprog / (float)cu)));
}

static void reset_derived(ppi_engine* E);

namespace ppi {

// Programs per FP tile: about 96 workgroups per CU, so that the launch ends in a short tail (a
// workgroup is a row tile x a chunk; on the regression workload 1,024 programs per chunk, 30
// workgroups per CU, measured 3% slower than 384-512), and at most 400 programs (at 7 waves per
// SIMD 400 measured 0.4% faster than 477 on the regression workload and 0.8% faster than 715 on
// the classification one, with 600 and more slower on both); at least 64 programs where that still
// leaves 4 workgroups per CU (fewer, and the LDS tile is staged for too little work); then equal
// chunks.
int32_t fp_default_chunk(int64_t nprog, int64_t nrb, int cu) {
    const int64_t items = nprog * nrb;
    const int64_t target = (int64_t)cu * 96;
    const int64_t floor = items >= (int64_t)cu * 4 * 64 ? 64 : 8;
    const int64_t per = (items + target - 1) / target;
    int64_t chunk = std::max<int64_t>(floor, std::min<int64_t>(kDefaultMaxChunk, per));
    const int64_t nch = (nprog + chunk - 1) / chunk;
    chunk = std::max<int64_t>(1, (nprog + nch - 1) / std::max<int64_t>(1, nch));
    return (int32_t)std::min<int64_t>(chunk, kMaxChunk);
}

bool compute_plan(const ppi_engine* E, const CompileStats& cs, int32_t nprog, int requested_form,
                         LaunchPlan* L, std::string* why) {
    int need_f = cs.max_col + 1;
    int F = need_f <= 8 ? 8 : (need_f <= 16 ? 16 : (need_f <= 32 ? 32 : 0));
    int S = cs.max_slots <= 9 ? 9 : 16;
    if (cs.max_slots > kMaxSlots) {
        *why = "program needs more than 16 register slots";
        return false;
    }
    int form = requested_form;
    if (form == PPI_FORM_AUTO) form = PPI_FORM_FP;
    // The FP and ordered forms stage at most 16 columns in LDS and read a wider dataset's other columns
    // through the caches: 6 waves per SIMD instead of the 4 a 32-column tile leaves room for (7% faster
    // on the 64-pair workload's 17-32 column datasets).  The PDP form holds its rows' features in
    // VGPRs and needs every column it reads among its F.
    if (form != PPI_FORM_PDP && form != PPI_FORM_PP && need_f > 16) F = 16;
    int R = E->cfg.rows_per_lane > 0 ? E->cfg.rows_per_lane : (form == PPI_FORM_PDP ? 2 : 4);
    auto shape_supported = [form](int r, int f, int s) {
        return form == PPI_FORM_PDP ? pdp_shape_supported(r, f, s) : ppi::shape_supported(r, f, s);
    };
    if (!shape_supported(R, F, S)) {
        // Nearest supported shape: keep F and S, lower R; then widen F.
        const int Rs[4] = {8, 4, 2, 1};
        bool found = false;
        for (int fi = 0; fi < 2 && !found; ++fi) {
            int FF = fi == 0 ? F : (F == 0 ? 0 : 32);
            for (int k = 0; k < 4 && !found; ++k)
                if (Rs[k] <= R && shape_supported(Rs[k], FF, S)) { R = Rs[k]; F = FF; found = true; }
            for (int k = 3; k >= 0 && !found; --k)
                if (shape_supported(Rs[k], FF, S)) { R = Rs[k]; F = FF; found = true; }
        }
        if (!found && shape_supported(1, 0, 16)) { R = 1; F = 0; S = 16; found = true; }
        if (!found) { *why = "no kernel shape"; return false; }
    }
    // The 4-slot stack where the shape has one: 20 VGPRs fewer at 4 rows, 40 at 8.
    if (S == 9 && cs.max_slots <= 4 && shape_supported(R, F, 4)) S = 4;
    L->form = form;
    L->R = R;
    L->F = F;
    L->S = S;
    int64_t rows_per_block = fp_rows_per_block(R);
    int64_t nrb = (E->nlin + rows_per_block - 1) / rows_per_block;
    int chunk = E->cfg.programs_per_tile;
    if (chunk <= 0) chunk = fp_default_chunk(nprog, nrb, E->dev_cu);
    L->chunk = std::min(chunk, kMaxChunk);
    L->nrb = (int)nrb;
    return true;
}

}  // namespace ppi

static bool plan(ppi_engine* E, std::string* why) {
    LaunchPlan L;
    if (!compute_plan(E, E->tape.stats, E->nprog, E->cfg.form, &L, why)) return false;
    E->form = L.form;
    E->R = L.R;
    E->F = L.F;
    E->S = L.S;
    E->chunk = L.chunk;
    E->nrb = L.nrb;
    return true;
}

// best_individual's work-group size for `nprog` programs (accelerator.cc:331-337).
int ppi::best_local_for(const ppi_engine* E, int32_t nprog) {
    const int cu = E->cfg.compute_units > 0 ? E->cfg.compute_units : E->dev_cu;
    const int mls2 = (int)std::min<size_t>((size_t)E->dev_max_wg, E->dev_lds / 8);
    return std::max(1, std::min(mls2, (int)std::ceil(nprog / (float)cu)));
}

extern "C" {

int32_t ppi_abi_version(void) { return PPI_ABI_VERSION; }

const char* ppi_last_error(void) { return g_last_error.c_str(); }

ppi_status ppi_device_count(int32_t* count) {
    if (!count) { set_error("count is NULL"); return PPI_E_ARG; }
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) { *count = 0; set_error(hipGetErrorString(e)); return PPI_E_NODEVICE; }
    *count = n;
    return PPI_OK;
}

ppi_status ppi_engine_create(ppi_engine** out, int32_t device) {
    if (!out) { set_error("out is NULL"); return PPI_E_ARG; }
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
        set_error("no HIP device");
        return PPI_E_NODEVICE;
    }
    if (device < 0) device = 0;
    if (device >= n) {
        set_error("valid device range: [0, " + std::to_string(n - 1) + "]");
        return PPI_E_ARG;
    }
    ppi_engine* E = new ppi_engine();
    E->device = device;
    hipError_t e = hipSetDevice(device);
    hipDeviceProp_t prop;
    if (e == hipSuccess) e = hipGetDeviceProperties(&prop, device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&E->stream, hipStreamNonBlocking);
    for (int k = 0; k < 4 && e == hipSuccess; ++k) e = hipEventCreate(&E->ev[k]);
    if (e != hipSuccess) {
        set_error(std::string("device init: ") + hipGetErrorString(e));
        delete E;
        return PPI_E_DEVICE;
    }
    // The no-allocation entry point's work area on this device (its address, once per device).
    if (ppi_status ss = scratch_setup(device)) {
        for (auto& ev : E->ev)
            if (ev) (void)hipEventDestroy(ev);
        (void)hipStreamDestroy(E->stream);
        delete E;
        return ss;
    }
    E->dev_cu = prop.multiProcessorCount;
    E->dev_max_wg = prop.maxThreadsPerBlock;
    E->dev_lds = prop.sharedMemPerBlock;
    E->arch = prop.gcnArchName;
    std::memset(&E->cfg, 0, sizeof(E->cfg));
    E->cfg.strategy = PPI_STRATEGY_SEQ;
    E->cfg.error_fn = PPI_EF_ABS;
    *out = E;
    return PPI_OK;
}

void ppi_engine_destroy(ppi_engine* E) {
    if (!E) return;
    (void)hipSetDevice(E->device);
    if (E->stream) (void)hipStreamSynchronize(E->stream);
    void* dev[] = {E->dX, E->dMeta, E->dCode, E->dPartial, E->dBadLo, E->dBadHi, E->dRaw, E->dVec,
                   E->dOrder, E->dCost, E->dEf, E->dPred, E->dDerOps, E->dDerSrc, E->dRowErr, E->dOSum, E->dOStop,
                   E->dRankVec, E->dRankGv, E->dRankGi, E->dRankIdx, E->dDeepCount, E->dDeepProg, E->dDeepErr};
    for (void* p : dev)
        if (p) (void)hipFree(p);
    E->folds.release();
    if (E->hRaw) (void)hipHostFree(E->hRaw);
    if (E->hVec) (void)hipHostFree(E->hVec);
    for (auto& ev : E->ev)
        if (ev) (void)hipEventDestroy(ev);
    if (E->stream) (void)hipStreamDestroy(E->stream);
    delete E;
}

ppi_status ppi_engine_configure(ppi_engine* E, const ppi_config* cfg, const ppi_symbol* table, int32_t ntable) {
    if (!E || !cfg) { set_error("engine or config is NULL"); return PPI_E_ARG; }
    if (cfg->strategy < PPI_STRATEGY_SEQ || cfg->strategy > PPI_STRATEGY_PDP) { set_error("bad strategy"); return PPI_E_ARG; }
    if (cfg->form < PPI_FORM_AUTO || cfg->form > PPI_FORM_ORDERED) { set_error("bad form"); return PPI_E_ARG; }
    if (cfg->form == PPI_FORM_ORDERED && cfg->strategy != PPI_STRATEGY_SEQ && cfg->strategy != PPI_STRATEGY_PP) {
        set_error("the ordered form serves the row-order strategies (SEQ, PP) only");
        return PPI_E_ARG;
    }
    if (cfg->ordered_pass_rows < 0) { set_error("bad ordered_pass_rows"); return PPI_E_ARG; }
    if (cfg->max_local_size < 0 || cfg->max_local_size > kPdpMaxLanes) {
        set_error("max_local_size must be in [0, " + std::to_string(kPdpMaxLanes) + "]");
        return PPI_E_ARG;
    }
    if (cfg->max_local_size == 0 && std::min<size_t>((size_t)E->dev_max_wg, E->dev_lds / 4) > (size_t)kPdpMaxLanes) {
        set_error("this device's work-group limit exceeds " + std::to_string(kPdpMaxLanes) + "; set max_local_size");
        return PPI_E_ARG;
    }
    if (cfg->error_fn < PPI_EF_ABS || cfg->error_fn > PPI_EF_EXPR) { set_error("bad error_fn"); return PPI_E_ARG; }
    std::vector<EfIns> ef;
    if (cfg->error_fn == PPI_EF_EXPR) {
        std::string err;
        if (!cfg->ef_expr) { set_error("PPI_EF_EXPR needs ef_expr"); return PPI_E_ARG; }
        if (!ef_compile(cfg->ef_expr, &ef, &err)) { set_error("error function: " + err); return PPI_E_ARG; }
    }
    bool cost = cfg->error_fn == PPI_EF_COST || cfg->error_fn == PPI_EF_COST_CLAMP;
    if (cost && (cfg->cost_n <= 0 || !cfg->cost)) { set_error("COST error function needs cost_n > 0 and cost"); return PPI_E_ARG; }
    if (cfg->rows_per_lane != 0 && cfg->rows_per_lane != 1 && cfg->rows_per_lane != 2 && cfg->rows_per_lane != 4 &&
        cfg->rows_per_lane != 8) {
        set_error("rows_per_lane must be 0, 1, 2, 4 or 8");
        return PPI_E_ARG;
    }
    if (cfg->programs_per_tile < 0 || cfg->programs_per_tile > kMaxChunk) {
        set_error("programs_per_tile must be in [0, " + std::to_string(kMaxChunk) + "]");
        return PPI_E_ARG;
    }
    if (ntable < 0 || (ntable > 0 && !table)) { set_error("bad symbol table"); return PPI_E_ARG; }
    for (int32_t k = 0; k < ntable; ++k) {
        int op = table[k].op;
        bool ok = (op >= PPI_OP_UNKNOWN && op <= PPI_OP_VALUE) || (op >= PPI_OP_IF_THEN_ELSE && op <= PPI_OP_DIV) ||
                  (op >= PPI_OP_NOT && op <= PPI_OP_TAN);
        if (!ok) { set_error("symbol table entry " + std::to_string(k) + ": unknown opcode " + std::to_string(op)); return PPI_E_ARG; }
    }
    (void)hipSetDevice(E->device);
    E->cfg = *cfg;
    E->cost.assign(cfg->cost, cfg->cost + (cost ? (size_t)cfg->cost_n * cfg->cost_n : 0));
    E->cfg.cost = nullptr;
    E->cfg.ef_expr = nullptr;
    E->ef = ef;
    if (!ef.empty()) {
        size_t cap = E->cap_ef;
        ppi_status s = ensure(&E->dEf, &cap, ef.size());
        if (s) return s;
        E->cap_ef = cap;
        PPI_HIP(hipMemcpy(E->dEf, ef.data(), ef.size() * sizeof(EfIns), hipMemcpyHostToDevice));
    }
    if (cost) {
        size_t cap = E->cap_cost;
        ppi_status s = ensure(&E->dCost, &cap, E->cost.size());
        if (s) return s;
        E->cap_cost = cap;
        PPI_HIP(hipMemcpy(E->dCost, E->cost.data(), E->cost.size() * sizeof(float), hipMemcpyHostToDevice));
    }
    E->symtab.build(table, ntable);
    E->configured = true;
    E->nprog = 0;  // programs compile against the table: re-upload after reconfiguring
    reset_derived(E);  // derived values depend on PROTECTED / NATIVE
    reference_geometry(E);
    return PPI_OK;
}

// Derived columns (chains of unary operators over input columns, tape.h): at most
// 4096, and at most 2 GiB of dataset memory with them.
static void reset_derived(ppi_engine* E) {
    E->derived.reset(E->ncol, derived_capacity(E->ldx, E->ncol));
    E->derived_ready = 0;
    E->derived_epoch0 = E->derived.epoch;
}

// Computes the derived columns the last compile added.
static ppi_status stage_derived(ppi_engine* E) {
    return stage_derived_columns(&E->dX, &E->cap_X, E->ldx, E->ncol, E->derived, &E->derived_ready, &E->dDerOps,
                                 &E->dDerSrc, &E->cap_der, E->cfg.protected_ops != 0, E->cfg.native_math != 0,
                                 E->stream);
}

static ppi_status stage_dataset(ppi_engine* E, int32_t nlin, int32_t ncol, const float* const* rows, const float* block) {
    if (nlin <= 0 || ncol <= 0) { set_error("dataset needs nlin > 0 and ncol > 0"); return PPI_E_ARG; }
    (void)hipSetDevice(E->device);
    const int64_t ldx = ((int64_t)nlin + 2047) / 2048 * 2048;
    std::vector<float> colmajor((size_t)ldx * ncol, 0.0f);
    for (int32_t c = 0; c < ncol; ++c) {
        float* dst = colmajor.data() + (size_t)c * ldx;
        for (int32_t r = 0; r < nlin; ++r) dst[r] = rows ? rows[r][c] : block[(size_t)r * ncol + c];
    }
    size_t cap = E->cap_X;
    ppi_status s = ensure(&E->dX, &cap, colmajor.size());
    if (s) return s;
    E->cap_X = cap;
    PPI_HIP(hipMemcpy(E->dX, colmajor.data(), colmajor.size() * sizeof(float), hipMemcpyHostToDevice));
    if (nlin >= kDeepMinRows) {
        if (!E->dDeepCount && (s = ensure(&E->dDeepCount, &cap, 1))) return s;
        if (!E->dDeepProg && (s = ensure(&E->dDeepProg, &cap, (size_t)kDeepCap))) return s;
        cap = E->cap_deep;
        if ((s = ensure(&E->dDeepErr, &cap, (size_t)kDeepCap * ldx))) return s;
        E->cap_deep = cap;
    }
    E->nlin = nlin;
    E->ncol = ncol;
    E->ldx = ldx;
    E->nprog = 0;  // attribute bounds depend on ncol: re-upload programs
    reset_derived(E);
    reference_geometry(E);
    return PPI_OK;
}

ppi_status ppi_engine_set_dataset(ppi_engine* E, const float* const* input, int32_t nlin, int32_t ncol) {
    if (!E || !input) { set_error("engine or input is NULL"); return PPI_E_ARG; }
    return stage_dataset(E, nlin, ncol, input, nullptr);
}

ppi_status ppi_engine_set_dataset_rowmajor(ppi_engine* E, const float* data, int32_t nlin, int32_t ncol) {
    if (!E || !data) { set_error("engine or data is NULL"); return PPI_E_ARG; }
    return stage_dataset(E, nlin, ncol, nullptr, data);
}

ppi_status ppi_engine_upload(ppi_engine* E, const int32_t* phenotype, const float* ephemeral, const int32_t* size,
                             int32_t stride, int32_t n) {
    if (!E) { set_error("engine is NULL"); return PPI_E_ARG; }
    if (!E->configured) { set_error("configure the engine first"); return PPI_E_STATE; }
    if (!E->dX) { set_error("set a dataset first"); return PPI_E_STATE; }
    if (n <= 0 || stride <= 0 || !phenotype || !ephemeral || !size) { set_error("bad program arrays"); return PPI_E_ARG; }
    (void)hipSetDevice(E->device);
    double t0 = now_ms();
    std::string err;
    int threads = (int)std::min<unsigned>(16u, std::max(1u, std::thread::hardware_concurrency()));
    // The FP and ordered forms read unary-of-input chains from precomputed derived columns.
    const bool fp_form = E->cfg.form == PPI_FORM_AUTO || E->cfg.form == PPI_FORM_FP || E->cfg.form == PPI_FORM_ORDERED;
    DerivedColumns* der = (fp_form && E->derived.cap > 0) ? &E->derived : nullptr;
    if (!compile_population_mt(E->symtab, phenotype, ephemeral, size, stride, n, E->ncol, false, threads, &E->tape, &err,
                               der, E->cfg.no_constant_folding == 0)) {
        set_error(err);
        return PPI_E_PROGRAM;
    }
    ppi_status fs = evaluate_folds(&E->tape, &E->folds, E->cfg.protected_ops != 0, E->cfg.native_math != 0, E->stream);
    if (fs) return fs;
    E->tape.code.insert(E->tape.code.end(), kTapePad, Ins{H_END, kEmpty});  // window / look-ahead padding
    E->nprog = n;
    reference_geometry(E);
    std::string why;
    if (!plan(E, &why)) { set_error(why); return PPI_E_PROGRAM; }
    ppi_status s;
    size_t cap;
    cap = E->cap_meta; if ((s = ensure(&E->dMeta, &cap, E->tape.meta.size()))) return s; E->cap_meta = cap;
    cap = E->cap_code; if ((s = ensure(&E->dCode, &cap, E->tape.code.size()))) return s; E->cap_code = cap;
    cap = E->cap_raw; if ((s = ensure(&E->dRaw, &cap, (size_t)n))) return s;
    cap = E->cap_raw; if ((s = ensure(&E->dVec, &cap, (size_t)n))) return s; E->cap_raw = cap;
    cap = E->cap_bad; if ((s = ensure(&E->dBadLo, &cap, (size_t)n))) return s;
    cap = E->cap_bad; if ((s = ensure(&E->dBadHi, &cap, (size_t)n))) return s; E->cap_bad = cap;
    cap = E->cap_hres; if ((s = ensure_host(&E->hRaw, &cap, (size_t)n))) return s;
    cap = E->cap_hres; if ((s = ensure_host(&E->hVec, &cap, (size_t)n))) return s; E->cap_hres = cap;
    if (E->form == PPI_FORM_FP) {
        cap = E->cap_partial;
        if ((s = ensure(&E->dPartial, &cap, (size_t)E->nrb * n))) return s;
        E->cap_partial = cap;
    }
    if (E->form == PPI_FORM_ORDERED) {
        // Row errors of one pass: nrb_pass row tiles x n programs x 64*R rows; by default as many
        // rows as fit min(16 GiB, half the free device memory), so most populations take one pass.
        const int64_t tile = fp_rows_per_block(E->R);
        size_t free_b = 0, total_b = 0;
        if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) free_b = (size_t)4 << 30;
        const int64_t budget = std::min<int64_t>((int64_t)16 << 30, (int64_t)(free_b / 2));
        int64_t per = E->cfg.ordered_pass_rows > 0 ? (E->cfg.ordered_pass_rows + tile - 1) / tile
                                                   : budget / ((int64_t)n * tile * (int64_t)sizeof(float));
        E->nrb_pass = (int)std::max<int64_t>(1, std::min<int64_t>(per, E->nrb));
        cap = E->cap_rowerr;
        if ((s = ensure(&E->dRowErr, &cap, (size_t)E->nrb_pass * n * tile))) return s;
        E->cap_rowerr = cap;
        cap = E->cap_osum; if ((s = ensure(&E->dOSum, &cap, (size_t)n))) return s;
        cap = E->cap_osum; if ((s = ensure(&E->dOStop, &cap, (size_t)n))) return s; E->cap_osum = cap;
    }
    if (der && (s = stage_derived(E))) return s;
    PPI_HIP(hipMemcpyAsync(E->dMeta, E->tape.meta.data(), E->tape.meta.size() * sizeof(ProgMeta), hipMemcpyHostToDevice, E->stream));
    PPI_HIP(hipMemcpyAsync(E->dCode, E->tape.code.data(), E->tape.code.size() * sizeof(Ins), hipMemcpyHostToDevice, E->stream));
    if (E->form == PPI_FORM_PP) {
        std::vector<int32_t> order((size_t)n);
        std::iota(order.begin(), order.end(), 0);
        std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) {
            return E->tape.meta[a].ncode > E->tape.meta[b].ncode;
        });
        cap = E->cap_order;
        if ((s = ensure(&E->dOrder, &cap, (size_t)n))) return s;
        E->cap_order = cap;
        PPI_HIP(hipMemcpyAsync(E->dOrder, order.data(), order.size() * sizeof(int32_t), hipMemcpyHostToDevice, E->stream));
    }
    PPI_HIP(hipStreamSynchronize(E->stream));
    E->stats.upload_ms = now_ms() - t0;
    E->stats.node_evals = (double)E->tape.stats.total_size * E->nlin;
    E->stats.executed_nodes = (double)E->tape.stats.total_code * E->nlin;
    E->stats.max_slots = E->tape.stats.max_slots;
    E->stats.folded_subtrees = (int32_t)E->tape.folds.meta.size();
    E->stats.folded_nodes = (int32_t)E->tape.stats.folded_nodes;
    return PPI_OK;
}

}  // extern "C"

EvalArgs ppi::make_args(ppi_engine* E, float alpha) {
    EvalArgs A{};
    A.meta = E->dMeta;
    A.code = E->dCode;
    A.X = E->dX;
    A.ldx = E->ldx;
    A.nlin = E->nlin;
    A.ncol = E->ncol;
    A.ycol = E->ncol - 1;
    A.nprog = E->nprog;
    A.chunk = E->chunk;
    A.strategy = E->cfg.strategy;
    A.error_fn = E->cfg.error_fn;
    A.reducemax = E->cfg.reducemax;
    A.cost = E->dCost;
    A.cost_n = E->cfg.cost_n;
    A.pdp_lanes = std::max(1, E->pdp_local);
    A.dp_groups = E->dp_groups;
    A.dp_local = std::max(1, E->dp_local);
    A.alpha = alpha;
    A.order = E->dOrder;
    A.partial = E->dPartial;
    A.bad_lo = E->dBadLo;
    A.bad_hi = E->dBadHi;
    A.raw = E->dRaw;
    A.vec = E->dVec;
    A.rb0 = 0;
    A.nrb_pass = E->nrb;
    A.rowerr = E->dRowErr;
    A.osum = E->dOSum;
    A.ostop = E->dOStop;
    A.ef = E->dEf;
    A.ef_n = (int32_t)E->ef.size();
    if (E->nlin >= kDeepMinRows && E->dDeepErr) {
        A.deep_count = E->dDeepCount;
        A.deep_prog = E->dDeepProg;
        A.deep_cap = kDeepCap;
    }
    return A;
}

// The recheck and its deep list on the engine stream.
static ppi_status enqueue_recheck(ppi_engine* E, const EvalArgs& A) {
    const bool P = E->cfg.protected_ops != 0, N = E->cfg.native_math != 0;
    if (A.deep_cap > 0) PPI_HIP(launch_fill_u32(A.deep_count, 0u, 1, E->stream));
    PPI_HIP(launch_recheck(A, P, N, E->stream));
    if (A.deep_cap > 0) {
        DeepArgs D{};
        D.base = A;
        D.err = E->dDeepErr;
        D.stride = E->ldx;
        D.max_rows = A.nlin;
        PPI_HIP(launch_deep(D, E->dev_cu * 4, P, N, E->stream));
    }
    return PPI_OK;
}

// Enqueue one evaluation on the engine stream: programs, work and result buffers as `A` names them,
// kernel form and shape as `L` says.
ppi_status ppi::enqueue_with(ppi_engine* E, EvalArgs A, const LaunchPlan& L, bool time_main) {
    const bool P = E->cfg.protected_ops != 0, N = E->cfg.native_math != 0;
    PPI_HIP(launch_fill_u32(A.bad_lo, 0xffffffffu, A.nprog, E->stream));
    PPI_HIP(launch_fill_u32(A.bad_hi, 0u, A.nprog, E->stream));
    if (time_main) PPI_HIP(hipEventRecord(E->ev[2], E->stream));
    if (A.error_fn == PPI_EF_EXPR) {
        // A user ERROR(X,Y): every program through the exact-order recomputation (the
        // reference's own partition and row order for the strategy), the expression
        // interpreted per row.
        PPI_HIP(launch_ef_prepare(A, E->stream));
        ppi_status s = enqueue_recheck(E, A);
        if (s) return s;
        if (time_main) PPI_HIP(hipEventRecord(E->ev[3], E->stream));
        return PPI_OK;
    }
    if (L.form == PPI_FORM_FP) {
        A.part_nrb = L.nrb;
        A.part_rows = fp_rows_per_block(L.R);
        PPI_HIP(launch_fp(A, L.R, L.F, L.S, P, N, E->stream));
        if (time_main) PPI_HIP(hipEventRecord(E->ev[3], E->stream));
        PPI_HIP(launch_finalize(A, L.nrb, E->stream));
    } else if (L.form == PPI_FORM_ORDERED) {
        // Passes of nrb_pass row tiles: per-row errors, then each program's row-order sum continues.
        PPI_HIP(launch_fill_u32(reinterpret_cast<uint32_t*>(A.osum), 0u, A.nprog, E->stream));  // 0.0f
        PPI_HIP(launch_fill_u32(A.ostop, 0u, A.nprog, E->stream));
        for (int rb0 = 0; rb0 < L.nrb; rb0 += L.nrb_pass) {
            A.rb0 = rb0;
            A.nrb_pass = std::min(L.nrb_pass, L.nrb - rb0);
            PPI_HIP(launch_fp(A, L.R, L.F, L.S, P, N, E->stream, true));
            PPI_HIP(launch_ordered_sum(A, fp_rows_per_block(L.R), rb0 + A.nrb_pass >= L.nrb, E->stream));
        }
        if (time_main) PPI_HIP(hipEventRecord(E->ev[3], E->stream));
    } else if (L.form == PPI_FORM_PDP) {
        PPI_HIP(launch_pdp(A, L.R, L.F, L.S, P, N, E->stream));
        if (time_main) PPI_HIP(hipEventRecord(E->ev[3], E->stream));
    } else {
        PPI_HIP(launch_pp(A, P, N, E->stream));
        if (time_main) PPI_HIP(hipEventRecord(E->ev[3], E->stream));
    }
    // The programs whose result depends on the reference's partition or order of the rows.
    return enqueue_recheck(E, A);
}

// Enqueue one evaluation of the resident programs on the engine stream.
static ppi_status enqueue_eval(ppi_engine* E, float alpha, bool time_main) {
    LaunchPlan L;
    L.form = E->form;
    L.R = E->R;
    L.F = E->F;
    L.S = E->S;
    L.chunk = E->chunk;
    L.nrb = E->nrb;
    L.nrb_pass = E->nrb_pass;
    return enqueue_with(E, make_args(E, alpha), L, time_main);
}

extern "C" {

ppi_status ppi_engine_launch(ppi_engine* E) {
    if (!E) { set_error("engine is NULL"); return PPI_E_ARG; }
    if (E->nprog <= 0) { set_error("upload programs first"); return PPI_E_STATE; }
    (void)hipSetDevice(E->device);
    PPI_HIP(hipEventRecord(E->ev[0], E->stream));
    ppi_status s = enqueue_eval(E, 0.0f, true);
    if (s) return s;
    PPI_HIP(hipMemcpyAsync(E->hRaw, E->dRaw, (size_t)E->nprog * sizeof(float), hipMemcpyDeviceToHost, E->stream));
    PPI_HIP(hipEventRecord(E->ev[1], E->stream));
    E->launched = true;
    return PPI_OK;
}

ppi_status ppi_engine_wait(ppi_engine* E, float alpha, float* vector, int32_t* index, int32_t* best_size) {
    if (!E) { set_error("engine is NULL"); return PPI_E_ARG; }
    if (!E->launched) { set_error("nothing launched"); return PPI_E_STATE; }
    E->launched = false;
    PPI_HIP(hipEventSynchronize(E->ev[1]));
    float ms = 0.0f;
    PPI_HIP(hipEventElapsedTime(&ms, E->ev[0], E->ev[1]));
    E->stats.kernel_ms = ms;
    PPI_HIP(hipEventElapsedTime(&ms, E->ev[2], E->ev[3]));
    E->stats.main_kernel_ms = ms;
    E->stats.form = E->form;
    E->stats.rows_per_lane = E->R;
    E->stats.programs_per_tile = E->chunk;
    double t0 = now_ms();
    const int n = E->nprog;
    // The kernels ran with alpha = 0; the alpha*size term is added here in the
    // reference's own order and precision (host epilogues, accelerator.cc:801,849).
    const float fn = (float)E->nlin;
    for (int i = 0; i < n; ++i) {
        const float raw = E->hRaw[i];
        const ProgMeta& m = E->tape.meta[i];
        const float as = alpha * (float)m.size;
        float v;
        switch (E->cfg.strategy) {
        case PPI_STRATEGY_PP:
        case PPI_STRATEGY_PDP:
            v = raw + as;
            break;
        default:
            // SEQ / FP: the kernel wrote sum/nlin + 0*size (or FLT_MAX); recover the
            // alpha-inclusive value exactly as sum/nlin + alpha*size.
            v = (raw == FLT_MAX || (m.flags & kEmpty)) ? raw : raw + as;
            break;
        }
        E->hVec[i] = v;
        (void)fn;
    }
    if (vector) std::memcpy(vector, E->hVec, (size_t)n * sizeof(float));
    if (index && best_size && *best_size > 0) {
        if (E->cfg.strategy == PPI_STRATEGY_SEQ) {
            if (*best_size > n) *best_size = n;
            pick_n_best(*best_size, index, n, E->hVec, nullptr);
        } else if (E->cfg.strategy == PPI_STRATEGY_FP) {
            best_acc(E->hVec, n, E->best_local, index, best_size);
        } else {
            best_acc(E->hRaw, n, E->best_local, index, best_size);
        }
    }
    E->stats.download_ms = now_ms() - t0;
    return PPI_OK;
}

ppi_status ppi_engine_evaluate(ppi_engine* E, const int32_t* phenotype, const float* ephemeral, const int32_t* size,
                               int32_t stride, int32_t n, float alpha, float* vector, int32_t* index,
                               int32_t* best_size) {
    ppi_status s = ppi_engine_upload(E, phenotype, ephemeral, size, stride, n);
    if (s) return s;
    s = ppi_engine_launch(E);
    if (s) return s;
    return ppi_engine_wait(E, alpha, vector, index, best_size);
}

ppi_status ppi_engine_predict(ppi_engine* E, const int32_t* phenotype, const float* ephemeral, int32_t size, float* out) {
    if (!E || !phenotype || !ephemeral || !out || size < 0) { set_error("bad arguments"); return PPI_E_ARG; }
    if (!E->configured || !E->dX) { set_error("configure the engine and set a dataset first"); return PPI_E_STATE; }
    (void)hipSetDevice(E->device);
    Tape t;
    std::string err;
    int32_t sz = size;
    int32_t stride = std::max(1, size);
    if (!compile_population(E->symtab, phenotype, ephemeral, &sz, stride, 1, E->ncol, false, &t, &err)) {
        set_error(err);
        return PPI_E_PROGRAM;
    }
    if (t.stats.max_slots > kMaxSlots) { set_error("program needs more than 16 register slots"); return PPI_E_PROGRAM; }
    t.code.insert(t.code.end(), kTapePad, Ins{H_END, kEmpty});
    ppi_status s;
    size_t cap;
    cap = E->cap_meta; if ((s = ensure(&E->dMeta, &cap, 1))) return s; E->cap_meta = cap;
    cap = E->cap_code; if ((s = ensure(&E->dCode, &cap, t.code.size()))) return s; E->cap_code = cap;
    cap = E->cap_pred; if ((s = ensure(&E->dPred, &cap, (size_t)E->nlin))) return s; E->cap_pred = cap;
    PPI_HIP(hipMemcpyAsync(E->dMeta, t.meta.data(), sizeof(ProgMeta), hipMemcpyHostToDevice, E->stream));
    PPI_HIP(hipMemcpyAsync(E->dCode, t.code.data(), t.code.size() * sizeof(Ins), hipMemcpyHostToDevice, E->stream));
    E->nprog = 0;  // the resident program set was replaced
    EvalArgs A = make_args(E, 0.0f);
    A.nprog = 1;
    int F = E->ncol <= 32 ? 32 : 0;
    PPI_HIP(launch_predict(A, F, 16, E->cfg.protected_ops != 0, E->cfg.native_math != 0, E->dPred, E->stream));
    PPI_HIP(hipMemcpyAsync(out, E->dPred, (size_t)E->nlin * sizeof(float), hipMemcpyDeviceToHost, E->stream));
    PPI_HIP(hipStreamSynchronize(E->stream));
    return PPI_OK;
}

ppi_status ppi_engine_evaluate_noalloc(ppi_engine* E, const int32_t* phenotype, const float* ephemeral,
                                      const int32_t* size, int32_t stride, int32_t n, float alpha, float* vector,
                                      int32_t* index, int32_t* best_size) {
    if (!E) { set_error("engine is NULL"); return PPI_E_ARG; }
    if (!E->configured) { set_error("configure the engine first"); return PPI_E_STATE; }
    if (!E->dX) { set_error("set a dataset first"); return PPI_E_STATE; }
    if (n <= 0 || stride <= 0 || !phenotype || !ephemeral || !size || !vector) {
        set_error("bad program arrays");
        return PPI_E_ARG;
    }
    int form = E->cfg.form == PPI_FORM_AUTO ? PPI_FORM_FP : E->cfg.form;
    if (form != PPI_FORM_FP && form != PPI_FORM_PDP) {
        set_error("the no-allocation entry point runs the FP and PDP forms (configured form " +
                  std::to_string(E->cfg.form) + ")");
        return PPI_E_UNSUPPORTED;
    }
    (void)hipSetDevice(E->device);
    uint8_t* base = static_cast<uint8_t*>(scratch_region(E->device));
    if (!base) { set_error("no scratch region on this device"); return PPI_E_STATE; }
    // Host-side compile (ordinary heap memory); no derived columns, which would grow the dataset buffer.
    Tape t;
    std::string err;
    int threads = (int)std::min<unsigned>(16u, std::max(1u, std::thread::hardware_concurrency()));
    if (!compile_population_mt(E->symtab, phenotype, ephemeral, size, stride, n, E->ncol, false, threads, &t, &err,
                               nullptr, E->cfg.no_constant_folding == 0)) {
        set_error(err);
        return PPI_E_PROGRAM;
    }
    t.code.insert(t.code.end(), kTapePad, Ins{H_END, kEmpty});
    LaunchPlan L;
    if (!compute_plan(E, t.stats, n, form, &L, &err)) { set_error(err); return PPI_E_PROGRAM; }
    // Sections of the scratch region, 256-byte aligned.
    size_t off = 0;
    auto carve = [&off](size_t bytes) {
        const size_t o = off;
        off = (off + bytes + 255) & ~(size_t)255;
        return o;
    };
    const size_t nf = t.folds.meta.size();
    const size_t o_meta = carve((size_t)n * sizeof(ProgMeta));
    const size_t o_code = carve(t.code.size() * sizeof(Ins));
    const size_t o_part = carve(L.form == PPI_FORM_FP ? (size_t)L.nrb * (size_t)n * sizeof(float) : 0);
    const size_t o_blo = carve((size_t)n * sizeof(uint32_t));
    const size_t o_bhi = carve((size_t)n * sizeof(uint32_t));
    const size_t o_raw = carve((size_t)n * sizeof(float));
    const size_t o_vec = carve((size_t)n * sizeof(float));
    const size_t o_fmeta = carve(nf * sizeof(ProgMeta));
    const size_t o_fcode = carve(t.folds.code.size() * sizeof(Ins));
    const size_t o_fout = carve(nf * sizeof(float));
    if (off > kScratchBytes) {
        set_error("evaluate_noalloc: " + std::to_string(n) + " programs over " + std::to_string(E->nlin) +
                  " rows need " + std::to_string(off) + " bytes of device scratch; the limit is " +
                  std::to_string(kScratchBytes) + " bytes");
        return PPI_E_ARG;
    }
    const bool P = E->cfg.protected_ops != 0, N = E->cfg.native_math != 0;
    if (nf > 0) {
        // Constant subtrees: evaluated once on the device, values written into the tape.
        std::vector<float> vals(nf);
        auto* fmeta = reinterpret_cast<ProgMeta*>(base + o_fmeta);
        auto* fcode = reinterpret_cast<Ins*>(base + o_fcode);
        auto* fout = reinterpret_cast<float*>(base + o_fout);
        PPI_HIP(hipMemcpyAsync(fmeta, t.folds.meta.data(), nf * sizeof(ProgMeta), hipMemcpyHostToDevice, E->stream));
        PPI_HIP(hipMemcpyAsync(fcode, t.folds.code.data(), t.folds.code.size() * sizeof(Ins), hipMemcpyHostToDevice,
                               E->stream));
        PPI_HIP(launch_fold(fmeta, fcode, (int32_t)nf, fout, P, N, E->stream));
        PPI_HIP(hipMemcpyAsync(vals.data(), fout, nf * sizeof(float), hipMemcpyDeviceToHost, E->stream));
        PPI_HIP(hipStreamSynchronize(E->stream));
        apply_folds(&t, vals.data());
    }
    EvalArgs A = make_args(E, 0.0f);
    A.meta = reinterpret_cast<const ProgMeta*>(base + o_meta);
    A.code = reinterpret_cast<const Ins*>(base + o_code);
    A.nprog = n;
    A.chunk = L.chunk;
    A.order = nullptr;
    A.partial = L.form == PPI_FORM_FP ? reinterpret_cast<float*>(base + o_part) : nullptr;
    A.bad_lo = reinterpret_cast<uint32_t*>(base + o_blo);
    A.bad_hi = reinterpret_cast<uint32_t*>(base + o_bhi);
    A.raw = reinterpret_cast<float*>(base + o_raw);
    A.vec = reinterpret_cast<float*>(base + o_vec);
    A.deep_cap = 0;  // every buffer of this call lives in the scratch array: no deep list
    A.rb0 = 0;
    A.nrb_pass = L.nrb;
    A.rowerr = nullptr;
    A.osum = nullptr;
    A.ostop = nullptr;
    PPI_HIP(hipMemcpyAsync(base + o_meta, t.meta.data(), t.meta.size() * sizeof(ProgMeta), hipMemcpyHostToDevice,
                           E->stream));
    PPI_HIP(hipMemcpyAsync(base + o_code, t.code.data(), t.code.size() * sizeof(Ins), hipMemcpyHostToDevice,
                           E->stream));
    ppi_status s = enqueue_with(E, A, L, false);
    if (s) return s;
    PPI_HIP(hipMemcpyAsync(vector, A.raw, (size_t)n * sizeof(float), hipMemcpyDeviceToHost, E->stream));
    PPI_HIP(hipStreamSynchronize(E->stream));
    // The epilogue of ppi_engine_wait: PP / PDP rank the kernel's values before alpha*size is added,
    // SEQ / FP rank the alpha-inclusive vector (accelerator.cc:843-889, sequential.cc:197-208).
    const int strategy = E->cfg.strategy;
    const bool rank = index && best_size && *best_size > 0;
    if (rank && (strategy == PPI_STRATEGY_PP || strategy == PPI_STRATEGY_PDP))
        best_acc(vector, n, best_local_for(E, n), index, best_size);
    for (int i = 0; i < n; ++i) {
        const float raw = vector[i];
        const float as = alpha * (float)t.meta[i].size;
        if (strategy == PPI_STRATEGY_PP || strategy == PPI_STRATEGY_PDP)
            vector[i] = raw + as;
        else
            vector[i] = (raw == FLT_MAX || (t.meta[i].flags & kEmpty)) ? raw : raw + as;
    }
    if (rank && strategy == PPI_STRATEGY_SEQ) {
        if (*best_size > n) *best_size = n;
        pick_n_best(*best_size, index, n, vector, nullptr);
    } else if (rank && strategy == PPI_STRATEGY_FP) {
        best_acc(vector, n, best_local_for(E, n), index, best_size);
    }
    return PPI_OK;
}

ppi_status ppi_engine_evaluate_ranked(ppi_engine* E, const int32_t* phenotype, const float* ephemeral,
                                      const int32_t* size, int32_t stride, int32_t n, float alpha, float* vector,
                                      int32_t* index, int32_t* best_size) {
    if (!vector) { set_error("vector is NULL"); return PPI_E_ARG; }
    if (index && best_size && *best_size > kRankMax) {
        set_error("best_size above " + std::to_string(kRankMax));
        return PPI_E_ARG;
    }
    ppi_status s = ppi_engine_upload(E, phenotype, ephemeral, size, stride, n);
    if (s) return s;
    if (E->cap_rank < (size_t)n || !E->dRankVec) {
        size_t c = 0;
        if ((s = ensure(&E->dRankVec, &c, (size_t)n))) return s;
        c = 0;
        if ((s = ensure(&E->dRankGv, &c, (size_t)n))) return s;
        c = 0;
        if ((s = ensure(&E->dRankGi, &c, (size_t)n))) return s;
        c = 0;
        if ((s = ensure(&E->dRankIdx, &c, (size_t)kRankMax))) return s;
        E->cap_rank = (size_t)n;
    }
    if ((s = enqueue_eval(E, 0.0f, true))) return s;
    // the epilogue of ppi_engine_wait, on the device
    int32_t k = (index && best_size && *best_size > 0) ? *best_size : 0;
    if (k > 0) {
        const int32_t limit = E->cfg.strategy == PPI_STRATEGY_SEQ ? n : (n + E->best_local - 1) / E->best_local;
        k = std::min(k, limit);
    }
    PPI_HIP(launch_rank(E->dRaw, E->dMeta, n, E->cfg.strategy, alpha, E->best_local, k, E->dRankVec, E->dRankGv,
                        E->dRankGi, E->dRankIdx, E->stream));
    PPI_HIP(hipMemcpyAsync(vector, E->dRankVec, (size_t)n * sizeof(float), hipMemcpyDeviceToHost, E->stream));
    if (k > 0) PPI_HIP(hipMemcpyAsync(index, E->dRankIdx, (size_t)k * sizeof(int32_t), hipMemcpyDeviceToHost, E->stream));
    PPI_HIP(hipStreamSynchronize(E->stream));
    if (index && best_size && *best_size > 0) *best_size = k;
    return PPI_OK;
}

ppi_status ppi_engine_resident_tape(const ppi_engine* E, uint32_t* meta, uint32_t* code, int64_t* n_code) {
    if (!E || !n_code) { set_error("bad arguments"); return PPI_E_ARG; }
    if (E->nprog <= 0) { set_error("upload programs first"); return PPI_E_STATE; }
    *n_code = (int64_t)E->tape.code.size();
    if (meta) std::memcpy(meta, E->tape.meta.data(), E->tape.meta.size() * sizeof(ProgMeta));
    if (code) std::memcpy(code, E->tape.code.data(), E->tape.code.size() * sizeof(Ins));
    return PPI_OK;
}

ppi_status ppi_ef_evaluate(const char* text, const float* X, const float* Y, int32_t n, float* out) {
    if (!text || n < 0 || (n > 0 && (!X || !Y || !out))) { set_error("bad arguments"); return PPI_E_ARG; }
    std::vector<EfIns> ef;
    std::string err;
    if (!ef_compile(text, &ef, &err)) { set_error("error function: " + err); return PPI_E_ARG; }
    for (int32_t i = 0; i < n; ++i) out[i] = ef_eval_host(ef.data(), (int)ef.size(), X[i], Y[i]);
    return PPI_OK;
}

ppi_status ppi_engine_get_stats(const ppi_engine* E, ppi_stats* out) {
    if (!E || !out) { set_error("engine or out is NULL"); return PPI_E_ARG; }
    *out = E->stats;
    out->form = E->form;
    out->rows_per_lane = E->R;
    out->programs_per_tile = E->chunk;
    out->derived_columns = E->derived_ready;
    out->derived_refills = (int32_t)(E->derived.epoch - E->derived_epoch0);
    return PPI_OK;
}

ppi_status ppi_engine_time(ppi_engine* E, int32_t iters, double* ms_per_eval) {
    if (!E || !ms_per_eval || iters <= 0) { set_error("bad arguments"); return PPI_E_ARG; }
    if (E->nprog <= 0) { set_error("upload programs first"); return PPI_E_STATE; }
    (void)hipSetDevice(E->device);
    PPI_HIP(hipEventRecord(E->ev[0], E->stream));
    double main_ms = 0.0;
    for (int32_t k = 0; k < iters; ++k) {
        ppi_status s = enqueue_eval(E, 0.0f, true);
        if (s) return s;
        PPI_HIP(hipEventSynchronize(E->ev[3]));
        float ms = 0.0f;
        PPI_HIP(hipEventElapsedTime(&ms, E->ev[2], E->ev[3]));
        main_ms += ms;
    }
    PPI_HIP(hipEventRecord(E->ev[1], E->stream));
    PPI_HIP(hipEventSynchronize(E->ev[1]));
    float ms = 0.0f;
    PPI_HIP(hipEventElapsedTime(&ms, E->ev[0], E->ev[1]));
    *ms_per_eval = ms / iters;
    E->stats.main_kernel_ms = main_ms / iters;
    E->stats.kernel_ms = ms / iters;
    return PPI_OK;
}

}  // extern "C"
