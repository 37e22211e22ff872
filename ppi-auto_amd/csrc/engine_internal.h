// engine_internal.h — state of one engine object (opaque behind include/ppi_engine.h).
#ifndef PPI_ENGINE_INTERNAL_H
#define PPI_ENGINE_INTERNAL_H

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/ppi_engine.h"
#include "kernels.h"
#include "tape.h"

namespace ppi {
void set_error(const std::string& s);
// Derived columns (tape.h): the most a dataset of this shape may hold (256, and 2 GiB with its input columns).
int32_t derived_capacity(int64_t ldx, int32_t ncol);
// Grows X (keeping its ncol + *ready columns) to hold every column of `table` and computes the
// columns added since *ready on stream `st`.
ppi_status stage_derived_columns(float** X, size_t* cap_floats, int64_t ldx, int32_t ncol, DerivedColumns& table,
                                 int32_t* ready, uint32_t** dOps, int32_t** dSrc, size_t* cap_list, bool P, bool N,
                                 hipStream_t st);
// Programs per FP workgroup of the default plan for nprog programs over nrb row tiles on a device
// of cu compute units (the FP form's launch is nrb x ceil(nprog / chunk) workgroups).
int32_t fp_default_chunk(int64_t nprog, int64_t nrb, int cu);
// Device buffers of a tape's fold programs (tape.h: FoldSet), kept across uploads.
struct FoldBuffers {
    ProgMeta* meta = nullptr;
    size_t cap_meta = 0;
    Ins* code = nullptr;
    size_t cap_code = 0;
    float* out = nullptr;
    size_t cap_out = 0;
    std::vector<float> host;
    void release();
};
// Evaluates the tape's folds on stream `st` with the interpreters' operator code
// (P = PROTECTED, N = NATIVE) and writes their values into the tape (apply_folds).
ppi_status evaluate_folds(Tape* t, FoldBuffers* B, bool P, bool N, hipStream_t st);
void pick_n_best(int n, int32_t* best, int num, const float* errors, const int32_t* indices);
// Device work area of ppi_engine_evaluate_noalloc (kern_scratch.hip): a static array of
// kScratchBytes on every device the library runs on.  scratch_setup(device), with that device
// current, takes its device address once (hipGetSymbolAddress); scratch_region(device) returns it
// (null before set-up).  Calls that use it at the same time on one device race.
constexpr size_t kScratchBytes = (size_t)64 << 20;
ppi_status scratch_setup(int device);
void* scratch_region(int device);
void best_acc(const float* vec, int P, int local2, int32_t* index, int32_t* best_size);
}  // namespace ppi

struct ppi_engine {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};  // eval begin/end, main kernel begin/end
    std::string arch;
    int dev_cu = 0;
    int dev_max_wg = 0;
    size_t dev_lds = 0;

    bool configured = false;
    ppi_config cfg{};
    std::vector<float> cost;
    ppi::SymbolTable symtab;

    // dataset, column-major with rows padded to ldx; derived columns follow the ncol input columns
    int32_t nlin = 0, ncol = 0;
    int64_t ldx = 0;
    float* dX = nullptr;
    size_t cap_X = 0;
    ppi::DerivedColumns derived;
    int32_t derived_ready = 0;  // derived columns already computed on the device
    uint32_t derived_epoch0 = 0; // table epoch when the dataset or configuration was set
    uint32_t* dDerOps = nullptr;
    int32_t* dDerSrc = nullptr;
    size_t cap_der = 0;

    // resident programs
    ppi::Tape tape;
    int32_t nprog = 0;
    ppi::ProgMeta* dMeta = nullptr;
    size_t cap_meta = 0;
    ppi::Ins* dCode = nullptr;
    size_t cap_code = 0;
    int32_t* dOrder = nullptr;
    size_t cap_order = 0;

    // work and result buffers
    float* dPartial = nullptr;
    size_t cap_partial = 0;
    uint32_t* dBadLo = nullptr;
    uint32_t* dBadHi = nullptr;
    size_t cap_bad = 0;
    float* dRaw = nullptr;
    float* dVec = nullptr;
    size_t cap_raw = 0;
    float* dCost = nullptr;
    size_t cap_cost = 0;
    std::vector<ppi::EfIns> ef;  // PPI_EF_EXPR: the compiled ERROR(X,Y)
    ppi::EfIns* dEf = nullptr;
    size_t cap_ef = 0;
    float* dPred = nullptr;
    size_t cap_pred = 0;
    float* hRaw = nullptr;
    float* hVec = nullptr;
    size_t cap_hres = 0;
    // ordered form: one pass of per-row errors, running sums and stop flags
    float* dRowErr = nullptr;
    size_t cap_rowerr = 0;
    float* dOSum = nullptr;
    uint32_t* dOStop = nullptr;
    size_t cap_osum = 0;
    int nrb_pass = 1;
    // the recheck's deep list (datasets of kDeepMinRows rows or more): kDeepCap entries of ldx rows
    uint32_t* dDeepCount = nullptr;
    int32_t* dDeepProg = nullptr;
    float* dDeepErr = nullptr;
    size_t cap_deep = 0;
    // fold programs of the last upload (constant subtrees, evaluated once)
    ppi::FoldBuffers folds;

    // reference geometry (accelerator.cc:272-355) and the launch plan
    int dp_local = 1, dp_groups = 1, pdp_local = 1, best_local = 1;
    int form = PPI_FORM_FP, R = 4, F = 16, S = 8, chunk = 64, nrb = 1;

    bool launched = false;
    ppi_stats stats{};
    // device epilogue (ppi_engine_evaluate_ranked): alpha-inclusive vector, group winners, index
    float* dRankVec = nullptr;
    float* dRankGv = nullptr;
    int32_t* dRankGi = nullptr;
    int32_t* dRankIdx = nullptr;
    size_t cap_rank = 0;
};

namespace ppi {
// Kernel form and shape of one evaluation (engine.cc: compute_plan).
struct LaunchPlan {
    int form = PPI_FORM_FP, R = 4, F = 16, S = 9, chunk = 64, nrb = 1;
    int nrb_pass = 1;  // ordered form: row tiles per pass of stored row errors
};
// The form and shape for a compiled population of `nprog` programs (`cs` its compile statistics)
// on the engine's dataset and configuration, form `requested_form` (AUTO: FP).
bool compute_plan(const ppi_engine* E, const CompileStats& cs, int32_t nprog, int requested_form, LaunchPlan* L,
                  std::string* why);
// best_individual's work-group size for `nprog` programs (accelerator.cc:331-337).
int best_local_for(const ppi_engine* E, int32_t nprog);
// Kernel arguments naming the engine's own resident buffers.
EvalArgs make_args(ppi_engine* E, float alpha);
// Enqueues one evaluation (fills, form kernels, epilogue, recheck) on the engine stream with the
// buffers `A` names.
ppi_status enqueue_with(ppi_engine* E, EvalArgs A, const LaunchPlan& L, bool time_main);
}  // namespace ppi

#endif
