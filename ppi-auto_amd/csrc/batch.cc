// batch.cc — batched evaluation of many (population, dataset) pairs in one
// persistent launch (C ABI in include/ppi_engine.h, ppi_batch_*).
//
// Work items are (pair, row tile of 64*R rows, program range), where each pair's
// programs are cut into ranges of roughly equal tape length so that every item
// costs about the same; items are ordered longest first and pulled from an
// atomic counter by a grid sized to the kernel's occupancy.  Each pair runs in a
// kernel shape (rows per lane R, LDS feature slots F, register-stack slots S)
// chosen by the autotuner: pairs are keyed by their stack class, mean program
// length and row count, and each key's candidate shapes are timed on a sample of
// that key's own items.  Pairs that share a shape form one persistent launch; the
// launches run back to back.  The replay (queue resets, persistent kernels,
// per-program epilogue, download) is captured once as a hipGraph.  Populations are
// copied when added and compiled by build(), every pending one on its own thread
// of a pool.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <numeric>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/ppi_engine.h"
#include "engine_internal.h"
#include "kernels.h"
#include "tape.h"

using namespace ppi;

namespace {

#define B_HIP(call)                                                                            \
    do {                                                                                       \
        hipError_t e_ = (call);                                                                \
        if (e_ != hipSuccess) {                                                                \
            set_error(std::string(#call) + ": " + hipGetErrorString(e_));                      \
            return PPI_E_DEVICE;                                                               \
        }                                                                                      \
    } while (0)

double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// A staged dataset: column-major input columns, then the derived columns
// (unary operators over input columns, tape.h) its populations compiled to.
struct Dataset {
    int32_t nlin = 0, ncol = 0;
    int64_t ldx = 0;
    float* dX = nullptr;
    size_t cap_X = 0;  // floats allocated at dX
    std::unique_ptr<DerivedColumns> derived;
    int32_t derived_ready = 0;
    uint32_t* dDerOps = nullptr;
    int32_t* dDerSrc = nullptr;
    size_t cap_der = 0;
};

struct Pop {
    int32_t dataset = 0;
    int32_t n = 0;
    int32_t prog_base = 0;
    int64_t total_size = 0;
    int32_t max_slots = 0;   // register slots of the population's largest program (after compile)
    double mean_code = 0.0;  // tape instructions per program, its H_END included
    // The caller's arrays (copied by add_population) until build() compiles them.
    int32_t stride = 0;
    std::vector<int32_t> ph, sz;
    std::vector<float> eph;
    bool compiled = false;
    Tape t;
};

// A kernel shape of the persistent form: rows per lane, LDS feature slots, stack slots.
struct Shape {
    int R = 4, F = 16, S = 4;
    bool operator==(const Shape& o) const { return R == o.R && F == o.F && S == o.S; }
};

// The pairs that run in one shape: one persistent launch over their items.
struct Group {
    Shape shape;
    int grid = 0;
    int32_t item_off = 0, n_items = 0;
    double work = 0.0;  // instructions x rows of the group's items
};

constexpr int kMaxGroups = 16;  // queue counters allocated per batch

template <typename T>
ppi_status dalloc(T** p, size_t n) {
    if (*p) (void)hipFree(*p);
    *p = nullptr;
    hipError_t e = hipMalloc((void**)p, std::max<size_t>(n, 1) * sizeof(T));
    if (e != hipSuccess) {
        set_error(std::string("hipMalloc: ") + hipGetErrorString(e));
        *p = nullptr;
        return PPI_E_NOMEM;
    }
    return PPI_OK;
}

}  // namespace

struct ppi_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    FoldBuffers folds;  // fold programs of the population being added
    int dev_cu = 0, dev_max_wg = 0;
    size_t dev_lds = 0;

    bool configured = false;
    ppi_config cfg{};
    std::vector<float> cost;
    SymbolTable symtab;

    std::vector<Dataset> datasets;
    std::vector<Pop> pops;
    Tape tape;                       // all populations, concatenated

    // device state of the last build
    ProgMeta* dMeta = nullptr;
    Ins* dCode = nullptr;
    PairDev* dPairs = nullptr;
    WorkItem* dItems = nullptr;
    int32_t* dPairOf = nullptr;
    float* dPartial = nullptr;
    uint32_t* dBadLo = nullptr;
    uint32_t* dBadHi = nullptr;
    float* dRaw = nullptr;
    float* dVec = nullptr;
    float* dCost = nullptr;
    uint32_t* dCounter = nullptr;
    // the recheck's deep list (pairs of kDeepMinRows rows or more): kDeepCap entries of deep_stride rows
    uint32_t* dDeepCount = nullptr;
    int32_t* dDeepProg = nullptr;
    float* dDeepErr = nullptr;
    int64_t deep_stride = 0;
    int32_t deep_rows = 0;
    float* hRaw = nullptr;
    size_t hRawCap = 0;
    bool use_deep = false;  // this build queues deep-list programs (a dataset of kDeepMinRows rows or more)
    // Element capacity of each per-build device buffer above, by the buffer member's address: the
    // buffers are kept across builds and grown only when a build needs more.
    std::unordered_map<const void*, size_t> cap;

    std::vector<PairDev> pairs;
    std::vector<WorkItem> items;
    std::vector<Shape> pair_shape;  // per pair (population)
    std::vector<Group> groups;
    // Pairs evaluated by plain FP launches (the engine's kernel and chunking) instead of the queue:
    // those whose own launch fills the device (direct_min_items work items of a row tile x 64
    // programs or more); their arguments, built at build().
    int64_t direct_min_items = 0;  // 0: the default; < 0: none
    struct Direct { EvalArgs A; Shape shape; };
    std::vector<Direct> direct;
    // The autotuner's decisions by key (stack class, length class, rows class), kept across
    // builds until configure(): a rebuild with new populations times only keys it has not seen.
    std::vector<std::pair<int, Shape>> tuned;
    bool P = false, N = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool built = false, launched = false;
    ppi_batch_stats stats{};
};

namespace {

// The graph of the last build.  The per-build device buffers stay allocated for the next build
// (free_buffers releases them): freeing and re-allocating them every generation — the partials
// alone can be hundreds of MB — cost a device synchronization and fresh page mappings each time.
void free_build(ppi_batch* b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    b->deep_stride = 0; b->deep_rows = 0;
    b->use_deep = false;
    b->built = false;
}

void free_buffers(ppi_batch* b) {
    // (the queue counter lives as long as the batch object)
    void* ptrs[] = {b->dMeta, b->dCode, b->dPairs, b->dItems, b->dPairOf, b->dPartial, b->dBadLo, b->dBadHi,
                    b->dRaw, b->dVec, b->dDeepCount, b->dDeepProg, b->dDeepErr};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    b->dMeta = nullptr; b->dCode = nullptr; b->dPairs = nullptr; b->dItems = nullptr; b->dPairOf = nullptr;
    b->dPartial = nullptr; b->dBadLo = nullptr; b->dBadHi = nullptr; b->dRaw = nullptr; b->dVec = nullptr;
    b->dDeepCount = nullptr; b->dDeepProg = nullptr; b->dDeepErr = nullptr;
    b->cap.clear();
}

// A per-build buffer of at least n elements: kept if it already holds that many, else replaced.
template <typename T>
ppi_status dgrow(ppi_batch* b, T** p, size_t n) {
    size_t& c = b->cap[(const void*)p];
    if (*p && n <= c) return PPI_OK;
    ppi_status s = dalloc(p, n);
    c = s == PPI_OK ? std::max<size_t>(n, 1) : 0;
    return s;
}

// LDS feature slots of the persistent kernel: 16 whatever the pairs' widths.  Wider datasets read
// their columns 16 and up through the caches (LdsFeat), which at 6 waves per SIMD beats a 32-slot
// tile at the 4 waves its LDS allows (64-pair workload, 2 to 32 features: 13% faster).
int feature_slots(int max_col) {
    (void)max_col;
    return 16;
}

// Reference geometry of one pair (accelerator.cc:272-355), as the plain engine derives it.
void pair_geometry(const ppi_batch* b, int32_t nlin, int32_t nprog, int32_t* pdp_lanes, int32_t* dp_groups,
                   int32_t* dp_local_out, int32_t* best_local) {
    const int cu = b->cfg.compute_units > 0 ? b->cfg.compute_units : b->dev_cu;
    int mls1 = b->cfg.max_local_size > 0 ? b->cfg.max_local_size
                                          : (int)std::min<size_t>((size_t)b->dev_max_wg, b->dev_lds / 4);
    int mls2 = (int)std::min<size_t>((size_t)b->dev_max_wg, b->dev_lds / 8);
    int dp_local = std::max(1, std::min(mls1, (int)std::ceil(nlin / (float)cu)));
    *dp_groups = (nlin + dp_local - 1) / dp_local;
    *dp_local_out = dp_local;
    *pdp_lanes = std::max(1, nlin < mls1 ? nlin : mls1);
    *best_local = std::max(1, std::min(mls2, (int)std::ceil(nprog / (float)cu)));
}

// Work items for the pairs' shapes (pair_shape): each pair's programs are cut into
// ranges whose tape length (instructions + one per program for its epilogue) is about
// its group's `target`, and every range is crossed with every row tile of the pair.
// Pairs with include[k] == 0 get their partial block but no items, and with row_limit
// only the row tiles below row_limit[k] get items (the tuner's samples; ranges are then
// sized for that work).  Groups hold the pairs of one shape, largest first; a group's
// items are contiguous in `items`, longest first.
void plan_items(const ppi_batch* b, const std::vector<char>& include, double items_per_slot,
                std::vector<PairDev>* pairs, std::vector<WorkItem>* items, std::vector<Group>* groups,
                int64_t* partials, const std::vector<int64_t>* row_limit = nullptr) {
    pairs->clear();
    items->clear();
    groups->clear();
    const size_t np = b->pops.size();
    std::vector<int32_t> nrb(np), tiles(np);
    std::vector<int> gof(np, -1);
    for (size_t k = 0; k < np; ++k) {
        const Pop& p = b->pops[k];
        const Shape& sh = b->pair_shape[k];
        nrb[k] = (b->datasets[p.dataset].nlin + 64 * sh.R - 1) / (64 * sh.R);
        tiles[k] = row_limit ? (int32_t)std::min<int64_t>(nrb[k], ((*row_limit)[k] + 64 * sh.R - 1) / (64 * sh.R))
                             : nrb[k];
        if (!include[k]) continue;
        int g = 0;
        while (g < (int)groups->size() && !((*groups)[g].shape == sh)) ++g;
        if (g == (int)groups->size()) {
            Group G;
            G.shape = sh;
            G.grid = b->dev_cu * std::max(1, batch_occupancy(sh.R, sh.F, sh.S, b->P, b->N));
            groups->push_back(G);
        }
        gof[k] = g;
        (*groups)[g].work += p.mean_code * p.n * tiles[k] * 64.0 * sh.R;
    }
    int64_t part = 0;
    struct Cost { double c; WorkItem w; };
    std::vector<std::vector<Cost>> costed(groups->size());
    for (size_t k = 0; k < np; ++k) {
        const Pop& p = b->pops[k];
        const Dataset& d = b->datasets[p.dataset];
        PairDev pd{};
        pd.X = d.dX;
        pd.ldx = d.ldx;
        pd.nlin = d.nlin;
        pd.ncol = d.ncol;
        pd.prog_base = p.prog_base;
        pd.nprog = p.n;
        pd.partial_base = part;
        pd.nrb = nrb[k];
        pd.rows_per_tile = 64 * b->pair_shape[k].R;
        int32_t bl;
        pair_geometry(b, d.nlin, p.n, &pd.pdp_lanes, &pd.dp_groups, &pd.dp_local, &bl);
        pairs->push_back(pd);
        part += (int64_t)nrb[k] * p.n;
        if (gof[k] < 0) continue;
        const Group& G = (*groups)[gof[k]];
        // Program ranges of about `target` instructions each (an item covers one row tile).
        const double target = std::max(64.0, G.work / (64.0 * G.shape.R) / std::max(1.0, G.grid * items_per_slot));
        int32_t j0 = 0;
        double acc = 0.0;
        for (int32_t j = 0; j < p.n; ++j) {
            acc += b->tape.meta[p.prog_base + j].ncode + 1.0;
            if (acc >= target || j + 1 == p.n) {
                for (int32_t t = 0; t < tiles[k]; ++t)
                    costed[gof[k]].push_back({acc, WorkItem{(int32_t)k, t, j0, j + 1}});
                j0 = j + 1;
                acc = 0.0;
            }
        }
    }
    std::vector<int> order(groups->size());
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(),
                     [&](int x, int y) { return (*groups)[x].work > (*groups)[y].work; });
    std::vector<Group> sorted;
    for (int g : order) {
        auto& c = costed[g];
        std::stable_sort(c.begin(), c.end(), [](const Cost& x, const Cost& y) { return x.c > y.c; });
        Group G = (*groups)[g];
        G.item_off = (int32_t)items->size();
        G.n_items = (int32_t)c.size();
        for (auto& e : c) items->push_back(e.w);
        sorted.push_back(G);
    }
    *groups = sorted;
    *partials = part;
}

// Instructions x rows of a list of items (the tuner's unit of work).
double items_work(const ppi_batch* b, const WorkItem* w, size_t n) {
    double t = 0.0;
    for (size_t i = 0; i < n; ++i) {
        const Pop& p = b->pops[w[i].pair];
        double len = 0.0;
        for (int32_t j = w[i].p0; j < w[i].p1; ++j) len += b->tape.meta[p.prog_base + j].ncode + 1.0;
        t += len * 64.0 * b->pair_shape[w[i].pair].R;
    }
    return t;
}

// Compiles the populations added since the last build: derived columns planned
// per population (in the order they were added) and staged, then every pending
// population compiled on its own thread of a pool (the derived-column tables are
// only read there), then fold values, and the tapes appended to the batch tape in
// order.  A population's copied arrays are released once it is compiled.
ppi_status compile_pending(ppi_batch* b) {
    std::vector<size_t> pend;
    for (size_t k = 0; k < b->pops.size(); ++k)
        if (!b->pops[k].compiled) pend.push_back(k);
    if (pend.empty()) return PPI_OK;
    const bool P = b->cfg.protected_ops != 0, N = b->cfg.native_math != 0;
    const int hw = (int)std::min<unsigned>(16u, std::max(1u, std::thread::hardware_concurrency()));
    for (size_t k : pend) {
        Pop& p = b->pops[k];
        Dataset& d = b->datasets[p.dataset];
        plan_derived(b->symtab, p.ph.data(), p.eph.data(), p.sz.data(), p.stride, p.n, d.ncol, hw, d.derived.get());
    }
    for (auto& d : b->datasets) {
        if (d.derived_ready >= (int32_t)d.derived->cols.size()) continue;
        ppi_status s = stage_derived_columns(&d.dX, &d.cap_X, d.ldx, d.ncol, *d.derived, &d.derived_ready, &d.dDerOps,
                                             &d.dDerSrc, &d.cap_der, P, N, b->stream);
        if (s) return s;
    }
    const bool fold = b->cfg.no_constant_folding == 0;
    std::vector<std::string> errs(pend.size());
    std::vector<char> ok(pend.size(), 1);
    std::atomic<size_t> next{0};
    auto work = [&]() {
        for (size_t i = next++; i < pend.size(); i = next++) {
            Pop& p = b->pops[pend[i]];
            const Dataset& d = b->datasets[p.dataset];
            ok[i] = compile_population_planned(b->symtab, p.ph.data(), p.eph.data(), p.sz.data(), p.stride, p.n,
                                               d.ncol, &p.t, &errs[i], d.derived.get(), fold);
        }
    };
    const int workers = (int)std::min<size_t>((size_t)hw, pend.size());
    if (workers <= 1) {
        work();
    } else {
        std::vector<std::thread> pool;
        for (int w = 0; w < workers; ++w) pool.emplace_back(work);
        for (auto& th : pool) th.join();
    }
    for (size_t i = 0; i < pend.size(); ++i)
        if (!ok[i]) {
            set_error("population " + std::to_string(pend[i]) + ": " + errs[i]);
            return PPI_E_PROGRAM;
        }
    for (size_t k : pend) {
        Pop& p = b->pops[k];
        Tape& t = p.t;
        if (ppi_status fs = evaluate_folds(&t, &b->folds, P, N, b->stream)) return fs;
        if (t.stats.max_slots > kMaxSlots) {
            set_error("population " + std::to_string(k) + ": program needs more than 16 register slots");
            return PPI_E_PROGRAM;
        }
        p.max_slots = t.stats.max_slots;
        p.mean_code = (double)(t.stats.total_code + p.n) / std::max(1, p.n);
        const uint32_t off = (uint32_t)b->tape.code.size();
        for (auto m : t.meta) {
            m.code_off += off;
            b->tape.meta.push_back(m);
        }
        b->tape.code.insert(b->tape.code.end(), t.code.begin(), t.code.end());
        b->tape.stats.total_size += t.stats.total_size;
        b->tape.stats.total_code += t.stats.total_code;
        b->tape.stats.max_slots = std::max(b->tape.stats.max_slots, t.stats.max_slots);
        b->tape.stats.max_col = std::max(b->tape.stats.max_col, t.stats.max_col);
        p.compiled = true;
        p.t = Tape{};
        std::vector<int32_t>().swap(p.ph);
        std::vector<int32_t>().swap(p.sz);
        std::vector<float>().swap(p.eph);
    }
    return PPI_OK;
}

ppi_status upload_items(ppi_batch* b) {
    ppi_status s;
    if ((s = dgrow(b, &b->dPairs, b->pairs.size()))) return s;
    if ((s = dgrow(b, &b->dItems, b->items.size()))) return s;
    B_HIP(hipMemcpy(b->dPairs, b->pairs.data(), b->pairs.size() * sizeof(PairDev), hipMemcpyHostToDevice));
    B_HIP(hipMemcpy(b->dItems, b->items.data(), b->items.size() * sizeof(WorkItem), hipMemcpyHostToDevice));
    return PPI_OK;
}

BatchArgs make_args(ppi_batch* b, int32_t n_items) {
    BatchArgs B{};
    EvalArgs& A = B.base;
    A.meta = b->dMeta;
    A.code = b->dCode;
    A.chunk = 0;
    A.strategy = b->cfg.strategy;
    A.error_fn = b->cfg.error_fn;
    A.reducemax = b->cfg.reducemax;
    A.cost = b->dCost;
    A.cost_n = b->cfg.cost_n;
    A.alpha = 0.0f;
    A.partial = b->dPartial;
    A.bad_lo = b->dBadLo;
    A.bad_hi = b->dBadHi;
    A.raw = b->dRaw;
    A.vec = b->dVec;
    if (b->use_deep) {
        A.deep_count = b->dDeepCount;
        A.deep_prog = b->dDeepProg;
        A.deep_cap = kDeepCap;
    }
    B.pairs = b->dPairs;
    B.items = b->dItems;
    B.n_items = n_items;
    B.counter = b->dCounter;
    return B;
}

int32_t total_programs(const ppi_batch* b) {
    return b->pops.empty() ? 0 : b->pops.back().prog_base + b->pops.back().n;
}

// The recheck and its deep list.
ppi_status enqueue_recheck(ppi_batch* b, const BatchArgs& B, int32_t n) {
    if (B.base.deep_cap > 0) B_HIP(launch_fill_u32(b->dDeepCount, 0u, 1, b->stream));
    B_HIP(launch_batch_recheck(B, b->dPairOf, n, b->P, b->N, b->stream));
    if (B.base.deep_cap > 0) {
        DeepArgs D{};
        D.base = B.base;
        D.pairs = b->dPairs;
        D.pair_of = b->dPairOf;
        D.err = b->dDeepErr;
        D.stride = b->deep_stride;
        D.max_rows = b->deep_rows;
        B_HIP(launch_deep(D, b->dev_cu * 4, b->P, b->N, b->stream));
    }
    return PPI_OK;
}

// The direct pairs' FP launches (with_direct), then the persistent launches of the groups, back
// to back, each with its own queue counter (reset beforehand by the caller).
ppi_status enqueue_groups(ppi_batch* b, const std::vector<Group>& groups, bool with_direct = true) {
    if (with_direct)
        for (const auto& d : b->direct)
            B_HIP(launch_fp(d.A, d.shape.R, d.shape.F, d.shape.S, b->P, b->N, b->stream));
    for (size_t g = 0; g < groups.size(); ++g) {
        const Group& G = groups[g];
        if (G.n_items == 0) continue;
        BatchArgs B = make_args(b, G.n_items);
        B.items = b->dItems + G.item_off;
        B.counter = b->dCounter + g;
        B_HIP(launch_batch(B, G.grid, G.shape.R, G.shape.F, G.shape.S, b->P, b->N, b->stream));
    }
    return PPI_OK;
}

// Queue resets, persistent kernels, epilogue (optionally the download).
ppi_status enqueue_run(ppi_batch* b, const std::vector<Group>& groups, bool with_finalize, bool with_download,
                       bool with_direct = true) {
    const int32_t n = total_programs(b);
    BatchArgs B = make_args(b, 0);
    B_HIP(launch_fill_u32(b->dBadLo, 0xffffffffu, n, b->stream));
    B_HIP(launch_fill_u32(b->dBadHi, 0u, n, b->stream));
    B_HIP(launch_fill_u32(b->dCounter, 0u, (int32_t)groups.size(), b->stream));
    ppi_status gs = enqueue_groups(b, groups, with_direct);
    if (gs) return gs;
    if (with_finalize) {
        B_HIP(launch_batch_finalize(B, b->dPairOf, n, b->stream));
        ppi_status s = enqueue_recheck(b, B, n);
        if (s) return s;
    }
    if (with_download)
        B_HIP(hipMemcpyAsync(b->hRaw, b->dRaw, (size_t)n * sizeof(float), hipMemcpyDeviceToHost, b->stream));
    return PPI_OK;
}

}  // namespace

extern "C" {

ppi_status ppi_batch_create(ppi_batch** out, int32_t device) {
    if (!out) { set_error("out is NULL"); return PPI_E_ARG; }
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) { set_error("no HIP device"); return PPI_E_NODEVICE; }
    if (device < 0 || device >= n) { set_error("bad device index"); return PPI_E_ARG; }
    ppi_batch* b = new ppi_batch();
    b->device = device;
    hipDeviceProp_t prop;
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = hipGetDeviceProperties(&prop, device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    for (int k = 0; k < 4 && e == hipSuccess; ++k) e = hipEventCreate(&b->ev[k]);
    if (e == hipSuccess) e = hipMalloc((void**)&b->dCounter, kMaxGroups * sizeof(uint32_t));
    if (e != hipSuccess) {
        set_error(std::string("device init: ") + hipGetErrorString(e));
        ppi_batch_destroy(b);
        return PPI_E_DEVICE;
    }
    b->dev_cu = prop.multiProcessorCount;
    b->dev_max_wg = prop.maxThreadsPerBlock;
    b->dev_lds = prop.sharedMemPerBlock;
    b->cfg.strategy = PPI_STRATEGY_SEQ;
    b->cfg.error_fn = PPI_EF_SQUARED;
    *out = b;
    return PPI_OK;
}

void ppi_batch_destroy(ppi_batch* b) {
    if (!b) return;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    free_build(b);
    free_buffers(b);
    if (b->dCounter) (void)hipFree(b->dCounter);
    for (auto& d : b->datasets) {
        if (d.dX) (void)hipFree(d.dX);
        if (d.dDerOps) (void)hipFree(d.dDerOps);
        if (d.dDerSrc) (void)hipFree(d.dDerSrc);
    }
    if (b->dCost) (void)hipFree(b->dCost);
    b->folds.release();
    if (b->hRaw) (void)hipHostFree(b->hRaw);
    for (auto& ev : b->ev)
        if (ev) (void)hipEventDestroy(ev);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

ppi_status ppi_batch_configure(ppi_batch* b, const ppi_config* cfg, const ppi_symbol* table, int32_t ntable) {
    if (!b || !cfg || ntable < 0 || (ntable > 0 && !table)) { set_error("bad arguments"); return PPI_E_ARG; }
    if (cfg->strategy < PPI_STRATEGY_SEQ || cfg->strategy > PPI_STRATEGY_PDP) { set_error("bad strategy"); return PPI_E_ARG; }
    if (cfg->error_fn == PPI_EF_EXPR) {
        set_error("the batched evaluator takes the built-in error functions (a user ERROR(X,Y) runs through the engine)");
        return PPI_E_UNSUPPORTED;
    }
    if (cfg->error_fn < PPI_EF_ABS || cfg->error_fn > PPI_EF_COST_CLAMP) { set_error("bad error_fn"); return PPI_E_ARG; }
    const bool cost = cfg->error_fn == PPI_EF_COST || cfg->error_fn == PPI_EF_COST_CLAMP;
    if (cost && (cfg->cost_n <= 0 || !cfg->cost)) { set_error("COST error function needs a matrix"); return PPI_E_ARG; }
    if (cfg->max_local_size < 0 || cfg->max_local_size > kPdpMaxLanes) {
        set_error("max_local_size must be in [0, " + std::to_string(kPdpMaxLanes) + "]");
        return PPI_E_ARG;
    }
    if (cfg->max_local_size == 0 && std::min<size_t>((size_t)b->dev_max_wg, b->dev_lds / 4) > (size_t)kPdpMaxLanes) {
        set_error("this device's work-group limit exceeds " + std::to_string(kPdpMaxLanes) + "; set max_local_size");
        return PPI_E_ARG;
    }
    if (cfg->rows_per_lane != 0 && cfg->rows_per_lane != 2 && cfg->rows_per_lane != 4 && cfg->rows_per_lane != 8) {
        set_error("batched rows_per_lane must be 0 (autotune), 2, 4 or 8");
        return PPI_E_ARG;
    }
    (void)hipSetDevice(b->device);
    b->cfg = *cfg;
    b->cost.assign(cfg->cost, cfg->cost + (cost ? (size_t)cfg->cost_n * cfg->cost_n : 0));
    b->cfg.cost = nullptr;
    if (cost) {
        ppi_status s = dalloc(&b->dCost, b->cost.size());
        if (s) return s;
        B_HIP(hipMemcpy(b->dCost, b->cost.data(), b->cost.size() * sizeof(float), hipMemcpyHostToDevice));
    }
    b->symtab.build(table, ntable);
    b->configured = true;
    b->tuned.clear();
    // Derived columns depend on PROTECTED / NATIVE: recompute them on demand.
    for (auto& d : b->datasets) {
        d.derived->reset(d.ncol, derived_capacity(d.ldx, d.ncol));
        d.derived_ready = 0;
    }
    b->pops.clear();
    b->tape = Tape{};
    free_build(b);
    return PPI_OK;
}

ppi_status ppi_batch_add_dataset(ppi_batch* b, const float* data, int32_t nlin, int32_t ncol, int32_t* id) {
    if (!b || !data || !id || nlin <= 0 || ncol <= 0) { set_error("bad arguments"); return PPI_E_ARG; }
    (void)hipSetDevice(b->device);
    Dataset d;
    d.nlin = nlin;
    d.ncol = ncol;
    d.ldx = ((int64_t)nlin + 2047) / 2048 * 2048;
    std::vector<float> cm((size_t)d.ldx * ncol, 0.0f);
    for (int32_t c = 0; c < ncol; ++c)
        for (int32_t r = 0; r < nlin; ++r) cm[(size_t)c * d.ldx + r] = data[(size_t)r * ncol + c];
    ppi_status s = dalloc(&d.dX, cm.size());
    if (s) return s;
    d.cap_X = cm.size();
    d.derived.reset(new DerivedColumns());
    d.derived->reset(ncol, derived_capacity(d.ldx, ncol));
    d.derived->refill = false;  // populations compiled by earlier builds keep reading the columns
    hipError_t e = hipMemcpy(d.dX, cm.data(), cm.size() * sizeof(float), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(d.dX);
        set_error(std::string("dataset upload: ") + hipGetErrorString(e));
        return PPI_E_DEVICE;
    }
    b->datasets.push_back(std::move(d));
    *id = (int32_t)b->datasets.size() - 1;
    return PPI_OK;
}

ppi_status ppi_batch_add_population(ppi_batch* b, int32_t dataset_id, const int32_t* phenotype, const float* ephemeral,
                                    const int32_t* size, int32_t stride, int32_t n, int32_t* pop_id) {
    if (!b || !b->configured) { set_error("configure the batch first"); return PPI_E_STATE; }
    if (dataset_id < 0 || dataset_id >= (int32_t)b->datasets.size()) { set_error("unknown dataset"); return PPI_E_ARG; }
    if (!phenotype || !ephemeral || !size || stride <= 0 || n <= 0 || !pop_id) { set_error("bad program arrays"); return PPI_E_ARG; }
    Pop p;
    p.dataset = dataset_id;
    p.n = n;
    p.prog_base = total_programs(b);
    p.stride = stride;
    for (int32_t i = 0; i < n; ++i) {
        if (size[i] < 0 || size[i] > stride) {
            set_error("program " + std::to_string(i) + ": size " + std::to_string(size[i]) + " outside [0, " +
                      std::to_string(stride) + "]");
            return PPI_E_PROGRAM;
        }
        p.total_size += size[i];
    }
    // Compiled by build(), all pending populations at once on a thread pool.
    p.ph.assign(phenotype, phenotype + (size_t)n * stride);
    p.eph.assign(ephemeral, ephemeral + (size_t)n * stride);
    p.sz.assign(size, size + n);
    b->pops.push_back(std::move(p));
    b->built = false;
    *pop_id = (int32_t)b->pops.size() - 1;
    return PPI_OK;
}

ppi_status ppi_batch_clear_populations(ppi_batch* b) {
    if (!b) { set_error("batch is NULL"); return PPI_E_ARG; }
    b->pops.clear();
    b->tape = Tape{};
    free_build(b);
    return PPI_OK;
}

ppi_status ppi_batch_build(ppi_batch* b) {
    if (!b || !b->configured) { set_error("configure the batch first"); return PPI_E_STATE; }
    if (b->pops.empty()) { set_error("add a population first"); return PPI_E_STATE; }
    (void)hipSetDevice(b->device);
    const double t0 = now_ms();
    free_build(b);
    ppi_status s;
    if ((s = compile_pending(b))) return s;
    const int32_t n = total_programs(b);
    b->P = b->cfg.protected_ops != 0;
    b->N = b->cfg.native_math != 0;
    const size_t np = b->pops.size();

    std::vector<Ins> code = b->tape.code;
    code.insert(code.end(), kTapePad, Ins{H_END, kEmpty});
    if ((s = dgrow(b, &b->dMeta, b->tape.meta.size()))) return s;
    if ((s = dgrow(b, &b->dCode, code.size()))) return s;
    B_HIP(hipMemcpy(b->dMeta, b->tape.meta.data(), b->tape.meta.size() * sizeof(ProgMeta), hipMemcpyHostToDevice));
    B_HIP(hipMemcpy(b->dCode, code.data(), code.size() * sizeof(Ins), hipMemcpyHostToDevice));
    std::vector<int32_t> pair_of((size_t)n);
    for (size_t k = 0; k < np; ++k)
        std::fill(pair_of.begin() + b->pops[k].prog_base, pair_of.begin() + b->pops[k].prog_base + b->pops[k].n,
                  (int32_t)k);
    if ((s = dgrow(b, &b->dPairOf, (size_t)n))) return s;
    B_HIP(hipMemcpy(b->dPairOf, pair_of.data(), pair_of.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    if ((s = dgrow(b, &b->dBadLo, (size_t)n))) return s;
    if ((s = dgrow(b, &b->dBadHi, (size_t)n))) return s;
    if ((s = dgrow(b, &b->dRaw, (size_t)n))) return s;
    if ((s = dgrow(b, &b->dVec, (size_t)n))) return s;
    if (b->hRawCap < (size_t)n) {
        if (b->hRaw) (void)hipHostFree(b->hRaw);
        b->hRaw = nullptr;
        B_HIP(hipHostMalloc((void**)&b->hRaw, (size_t)n * sizeof(float), hipHostMallocDefault));
        b->hRawCap = (size_t)n;
    }

    // Candidate shapes of each pair, from its largest program's register slots: the 4-slot
    // stack at 4 and 8 rows per lane (its 2-row form is the 9-slot one), the 9-slot stack at
    // 2, 4 and 8, the 16-slot stack at 2 rows; 16 LDS feature slots (feature_slots), 32 at
    // 2 rows.  The shape at 4 rows per lane comes first (the default where nothing is timed).
    std::vector<std::vector<Shape>> cand(np);
    for (size_t k = 0; k < np; ++k) {
        const int ms = b->pops[k].max_slots;
        const int F = feature_slots(b->tape.stats.max_col);
        std::vector<Shape> c;
        if (ms <= 4) c = {{4, F, 4}, {8, F, 4}, {2, 32, 9}};
        else if (ms <= 9) c = {{4, F, 9}, {8, F, 9}, {2, 32, 9}};
        else c = {{2, 32, 16}};
        c.erase(std::remove_if(c.begin(), c.end(),
                               [&](const Shape& x) {
                                   return !batch_shape_supported(x.R, x.F, x.S) ||
                                          (b->cfg.rows_per_lane > 0 && x.R != b->cfg.rows_per_lane);
                               }),
                c.end());
        if (c.empty()) {
            set_error("population " + std::to_string(k) + ": no batched kernel shape for its programs (" +
                      std::to_string(ms) + " register slots) at rows_per_lane " +
                      std::to_string(b->cfg.rows_per_lane));
            return PPI_E_PROGRAM;
        }
        cand[k] = c;
    }
    // Partials sized for each pair's smallest candidate R.
    int64_t max_part = 0;
    for (size_t k = 0; k < np; ++k) {
        int rmin = 8;
        for (const Shape& x : cand[k]) rmin = std::min(rmin, x.R);
        const Dataset& d = b->datasets[b->pops[k].dataset];
        max_part += (int64_t)((d.nlin + 64 * rmin - 1) / (64 * rmin)) * b->pops[k].n;
    }
    if ((s = dgrow(b, &b->dPartial, (size_t)max_part))) return s;
    b->deep_rows = 0;
    b->deep_stride = 0;
    for (const Pop& p : b->pops) {
        const Dataset& d = b->datasets[p.dataset];
        b->deep_rows = std::max(b->deep_rows, d.nlin);
        b->deep_stride = std::max(b->deep_stride, d.ldx);
    }
    if (b->deep_rows >= kDeepMinRows) {
        if ((s = dgrow(b, &b->dDeepCount, 1))) return s;
        if ((s = dgrow(b, &b->dDeepProg, (size_t)kDeepCap))) return s;
        if ((s = dgrow(b, &b->dDeepErr, (size_t)kDeepCap * b->deep_stride))) return s;
        b->use_deep = true;
    }

    // Autotuner.  Pairs are keyed by (stack class, mean tape length per program, rows):
    // lengths below 24, 24-63 and 64 or more instructions; rows below 16,384, below 262,144,
    // and more.  For each key with more than one candidate shape, every candidate is timed
    // on the same sample of that key's work (a row prefix of each of its pairs, all their
    // programs), and the key's pairs take the fastest.  A key decided by an earlier build of
    // this batch keeps its shape.
    const double items_per_slot = 6.0;
    b->pair_shape.resize(np);
    for (size_t k = 0; k < np; ++k) b->pair_shape[k] = cand[k][0];
    std::vector<std::pair<int, std::vector<size_t>>> cells;
    for (size_t k = 0; k < np; ++k) {
        const Pop& p = b->pops[k];
        const int sc = cand[k][0].S;
        const int lc = p.mean_code < 24.0 ? 0 : (p.mean_code < 64.0 ? 1 : 2);
        const int32_t nl = b->datasets[p.dataset].nlin;
        const int rc = nl < 16384 ? 0 : (nl < 262144 ? 1 : 2);
        const int key = (sc * 3 + lc) * 3 + rc;
        auto it = std::find_if(cells.begin(), cells.end(), [&](const auto& c) { return c.first == key; });
        if (it == cells.end()) cells.push_back({key, {k}});
        else it->second.push_back(k);
    }
    std::memset(b->stats.tune_ms, 0, sizeof(b->stats.tune_ms));
    b->stats.autotuned = 0;
    b->stats.tuned_keys = 0;
    double top_work = -1.0;
    int64_t part;
    for (auto& cell : cells) {
        const std::vector<size_t>& ks = cell.second;
        const std::vector<Shape>& cs = cand[ks[0]];
        if (cs.size() < 2) continue;
        auto known = std::find_if(b->tuned.begin(), b->tuned.end(), [&](const auto& t) { return t.first == cell.first; });
        if (known != b->tuned.end()) {
            for (size_t k : ks) b->pair_shape[k] = known->second;
            b->stats.autotuned = 1;
            continue;
        }
        std::vector<char> inc(np, 0);
        double cell_work = 0.0;
        for (size_t k : ks) {
            inc[k] = 1;
            const Pop& p = b->pops[k];
            cell_work += p.mean_code * p.n * b->datasets[p.dataset].nlin;
        }
        // The same work for every candidate: rows [0, T_k) of each of the key's pairs, all their
        // programs, T_k a multiple of 512 rows (whole row tiles at 2, 4 and 8 rows per lane), the
        // prefix sized to about 8e9 instruction x rows (a few ms) so that the timed span runs
        // many items per workgroup, not just the ramp and the tail.
        const double frac = std::min(1.0, 8e9 / std::max(1.0, cell_work));
        std::vector<int64_t> rows_of(np, 0);
        for (size_t k : ks) {
            const int32_t nl = b->datasets[b->pops[k].dataset].nlin;
            rows_of[k] = std::min<int64_t>(nl, ((int64_t)std::ceil(frac * nl / 512.0)) * 512);
        }
        double tms[3] = {0.0, 0.0, 0.0}, tms_raw[3] = {0.0, 0.0, 0.0};
        for (const Shape& sh : cs) {
            for (size_t k : ks) b->pair_shape[k] = sh;
            plan_items(b, inc, items_per_slot, &b->pairs, &b->items, &b->groups, &part, &rows_of);
            const double work = items_work(b, b->items.data(), b->items.size());
            if ((s = upload_items(b))) return s;
            if ((s = enqueue_run(b, b->groups, false, false, false))) return s;  // warm-up
            // Repeat short samples so that the timed span is well above launch jitter.
            int reps = 1;
            float ms = 0.0f;
            for (;;) {
                B_HIP(hipEventRecord(b->ev[2], b->stream));
                for (int r = 0; r < reps; ++r)
                    if ((s = enqueue_run(b, b->groups, false, false, false))) return s;
                B_HIP(hipEventRecord(b->ev[3], b->stream));
                B_HIP(hipEventSynchronize(b->ev[3]));
                B_HIP(hipEventElapsedTime(&ms, b->ev[2], b->ev[3]));
                if (ms >= 2.0f || reps >= 64) break;
                reps *= 4;
            }
            const int slot = sh.R == 2 ? 0 : (sh.R == 4 ? 1 : 2);
            tms_raw[slot] = ms / reps;
            tms[slot] = ms / reps / std::max(1.0, work) * 1e9;
        }
        // The first candidate (4 rows per lane) stays unless another is faster by 3% on the
        // same work: the samples cannot resolve smaller differences reliably.
        Shape bs = cs[0];
        double best = tms_raw[cs[0].R == 2 ? 0 : (cs[0].R == 4 ? 1 : 2)];
        for (const Shape& sh : cs) {
            const double t = tms_raw[sh.R == 2 ? 0 : (sh.R == 4 ? 1 : 2)];
            if (t < 0.97 * best) { best = t; bs = sh; }
        }
        for (size_t k : ks) b->pair_shape[k] = bs;
        b->tuned.push_back({cell.first, bs});
        b->stats.autotuned = 1;
        ++b->stats.tuned_keys;
        if (cell_work > top_work) {
            top_work = cell_work;
            for (int i = 0; i < 3; ++i) b->stats.tune_ms[i] = tms[i];
        }
    }
    // Pairs whose own FP launch fills the device run as one: the engine's kernel, chunking and
    // XCD-aware tile order over the pair alone, with the partials in the pair's block (the layout
    // the queue's items write), so the epilogue below is the same for both.  The queue keeps the
    // rest, where its packing of many pairs into one launch pays.  On the 64-pair workload the
    // pairs' FP launches run at the plain engine's rate (7 waves per SIMD against the persistent
    // kernel's 6).
    std::vector<char> queued(np, 1);
    b->direct.clear();
    const int64_t min_items = b->direct_min_items != 0 ? b->direct_min_items : (int64_t)b->dev_cu * 96 * 64;
    if (min_items > 0) {
        for (size_t k = 0; k < np; ++k) {
            const Shape& sh = b->pair_shape[k];
            const Pop& p = b->pops[k];
            const Dataset& d = b->datasets[p.dataset];
            const int64_t nrb = (d.nlin + 64 * sh.R - 1) / (64 * sh.R);
            if ((int64_t)p.n * nrb < min_items || !shape_supported(sh.R, sh.F, sh.S)) continue;
            queued[k] = 0;
        }
    }
    plan_items(b, queued, items_per_slot, &b->pairs, &b->items, &b->groups, &part);
    for (size_t k = 0; k < np; ++k) {
        if (queued[k]) continue;
        const PairDev& pd = b->pairs[k];
        EvalArgs A = make_args(b, 0).base;  // the per-pair fields as batch_kernel sets them
        A.meta += pd.prog_base;
        A.X = pd.X;
        A.ldx = pd.ldx;
        A.nlin = pd.nlin;
        A.ncol = pd.ncol;
        A.ycol = pd.ncol - 1;
        A.nprog = pd.nprog;
        A.pdp_lanes = pd.pdp_lanes;
        A.dp_groups = pd.dp_groups;
        A.partial += pd.partial_base;
        A.bad_lo += pd.prog_base;
        A.bad_hi += pd.prog_base;
        A.chunk = fp_default_chunk(pd.nprog, pd.nrb, b->dev_cu);
        A.rb0 = 0;
        A.nrb_pass = pd.nrb;
        b->direct.push_back({A, b->pair_shape[k]});
    }
    if ((s = upload_items(b))) return s;
    if ((int)b->groups.size() > kMaxGroups) { set_error("too many kernel shapes"); return PPI_E_PROGRAM; }

    // Capture the replay: resets, persistent kernel, epilogue, download.
    B_HIP(hipStreamSynchronize(b->stream));
    B_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    s = enqueue_run(b, b->groups, true, true);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(b->stream, &g);
    if (s) return s;
    if (e != hipSuccess) { set_error(std::string("capture: ") + hipGetErrorString(e)); return PPI_E_DEVICE; }
    b->graph = g;
    B_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    b->built = true;
    b->stats.n_pairs = (int32_t)b->pops.size();
    b->stats.n_items = (int32_t)b->items.size();
    b->stats.direct_pairs = (int32_t)b->direct.size();
    // The group carrying the most work (groups are sorted by work), else the first direct pair.
    const Shape top = b->groups.empty() ? b->direct[0].shape : b->groups[0].shape;
    b->stats.grid = b->groups.empty() ? 0 : b->groups[0].grid;
    b->stats.rows_per_lane = top.R;
    b->stats.feature_slots = top.F;
    b->stats.n_shapes = (int32_t)b->groups.size();
    double w = 0.0;
    for (auto& p : b->pops) w += (double)p.total_size * b->datasets[p.dataset].nlin;
    b->stats.node_evals = w;
    b->stats.derived_columns = 0;
    for (auto& d : b->datasets) b->stats.derived_columns += d.derived_ready;
    b->stats.build_ms = now_ms() - t0;
    return PPI_OK;
}

ppi_status ppi_batch_launch(ppi_batch* b) {
    if (!b || !b->built) { set_error("build the batch first"); return PPI_E_STATE; }
    (void)hipSetDevice(b->device);
    B_HIP(hipEventRecord(b->ev[0], b->stream));
    B_HIP(hipGraphLaunch(b->exec, b->stream));
    B_HIP(hipEventRecord(b->ev[1], b->stream));
    b->launched = true;
    return PPI_OK;
}

ppi_status ppi_batch_wait(ppi_batch* b) {
    if (!b || !b->launched) { set_error("nothing launched"); return PPI_E_STATE; }
    b->launched = false;
    B_HIP(hipEventSynchronize(b->ev[1]));
    float ms = 0.0f;
    B_HIP(hipEventElapsedTime(&ms, b->ev[0], b->ev[1]));
    b->stats.graph_ms = ms;
    return PPI_OK;
}

ppi_status ppi_batch_results(ppi_batch* b, int32_t pop_id, float alpha, float* vector, int32_t* index,
                             int32_t* best_size) {
    if (!b || !b->built) { set_error("build the batch first"); return PPI_E_STATE; }
    if (pop_id < 0 || pop_id >= (int32_t)b->pops.size()) { set_error("unknown population"); return PPI_E_ARG; }
    const Pop& p = b->pops[pop_id];
    std::vector<float> vec((size_t)p.n);
    for (int32_t i = 0; i < p.n; ++i) {
        const float raw = b->hRaw[p.prog_base + i];
        const ProgMeta& m = b->tape.meta[p.prog_base + i];
        const float as = alpha * (float)m.size;
        if (b->cfg.strategy == PPI_STRATEGY_PP || b->cfg.strategy == PPI_STRATEGY_PDP) vec[i] = raw + as;
        else vec[i] = (raw == FLT_MAX || (m.flags & kEmpty)) ? raw : raw + as;
    }
    if (vector) std::memcpy(vector, vec.data(), vec.size() * sizeof(float));
    if (index && best_size && *best_size > 0) {
        int32_t pdp, dpg, dpl, bl;
        pair_geometry(b, b->datasets[p.dataset].nlin, p.n, &pdp, &dpg, &dpl, &bl);
        if (b->cfg.strategy == PPI_STRATEGY_SEQ) {
            if (*best_size > p.n) *best_size = p.n;
            pick_n_best(*best_size, index, p.n, vec.data(), nullptr);
        } else if (b->cfg.strategy == PPI_STRATEGY_FP) {
            best_acc(vec.data(), p.n, bl, index, best_size);
        } else {
            best_acc(b->hRaw + p.prog_base, p.n, bl, index, best_size);
        }
    }
    return PPI_OK;
}

ppi_status ppi_batch_time(ppi_batch* b, int32_t iters, double* ms_per_run) {
    if (!b || !b->built || iters <= 0 || !ms_per_run) { set_error("bad arguments"); return PPI_E_ARG; }
    (void)hipSetDevice(b->device);
    B_HIP(hipEventRecord(b->ev[0], b->stream));
    for (int32_t k = 0; k < iters; ++k) B_HIP(hipGraphLaunch(b->exec, b->stream));
    B_HIP(hipEventRecord(b->ev[1], b->stream));
    B_HIP(hipEventSynchronize(b->ev[1]));
    float ms = 0.0f;
    B_HIP(hipEventElapsedTime(&ms, b->ev[0], b->ev[1]));
    *ms_per_run = ms / iters;
    b->stats.graph_ms = *ms_per_run;
    // The evaluation kernels alone (direct FP launches and the persistent queue), outside the graph, for the
    // roofline figure.
    double kms = 0.0;
    for (int32_t k = 0; k < iters; ++k) {
        const int32_t n = total_programs(b);
        BatchArgs B = make_args(b, 0);
        B_HIP(launch_fill_u32(b->dBadLo, 0xffffffffu, n, b->stream));
        B_HIP(launch_fill_u32(b->dBadHi, 0u, n, b->stream));
        B_HIP(launch_fill_u32(b->dCounter, 0u, (int32_t)b->groups.size(), b->stream));
        B_HIP(hipEventRecord(b->ev[2], b->stream));
        ppi_status gs = enqueue_groups(b, b->groups);
        if (gs) return gs;
        B_HIP(hipEventRecord(b->ev[3], b->stream));
        B_HIP(launch_batch_finalize(B, b->dPairOf, n, b->stream));
        ppi_status s = enqueue_recheck(b, B, n);
        if (s) return s;
        B_HIP(hipEventSynchronize(b->ev[3]));
        float t = 0.0f;
        B_HIP(hipEventElapsedTime(&t, b->ev[2], b->ev[3]));
        kms += t;
    }
    B_HIP(hipStreamSynchronize(b->stream));
    b->stats.kernel_ms = kms / iters;
    return PPI_OK;
}

ppi_status ppi_batch_pair_shape(const ppi_batch* b, int32_t pop_id, int32_t* rows_per_lane, int32_t* feature_slots,
                                int32_t* stack_slots) {
    if (!b || !b->built) { set_error("build the batch first"); return PPI_E_STATE; }
    if (pop_id < 0 || pop_id >= (int32_t)b->pops.size()) { set_error("unknown population"); return PPI_E_ARG; }
    const Shape& sh = b->pair_shape[pop_id];
    if (rows_per_lane) *rows_per_lane = sh.R;
    if (feature_slots) *feature_slots = sh.F;
    if (stack_slots) *stack_slots = sh.S;
    return PPI_OK;
}

ppi_status ppi_batch_set_direct_threshold(ppi_batch* b, int64_t min_items) {
    if (!b) { set_error("batch is NULL"); return PPI_E_ARG; }
    b->direct_min_items = min_items;
    return PPI_OK;
}

ppi_status ppi_batch_get_stats(const ppi_batch* b, ppi_batch_stats* out) {
    if (!b || !out) { set_error("bad arguments"); return PPI_E_ARG; }
    *out = b->stats;
    return PPI_OK;
}

}  // extern "C"
