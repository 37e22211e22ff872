// adjoint.cc — reverse-mode derivatives of the fitness with respect to the
// ephemeral constants, and on-device constant refinement (C ABI in
// include/ppi_engine.h, ppi_adjoint_*).
//
// forward() evaluates with an internal engine (the plain drop-in path) and
// keeps the programs.  backward() compiles them into the adjoint tape (forward
// tape + per-segment store lists + reverse program, tape.h), then launches a
// persistent grid whose waves take (job, row range) units from a queue counter,
// heaviest job first within each range; the ranges are cut so that the units
// outnumber the device's wave slots 32 times over.  The partial store (VGPRs with
// forward recomputation per segment, or LDS with one forward pass for most
// programs), the rows per lane (1, or 2 for programs of at most 5 stack and
// adjoint slots) and, at one row per lane, whether a row's features sit in
// registers are picked by timing the candidates on a sample of the units.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/ppi_engine.h"
#include "engine_internal.h"
#include "kernels.h"
#include "tape.h"

using namespace ppi;

namespace {

#define A_HIP(call)                                                                            \
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

// Device buffer that grows on demand.
template <typename T>
struct DBuf {
    T* p = nullptr;
    size_t cap = 0;
    ppi_status reserve(size_t n) {
        if (n <= cap && p) return PPI_OK;
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
        size_t want = std::max<size_t>(n, 1);
        if (hipMalloc((void**)&p, want * sizeof(T)) != hipSuccess) {
            p = nullptr;
            set_error("device allocation of " + std::to_string(want * sizeof(T)) + " bytes failed");
            return PPI_E_NOMEM;
        }
        cap = want;
        return PPI_OK;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
};

template <typename T>
ppi_status upload(DBuf<T>& b, const std::vector<T>& v, hipStream_t st) {
    ppi_status s = b.reserve(v.size());
    if (s) return s;
    if (!v.empty()) A_HIP(hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, st));
    return PPI_OK;
}

}  // namespace

struct ppi_adjoint {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    int dev_cu = 0;
    ppi_engine* eng = nullptr;
    bool configured = false;
    ppi_config cfg{};

    // programs kept by forward()
    std::vector<int32_t> ph, sz;
    std::vector<float> eph;
    int32_t stride = 0, n = 0;
    float alpha = 0.0f;
    bool have_programs = false;

    // compiled adjoint tape for partial store `tape_ps` at `tape_r` rows per lane (-1: none)
    int tape_ps = -1, tape_r = 0;
    // rows per lane requested (0: chosen by measurement); sampled ms per candidate
    // ((rows per lane - 1) * 2 + store) * 2 + (features read from the dataset),
    // -1 = not sampled, -2 = no kernel shape for these programs
    int want_r = 0;
    double cand_ms[8] = {-1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0};
    AdjTape T;
    int R = 1, F = 12, S = 12, AS = kAdjSlots;
    int32_t nranges = 1;
    int64_t range_rows = 0;
    int grid = 0;
    std::vector<int32_t> order;

    DBuf<AdjJob> dJobs;
    DBuf<AdjSeg> dSegs;
    DBuf<RevIns> dRev;
    DBuf<AdjStore> dStores;
    DBuf<Ins> dCode;
    DBuf<int32_t> dOrder, dSample;
    DBuf<ProgMeta> dMeta;
    DBuf<AdjConst> dConsts;
    DBuf<uint32_t> dProgConsts;
    DBuf<float> dGpart, dFpart, dFit, dGraw, dGscale, dGout, dCur, dCbest, dGbest, dFbest, dLam;
    DBuf<uint32_t> dFbad, dBad, dQueue;

    ppi_adjoint_stats stats{};
};

namespace {

bool differentiable(const ppi_config& c, std::string* why) {
    if (c.error_fn != PPI_EF_ABS && c.error_fn != PPI_EF_SQUARED) {
        *why = "the adjoint supports the ABS and SQUARED error functions";
        return false;
    }
    if (c.reducemax) {
        *why = "the adjoint supports the sum reduction only (REDUCEMAX is off)";
        return false;
    }
    return true;
}

AdjArgs make_args(ppi_adjoint* a) {
    const ppi_engine* E = a->eng;
    AdjArgs A{};
    A.jobs = a->dJobs.p;
    A.segs = a->dSegs.p;
    A.rev = a->dRev.p;
    A.stores = a->dStores.p;
    A.code = a->dCode.p;
    A.order = a->dOrder.p;
    A.njobs = (int32_t)a->T.jobs.size();
    A.nranges = a->nranges;
    A.range_rows = a->range_rows;
    A.X = E->dX;
    A.ldx = E->ldx;
    A.nlin = E->nlin;
    A.ncol = E->ncol;
    A.ycol = E->ncol - 1;
    A.error_fn = a->cfg.error_fn;
    A.nprog = a->n;
    A.nconst = (int64_t)a->T.consts.size();
    A.gpart = a->dGpart.p;
    A.fpart = a->dFpart.p;
    A.fbad = a->dFbad.p;
    A.queue = a->dQueue.p;
    return A;
}

AdjFinal make_final(ppi_adjoint* a, const float* gscale, float* gout) {
    AdjFinal F{};
    F.meta = a->dMeta.p;
    F.consts = a->dConsts.p;
    F.nprog = a->n;
    F.nconst = (int64_t)a->T.consts.size();
    F.nranges = a->nranges;
    F.nlin = a->eng->nlin;
    F.alpha = a->alpha;
    F.fpart = a->dFpart.p;
    F.fbad = a->dFbad.p;
    F.gpart = a->dGpart.p;
    F.fit = a->dFit.p;
    F.bad = a->dBad.p;
    F.graw = a->dGraw.p;
    F.gscale = gscale;
    F.gout = gout;
    F.stride = a->stride;
    return F;
}

void forget_tuning(ppi_adjoint* a) {
    a->tape_ps = -1;
    a->tape_r = 0;
    for (double& t : a->cand_ms) t = -1.0;
}

// Compile the kept programs for partial store `ps` at `rows` rows per lane and
// upload the tape (skipped when that tape is current).
ppi_status compile_tape(ppi_adjoint* a, int ps, int rows) {
    if (a->tape_ps == ps && a->tape_r == rows) return PPI_OK;
    const double t0 = now_ms();
    std::string err;
    a->tape_ps = -1;  // the device tape is replaced below
    a->tape_r = 0;
    const int32_t B = adjoint_store_partials(ps, rows);
    if (!compile_adjoint(a->eng->symtab, a->ph.data(), a->eph.data(), a->sz.data(), a->stride, a->n, a->eng->ncol, B,
                         &a->T, &err)) {
        set_error(err);
        return PPI_E_PROGRAM;
    }
    AdjTape& T = a->T;
    for (const AdjConst& c : T.consts)
        if (c.ins >= T.fwd.code.size()) {
            set_error("program " + std::to_string(c.prog) + ": constant at " + std::to_string(c.pos) + " has no tape slot");
            return PPI_E_PROGRAM;
        }
    T.fwd.code.insert(T.fwd.code.end(), kTapePad, Ins{H_END, kEmpty});
    // Window padding: the kernels read reverse entries up to 127 past a segment's
    // first and store entries up to 63 past its first.
    T.rev.insert(T.rev.end(), 128, RevIns{0u, 0u});
    T.stores.insert(T.stores.end(), 64, AdjStore{0u, 0u});
    const size_t nj = T.jobs.size();
    a->order.resize(nj);
    std::iota(a->order.begin(), a->order.end(), 0);
    std::stable_sort(a->order.begin(), a->order.end(),
                     [&](int32_t x, int32_t y) { return T.jobs[(size_t)x].cost > T.jobs[(size_t)y].cost; });
    ppi_status s;
    hipStream_t st = a->stream;
    if ((s = upload(a->dJobs, T.jobs, st)) || (s = upload(a->dSegs, T.segs, st)) || (s = upload(a->dRev, T.rev, st)) ||
        (s = upload(a->dStores, T.stores, st)) || (s = upload(a->dCode, T.fwd.code, st)) ||
        (s = upload(a->dOrder, a->order, st)) || (s = upload(a->dMeta, T.fwd.meta, st)) ||
        (s = upload(a->dConsts, T.consts, st)) || (s = upload(a->dProgConsts, T.prog_consts, st)))
        return s;
    const size_t nc = T.consts.size();
    if ((s = a->dFit.reserve((size_t)a->n)) || (s = a->dBad.reserve((size_t)a->n)) || (s = a->dGraw.reserve(nc)) ||
        (s = a->dCur.reserve(nc)) || (s = a->dCbest.reserve(nc)) || (s = a->dGbest.reserve(nc)) ||
        (s = a->dFbest.reserve((size_t)a->n)) || (s = a->dLam.reserve((size_t)a->n)))
        return s;
    A_HIP(hipStreamSynchronize(st));
    a->tape_ps = ps;
    a->tape_r = rows;
    const int64_t nlin = a->eng->nlin;
    double fwd = 0.0, rev = 0.0;
    for (const AdjJob& j : T.jobs) {
        if (j.flags & kEmpty) continue;
        for (uint32_t k = 0; k < j.nseg; ++k) {
            const AdjSeg& sg = T.segs[j.seg_off + k];
            fwd += sg.fwd_end;
            rev += sg.rev_end - sg.rev_begin;
        }
    }
    a->stats.forward_ins = fwd * (double)nlin;
    a->stats.reverse_ins = rev * (double)nlin;
    a->stats.n_jobs = (int32_t)nj;
    a->stats.n_consts = (int32_t)nc;
    a->stats.max_segments = T.max_seg;
    a->stats.node_evals = (double)T.fwd.stats.total_size * (double)nlin;
    a->stats.compile_ms = now_ms() - t0;
    return PPI_OK;
}

// Tape for (ps, rows) plus the kernel shape and launch plan with features in
// VGPRs (feat 12), read from the dataset (feat 0), or the default (-1: in VGPRs
// when the programs read only the first 12 columns).  PPI_E_UNSUPPORTED: no
// kernel shape holds these programs that way (their stack or adjoint slots, or
// the columns they read; see kern_adjoint*.hip).
ppi_status prepare(ppi_adjoint* a, int ps, int rows, int feat = -1) {
    ppi_status s = compile_tape(a, ps, rows);
    if (s) return s;
    const AdjTape& T = a->T;
    const bool cols_fit = T.fwd.stats.max_col < 12;
    if (rows == 2) {
        // One register tuple of 10 per stack: 5 slots per row.
        a->S = 5;
        a->AS = 5;
        if (T.fwd.stats.max_slots > 5 || T.max_adj_slots > 5) {
            set_error("programs need more than 5 stack or adjoint slots: one row per lane");
            return PPI_E_UNSUPPORTED;
        }
    } else {
        a->S = T.fwd.stats.max_slots <= 12 ? 12 : 16;
        a->AS = kAdjSlots;
    }
    const bool vgpr_ok = cols_fit && rows == 1 && a->S == 12;  // two rows per lane: dataset reads only
    if (feat == 12 && !vgpr_ok) {
        set_error("features in registers: one row per lane, programs reading the first 12 columns, 12 stack slots");
        return PPI_E_UNSUPPORTED;
    }
    a->F = feat == 0 ? 0 : (vgpr_ok ? 12 : 0);
    a->R = rows;
    if (!adjoint_shape_supported(a->R, a->F, a->S, a->AS, ps)) {
        set_error("no adjoint kernel shape for these programs");
        return PPI_E_UNSUPPORTED;
    }
    // Persistent grid: the workgroups resident at once.  Row ranges: units
    // outnumber the wave slots 32 times over (at least 8 row tiles per range), so
    // the launch's tail, one unit long, is short.
    const bool P = a->cfg.protected_ops != 0, N = a->cfg.native_math != 0;
    const int occ = std::max(1, adjoint_occupancy(a->R, a->F, a->S, a->AS, ps, P, N));
    a->grid = a->dev_cu * occ;
    const int64_t slots = (int64_t)a->grid * 4;
    const int64_t nlin = a->eng->nlin;
    const int64_t nj = (int64_t)T.jobs.size();
    const int64_t tile = 64 * a->R;
    const int64_t max_ranges = std::max<int64_t>(1, nlin / (tile * 8));
    int64_t want = (32 * slots + nj - 1) / std::max<int64_t>(1, nj);
    want = std::max<int64_t>(1, std::min(want, max_ranges));
    int64_t rr = (nlin + want - 1) / want;
    rr = (rr + tile - 1) / tile * tile;
    a->range_rows = rr;
    a->nranges = (int32_t)((nlin + rr - 1) / rr);
    if ((uint64_t)nj * (uint64_t)a->nranges + (uint64_t)slots >= 0xffffffffull) {
        set_error("too many adjoint units");
        return PPI_E_ARG;
    }
    const size_t nc = T.consts.size();
    if ((s = a->dGpart.reserve((size_t)a->nranges * nc)) || (s = a->dFpart.reserve((size_t)a->nranges * a->n)) ||
        (s = a->dFbad.reserve((size_t)a->nranges * a->n)) || (s = a->dQueue.reserve(1)))
        return s;
    a->stats.rows_per_lane = a->R;
    a->stats.feature_slots = a->F;
    a->stats.ranges = a->nranges;
    return PPI_OK;
}

ppi_status enqueue_adjoint(ppi_adjoint* a, const AdjArgs& A, int grid) {
    const bool P = a->cfg.protected_ops != 0, N = a->cfg.native_math != 0;
    A_HIP(launch_fill_u32(a->dQueue.p, 0u, 1, a->stream));
    A_HIP(launch_adjoint(A, grid, a->R, a->F, a->S, a->AS, a->tape_ps, P, N, a->stream));
    return PPI_OK;
}

// Device ms of the adjoint kernel over a sample: an evenly spaced subset of the
// jobs (heaviest first) times ranges of 1024 rows spread over the whole dataset
// (up to 256 ranges), 32 units per wave slot of the persistent grid.  Short units,
// many per wave, keep the sample's tail (one unit) small, so shapes of different
// occupancy compare as in the full launch; the waves in flight stream rows from all
// over the dataset as the full launch's long ranges do.
ppi_status sample_ms(ppi_adjoint* a, double* ms) {
    const size_t nj = a->T.jobs.size();
    const int64_t rr = std::min<int64_t>(a->range_rows, 1024);
    const int32_t nr = (int32_t)std::max<int64_t>(1, std::min<int64_t>(256, a->eng->nlin / rr));
    const size_t units = (size_t)a->grid * 4 * 32;
    const size_t want = std::max<size_t>(1, std::min<size_t>(nj, units / (size_t)nr));
    std::vector<int32_t> pick(want);
    for (size_t k = 0; k < want; ++k) pick[k] = a->order[k * nj / want];
    // The per-range sums are indexed by range: hold nr ranges (the full launch may use fewer).
    const size_t nc = a->T.consts.size();
    ppi_status s;
    if ((s = a->dGpart.reserve((size_t)nr * nc)) || (s = a->dFpart.reserve((size_t)nr * a->n)) ||
        (s = a->dFbad.reserve((size_t)nr * a->n)))
        return s;
    if ((s = upload(a->dSample, pick, a->stream))) return s;
    AdjArgs A = make_args(a);
    A.order = a->dSample.p;
    A.njobs = (int32_t)want;
    A.nranges = nr;
    A.range_rows = rr;
    const int grid = (int)std::min<int64_t>(a->grid, ((int64_t)want * nr + 3) / 4);
    float best = 0.0f;
    for (int rep = 0; rep < 3; ++rep) {
        A_HIP(hipEventRecord(a->ev[2], a->stream));
        if ((s = enqueue_adjoint(a, A, grid))) return s;
        A_HIP(hipEventRecord(a->ev[3], a->stream));
        A_HIP(hipEventSynchronize(a->ev[3]));
        float t = 0.0f;
        A_HIP(hipEventElapsedTime(&t, a->ev[2], a->ev[3]));
        if (rep == 1 || (rep == 2 && t < best)) best = t;
    }
    *ms = best;
    return PPI_OK;
}

// Resolve PPI_PARTIALS_* and the rows per lane to a kernel: every candidate
// (store: both for AUTO; rows per lane: 1 and 2 unless one was requested;
// features in VGPRs or read from the dataset) whose shape holds the programs is
// timed on a sample of the units, and the fastest is prepared.  Timings are kept
// until the programs or data change.
ppi_status choose(ppi_adjoint* a, int32_t partials, int* ps_out) {
    int ps_lo = 0, ps_hi = 1;
    if (partials == PPI_PARTIALS_RECOMPUTE) ps_lo = ps_hi = 0;
    else if (partials == PPI_PARTIALS_STORE) ps_lo = ps_hi = 1;
    else if (partials != PPI_PARTIALS_AUTO) { set_error("bad partials mode"); return PPI_E_ARG; }
    const int r_lo = a->want_r ? a->want_r : 1, r_hi = a->want_r ? a->want_r : 2;
    int best = -1;
    double best_ms = 0.0;
    for (int r = r_lo; r <= r_hi; ++r)
        for (int ps = ps_lo; ps <= ps_hi; ++ps)
            for (int fi = 0; fi < 2; ++fi) {
                const int c = ((r - 1) * 2 + ps) * 2 + fi;
                if (a->cand_ms[c] == -1.0) {
                    ppi_status s = prepare(a, ps, r, fi ? 0 : 12);
                    if (s == PPI_E_UNSUPPORTED) {
                        a->cand_ms[c] = -2.0;
                        a->stats.tune_ms[c] = -2.0;
                        continue;
                    }
                    if (s) return s;
                    double t = 0.0;
                    if ((s = sample_ms(a, &t))) return s;
                    a->cand_ms[c] = t;
                    a->stats.tune_ms[c] = t;
                }
                if (a->cand_ms[c] < 0.0) continue;
                if (best < 0 || a->cand_ms[c] < best_ms) {
                    best = c;
                    best_ms = a->cand_ms[c];
                }
            }
    if (best < 0) {
        set_error(a->want_r == 2 ? "programs need more than 5 stack or adjoint slots: one row per lane"
                                 : "no adjoint kernel shape for these programs");
        return PPI_E_UNSUPPORTED;
    }
    const int ps = (best >> 1) & 1, rows = best / 4 + 1, feat = (best & 1) ? 0 : 12;
    ppi_status s = prepare(a, ps, rows, feat);
    if (s) return s;
    *ps_out = ps;
    a->stats.partials = ps ? PPI_PARTIALS_STORE : PPI_PARTIALS_RECOMPUTE;
    return PPI_OK;
}

ppi_status run_pass(ppi_adjoint* a, const float* gscale, float* gout, bool time_kernel) {
    AdjArgs A = make_args(a);
    ppi_status s;
    if (time_kernel) A_HIP(hipEventRecord(a->ev[2], a->stream));
    if ((s = enqueue_adjoint(a, A, a->grid))) return s;
    if (time_kernel) A_HIP(hipEventRecord(a->ev[3], a->stream));
    A_HIP(launch_adjoint_finalize(make_final(a, gscale, gout), a->stream));
    return PPI_OK;
}

ppi_status check_ready(ppi_adjoint* a) {
    if (!a) { set_error("adjoint is NULL"); return PPI_E_ARG; }
    if (!a->configured) { set_error("configure first"); return PPI_E_STATE; }
    if (!a->have_programs) { set_error("call forward first"); return PPI_E_STATE; }
    (void)hipSetDevice(a->device);
    return PPI_OK;
}

}  // namespace

extern "C" {

ppi_status ppi_adjoint_create(ppi_adjoint** out, int32_t device) {
    if (!out) { set_error("out is NULL"); return PPI_E_ARG; }
    *out = nullptr;
    ppi_engine* E = nullptr;
    ppi_status s = ppi_engine_create(&E, device);
    if (s) return s;
    ppi_adjoint* a = new ppi_adjoint();
    a->eng = E;
    a->device = E->device;
    a->dev_cu = E->dev_cu;
    hipError_t e = hipStreamCreateWithFlags(&a->stream, hipStreamNonBlocking);
    for (int k = 0; k < 4 && e == hipSuccess; ++k) e = hipEventCreate(&a->ev[k]);
    if (e != hipSuccess) {
        set_error(std::string("device init: ") + hipGetErrorString(e));
        ppi_adjoint_destroy(a);
        return PPI_E_DEVICE;
    }
    *out = a;
    return PPI_OK;
}

void ppi_adjoint_destroy(ppi_adjoint* a) {
    if (!a) return;
    (void)hipSetDevice(a->device);
    if (a->stream) (void)hipStreamSynchronize(a->stream);
    a->dJobs.release(); a->dSegs.release(); a->dRev.release(); a->dStores.release(); a->dCode.release();
    a->dOrder.release(); a->dSample.release(); a->dMeta.release(); a->dConsts.release(); a->dProgConsts.release();
    a->dGpart.release(); a->dFpart.release(); a->dFit.release(); a->dGraw.release(); a->dGscale.release();
    a->dGout.release(); a->dCur.release(); a->dCbest.release(); a->dGbest.release(); a->dFbest.release();
    a->dLam.release(); a->dFbad.release(); a->dBad.release(); a->dQueue.release();
    for (auto& ev : a->ev)
        if (ev) (void)hipEventDestroy(ev);
    if (a->stream) (void)hipStreamDestroy(a->stream);
    if (a->eng) ppi_engine_destroy(a->eng);
    delete a;
}

ppi_status ppi_adjoint_configure(ppi_adjoint* a, const ppi_config* cfg, const ppi_symbol* table, int32_t ntable) {
    if (!a || !cfg) { set_error("adjoint or config is NULL"); return PPI_E_ARG; }
    std::string why;
    if (!differentiable(*cfg, &why)) { set_error(why); return PPI_E_UNSUPPORTED; }
    ppi_status s = ppi_engine_configure(a->eng, cfg, table, ntable);
    if (s) return s;
    a->cfg = *cfg;
    a->cfg.cost = nullptr;
    a->configured = true;
    a->have_programs = false;
    forget_tuning(a);
    return PPI_OK;
}

ppi_status ppi_adjoint_set_dataset_rowmajor(ppi_adjoint* a, const float* data, int32_t nlin, int32_t ncol) {
    if (!a) { set_error("adjoint is NULL"); return PPI_E_ARG; }
    ppi_status s = ppi_engine_set_dataset_rowmajor(a->eng, data, nlin, ncol);
    if (s) return s;
    a->have_programs = false;
    forget_tuning(a);
    return PPI_OK;
}

ppi_status ppi_adjoint_forward(ppi_adjoint* a, const int32_t* phenotype, const float* ephemeral, const int32_t* size,
                               int32_t stride, int32_t n, float alpha, float* fitness) {
    if (!a) { set_error("adjoint is NULL"); return PPI_E_ARG; }
    if (!a->configured) { set_error("configure first"); return PPI_E_STATE; }
    if (n <= 0 || stride <= 0 || !phenotype || !ephemeral || !size) { set_error("bad program arrays"); return PPI_E_ARG; }
    std::vector<float> fit((size_t)n);
    ppi_status s = ppi_engine_evaluate(a->eng, phenotype, ephemeral, size, stride, n, alpha, fit.data(), nullptr, nullptr);
    if (s) return s;
    a->ph.assign(phenotype, phenotype + (size_t)n * stride);
    a->eph.assign(ephemeral, ephemeral + (size_t)n * stride);
    a->sz.assign(size, size + n);
    a->stride = stride;
    a->n = n;
    a->alpha = alpha;
    a->have_programs = true;
    forget_tuning(a);
    if (fitness) std::memcpy(fitness, fit.data(), (size_t)n * sizeof(float));
    return PPI_OK;
}

ppi_status ppi_adjoint_backward(ppi_adjoint* a, const float* grad_fitness, int32_t partials, float* grad) {
    ppi_status s = check_ready(a);
    if (s) return s;
    if (!grad) { set_error("grad is NULL"); return PPI_E_ARG; }
    int ps = 0;
    if ((s = choose(a, partials, &ps))) return s;
    const size_t total = (size_t)a->n * a->stride;
    if ((s = a->dGout.reserve(total))) return s;
    const float* gs = nullptr;
    if (grad_fitness) {
        if ((s = a->dGscale.reserve((size_t)a->n))) return s;
        A_HIP(hipMemcpyAsync(a->dGscale.p, grad_fitness, (size_t)a->n * sizeof(float), hipMemcpyHostToDevice, a->stream));
        gs = a->dGscale.p;
    }
    A_HIP(hipMemsetAsync(a->dGout.p, 0, total * sizeof(float), a->stream));
    A_HIP(hipEventRecord(a->ev[0], a->stream));
    if ((s = run_pass(a, gs, a->dGout.p, true))) return s;
    A_HIP(hipEventRecord(a->ev[1], a->stream));
    A_HIP(hipMemcpyAsync(grad, a->dGout.p, total * sizeof(float), hipMemcpyDeviceToHost, a->stream));
    A_HIP(hipStreamSynchronize(a->stream));
    float ms = 0.0f;
    A_HIP(hipEventElapsedTime(&ms, a->ev[0], a->ev[1]));
    a->stats.backward_ms = ms;
    A_HIP(hipEventElapsedTime(&ms, a->ev[2], a->ev[3]));
    a->stats.kernel_ms = ms;
    return PPI_OK;
}

ppi_status ppi_adjoint_descend(ppi_adjoint* a, int32_t steps, float step, int32_t partials, float* ephemeral_out,
                               float* fitness_out) {
    ppi_status s = check_ready(a);
    if (s) return s;
    if (steps < 0 || !(step > 0.0f) || !std::isfinite(step)) { set_error("need steps >= 0 and a finite step > 0"); return PPI_E_ARG; }
    int ps = 0;
    if ((s = choose(a, partials, &ps))) return s;
    // Plain-evaluator fitness of the starting point, for the final no-regression check.
    std::vector<float> f_start((size_t)a->n), f_end((size_t)a->n);
    if ((s = ppi_engine_evaluate(a->eng, a->ph.data(), a->eph.data(), a->sz.data(), a->stride, a->n, a->alpha,
                                 f_start.data(), nullptr, nullptr)))
        return s;
    const std::vector<float> eph_start = a->eph;
    AdjTape& T = a->T;
    const size_t nc = T.consts.size();
    std::vector<float> cur(nc);
    for (size_t k = 0; k < nc; ++k) cur[k] = a->eph[(size_t)T.consts[k].prog * a->stride + (size_t)T.consts[k].pos];
    if (nc) A_HIP(hipMemcpyAsync(a->dCur.p, cur.data(), nc * sizeof(float), hipMemcpyHostToDevice, a->stream));
    AdjDescend D{};
    D.prog_consts = a->dProgConsts.p;
    D.consts = a->dConsts.p;
    D.nprog = a->n;
    D.fit = a->dFit.p;
    D.bad = a->dBad.p;
    D.graw = a->dGraw.p;
    D.cur = a->dCur.p;
    D.cbest = a->dCbest.p;
    D.gbest = a->dGbest.p;
    D.fbest = a->dFbest.p;
    D.lam = a->dLam.p;
    D.code = a->dCode.p;
    D.step0 = step;
    A_HIP(hipEventRecord(a->ev[0], a->stream));
    for (int32_t it = 0; it <= steps; ++it) {
        if ((s = run_pass(a, nullptr, nullptr, false))) return s;
        D.iter = it;
        D.last = it == steps ? 1 : 0;
        A_HIP(launch_adjoint_descend(D, a->stream));
    }
    A_HIP(hipEventRecord(a->ev[1], a->stream));
    if (nc) A_HIP(hipMemcpyAsync(cur.data(), a->dCbest.p, nc * sizeof(float), hipMemcpyDeviceToHost, a->stream));
    A_HIP(hipStreamSynchronize(a->stream));
    float ms = 0.0f;
    A_HIP(hipEventElapsedTime(&ms, a->ev[0], a->ev[1]));
    a->stats.descend_ms = ms;
    a->stats.descend_steps = steps;
    for (size_t k = 0; k < nc; ++k) a->eph[(size_t)T.consts[k].prog * a->stride + (size_t)T.consts[k].pos] = cur[k];
    // The device tape now holds cbest.  Acceptance above used the adjoint pass's own
    // fitness, whose row sums run in another order than the plain evaluator's: a
    // program "improved" by less than that rounding keeps its original constants.
    if ((s = ppi_engine_evaluate(a->eng, a->ph.data(), a->eph.data(), a->sz.data(), a->stride, a->n, a->alpha,
                                 f_end.data(), nullptr, nullptr)))
        return s;
    int32_t reverted = 0;
    for (int32_t p = 0; p < a->n; ++p) {
        if (!(f_end[(size_t)p] > f_start[(size_t)p])) continue;
        const size_t base = (size_t)p * a->stride;
        std::memcpy(&a->eph[base], &eph_start[base], (size_t)a->stride * sizeof(float));
        f_end[(size_t)p] = f_start[(size_t)p];
        ++reverted;
    }
    if (reverted) {  // the device tape holds cbest for the reverted programs: recompile on next use
        a->tape_ps = -1;
        a->tape_r = 0;
    }
    a->stats.descend_reverted = reverted;
    if (fitness_out) std::memcpy(fitness_out, f_end.data(), f_end.size() * sizeof(float));
    if (ephemeral_out) std::memcpy(ephemeral_out, a->eph.data(), a->eph.size() * sizeof(float));
    return PPI_OK;
}

ppi_status ppi_adjoint_time(ppi_adjoint* a, int32_t iters, int32_t partials, double* ms_per_pass) {
    ppi_status s = check_ready(a);
    if (s) return s;
    if (iters <= 0 || !ms_per_pass) { set_error("bad arguments"); return PPI_E_ARG; }
    int ps = 0;
    if ((s = choose(a, partials, &ps))) return s;
    double kern = 0.0;
    A_HIP(hipEventRecord(a->ev[0], a->stream));
    for (int32_t k = 0; k < iters; ++k) {
        if ((s = run_pass(a, nullptr, nullptr, true))) return s;
        A_HIP(hipEventSynchronize(a->ev[3]));
        float t = 0.0f;
        A_HIP(hipEventElapsedTime(&t, a->ev[2], a->ev[3]));
        kern += t;
    }
    A_HIP(hipEventRecord(a->ev[1], a->stream));
    A_HIP(hipEventSynchronize(a->ev[1]));
    float ms = 0.0f;
    A_HIP(hipEventElapsedTime(&ms, a->ev[0], a->ev[1]));
    *ms_per_pass = ms / iters;
    a->stats.backward_ms = ms / iters;
    a->stats.kernel_ms = kern / iters;
    return PPI_OK;
}

ppi_status ppi_adjoint_set_rows_per_lane(ppi_adjoint* a, int32_t rows_per_lane) {
    if (!a) { set_error("adjoint is NULL"); return PPI_E_ARG; }
    if (rows_per_lane < 0 || rows_per_lane > 2) { set_error("rows_per_lane must be 0, 1 or 2"); return PPI_E_ARG; }
    a->want_r = rows_per_lane;
    return PPI_OK;
}

ppi_status ppi_adjoint_get_stats(const ppi_adjoint* a, ppi_adjoint_stats* out) {
    if (!a || !out) { set_error("adjoint or out is NULL"); return PPI_E_ARG; }
    *out = a->stats;
    return PPI_OK;
}

}  // extern "C"
