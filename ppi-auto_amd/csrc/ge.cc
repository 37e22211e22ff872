// ge.cc — grammatical evolution on the device: the runtime and C ABI of ppi_ge_* (include/ppi_engine.h).
//
// Replaces the host side of PPI's generation loop (src/ppi.cc: ppi_generate_population, ppi_evolve,
// ppi_evaluate with its decode loop) for one island: genomes, phenotypes, the compiled tape and the
// fitness vector stay in HBM, and a generation is a chain of kernels on the engine's stream —
// breeding plan (one thread per random stream), offspring, decode, compile, pack, evaluation (the
// engine's own kernels and semantics), the reference's epilogue and best-individual tracking.  The
// host reads back 16 bytes of compile totals per generation, to pick the evaluation kernel's shape.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/ppi_engine.h"
#include "engine_internal.h"
#include "ge.h"
#include "kernels.h"

using namespace ppi;

namespace {

#define GE_HIP(call)                                                                           \
    do {                                                                                       \
        hipError_t e_ = (call);                                                                \
        if (e_ != hipSuccess) {                                                                \
            set_error(std::string(#call) + ": " + hipGetErrorString(e_));                      \
            return PPI_E_DEVICE;                                                               \
        }                                                                                      \
    } while (0)

template <typename T>
ppi_status dalloc(T** p, size_t n) {
    if (*p) (void)hipFree(*p);
    *p = nullptr;
    if (hipMalloc((void**)p, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess) {
        *p = nullptr;
        set_error("device allocation of " + std::to_string(n * sizeof(T)) + " B failed");
        return PPI_E_NOMEM;
    }
    return PPI_OK;
}

template <typename T>
ppi_status dgrow(T** p, size_t* cap, size_t n) {
    if (*p && n <= *cap) return PPI_OK;
    ppi_status s = dalloc(p, n);
    *cap = s ? 0 : std::max<size_t>(n, 1);
    return s;
}

double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

// XorShift128+'s state update without the output addition is linear over GF(2): s0' = s1,
// s1' = x ^ s1 ^ (x >> 17) ^ (s1 >> 26) with x = s0 ^ (s0 << 23).  Matrices are 128 columns of
// (s0, s1) word pairs: column b < 64 is the image of s0's bit b, 64 + b of s1's bit b.
void ppi::ge_jump_tables(uint64_t* out) {
    struct V { uint64_t a, b; };
    auto step = [](V v) {
        uint64_t x = v.a;
        const uint64_t y = v.b;
        x ^= x << 23;
        return V{y, x ^ y ^ (x >> 17) ^ (y >> 26)};
    };
    auto apply = [](const std::vector<V>& M, V v) {
        V r{0, 0};
        for (int b = 0; b < 64; ++b) {
            if ((v.a >> b) & 1) { r.a ^= M[b].a; r.b ^= M[b].b; }
            if ((v.b >> b) & 1) { r.a ^= M[64 + b].a; r.b ^= M[64 + b].b; }
        }
        return r;
    };
    auto square = [&](const std::vector<V>& M) {
        std::vector<V> R(128);
        for (int j = 0; j < 128; ++j) R[j] = apply(M, M[j]);
        return R;
    };
    std::vector<V> M(128);
    for (int j = 0; j < 128; ++j) M[j] = step(j < 64 ? V{1ull << j, 0} : V{0, 1ull << (j - 64)});
    for (int k = 1; k < kDrawBlock; k <<= 1) M = square(M);  // T^kDrawBlock
    for (int lvl = 0; lvl < kJumpLevels; ++lvl) {
        for (int j = 0; j < 128; ++j) {
            out[(size_t)lvl * 256 + 2 * j] = M[j].a;
            out[(size_t)lvl * 256 + 2 * j + 1] = M[j].b;
        }
        M = square(M);
    }
}

struct ppi_ge {
    ppi_engine* E = nullptr;
    bool configured = false, initialized = false;
    ppi_ge_params p{};
    GeShape S{};
    int n_it_cap = 0;
    // grammar
    int32_t *d_first = nullptr, *d_nrules = nullptr, *d_len = nullptr, *d_sym = nullptr;
    GeDevGrammar G{};
    int32_t* d_symop = nullptr;
    float* d_symval = nullptr;
    // population state: genomes [2][P*W] (cur = the last evaluated), fitness [2][P]
    uint64_t* d_rng = nullptr;
    uint32_t* d_gen[2] = {nullptr, nullptr};
    float* d_fit[2] = {nullptr, nullptr};
    int cur = 0;
    uint32_t* d_best_genome = nullptr;
    float* d_best_fit = nullptr;
    int32_t* d_best_idx = nullptr;
    uint32_t* d_immigrants = nullptr;
    int32_t n_immigrants = 0;
    GePlan* d_plan = nullptr;
    int32_t* d_flips = nullptr;
    // parallel random streams: jump tables, draws, block states, iteration chains, consumed counts
    uint64_t* d_jump = nullptr;
    double* d_draw = nullptr;
    uint64_t* d_bstate = nullptr;
    int32_t* d_nxt = nullptr;
    int64_t* d_end = nullptr;
    GeDraws breed_draws{}, gen_draws{};
    // decode
    int32_t *d_ph = nullptr, *d_size = nullptr, *d_frames = nullptr;
    float* d_eph = nullptr;
    int64_t fstride = 0;
    // compile
    GeCompileScratch X{};
    GeCompiled* d_comp = nullptr;
    GeCompileTotals* d_tot = nullptr;
    GeCompileTotals h_tot{};
    uint32_t* d_offs = nullptr;
    ProgMeta* d_meta = nullptr;
    Ins* d_code = nullptr;
    // evaluation and epilogue
    float *d_partial = nullptr, *d_raw = nullptr, *d_vec = nullptr, *d_rowerr = nullptr, *d_osum = nullptr;
    uint32_t *d_bad_lo = nullptr, *d_bad_hi = nullptr, *d_ostop = nullptr;
    size_t cap_partial = 0, cap_rowerr = 0;
    float* d_gval = nullptr;
    int32_t* d_gidx = nullptr;
    hipEvent_t ev[6] = {};
    ppi_ge_stats st{};
    int32_t last_slots = 0;

    void release() {
        void* ptrs[] = {d_first, d_nrules, d_len, d_sym, d_symop, d_symval, d_rng, d_gen[0], d_gen[1], d_fit[0],
                        d_fit[1], d_best_genome, d_best_fit, d_best_idx, d_immigrants, d_plan, d_flips, d_jump,
                        d_draw, d_bstate, d_nxt, d_end, d_ph,
                        d_size, d_frames, d_eph, X.node, X.stk, X.fval, X.stage, d_comp, d_tot, d_offs, d_meta,
                        d_code, d_partial, d_raw, d_vec, d_rowerr, d_osum, d_bad_lo, d_bad_hi, d_ostop, d_gval,
                        d_gidx};
        for (void* q : ptrs)
            if (q) (void)hipFree(q);
        d_first = d_nrules = d_len = d_sym = d_symop = nullptr;
        d_symval = nullptr;
        d_rng = nullptr;
        d_gen[0] = d_gen[1] = nullptr;
        d_fit[0] = d_fit[1] = nullptr;
        d_best_genome = nullptr;
        d_best_fit = nullptr;
        d_best_idx = nullptr;
        d_immigrants = nullptr;
        d_plan = nullptr;
        d_flips = nullptr;
        d_jump = nullptr;
        d_draw = nullptr;
        d_bstate = nullptr;
        d_nxt = nullptr;
        d_end = nullptr;
        d_ph = d_size = d_frames = nullptr;
        d_eph = nullptr;
        X = GeCompileScratch{};
        d_comp = nullptr;
        d_tot = nullptr;
        d_offs = nullptr;
        d_meta = nullptr;
        d_code = nullptr;
        d_partial = d_raw = d_vec = d_rowerr = d_osum = nullptr;
        d_bad_lo = d_bad_hi = d_ostop = nullptr;
        cap_partial = cap_rowerr = 0;
        d_gval = nullptr;
        d_gidx = nullptr;
        configured = initialized = false;
    }
};

namespace {

// Decode, compile and evaluate population `which`; its fitness goes to d_fit[which] and the best
// individual so far is updated.  ev[1] .. ev[5] bracket the phases.
ppi_status evaluate_population(ppi_ge* g, int which) {
    ppi_engine* E = g->E;
    hipStream_t st = E->stream;
    const GeShape& S = g->S;
    const int P = S.P;
    GE_HIP(hipEventRecord(g->ev[1], st));
    GE_HIP(ge_launch_decode(S, g->G, g->d_gen[which], g->d_ph, g->d_eph, g->d_size, g->d_frames, g->fstride, st));
    GE_HIP(hipEventRecord(g->ev[2], st));
    GE_HIP(hipMemsetAsync(g->d_tot, 0, sizeof(GeCompileTotals), st));
    GE_HIP(ge_launch_compile(g->d_ph, g->d_eph, g->d_size, S.max_size, P, g->d_symop, g->d_symval, E->ncol,
                             E->cfg.no_constant_folding == 0, E->cfg.protected_ops != 0, E->cfg.native_math != 0,
                             g->X, g->d_comp, g->d_tot, st));
    GE_HIP(ge_launch_pack(g->X, g->d_comp, g->d_size, P, g->d_offs, g->d_meta, g->d_code, g->d_tot, st));
    GE_HIP(hipEventRecord(g->ev[3], st));
    GE_HIP(hipMemcpyAsync(&g->h_tot, g->d_tot, sizeof(GeCompileTotals), hipMemcpyDeviceToHost, st));
    GE_HIP(hipStreamSynchronize(st));
    if (g->h_tot.error & 1) {
        set_error("a program needs more than " + std::to_string(kMaxSlots) + " register slots");
        return PPI_E_PROGRAM;
    }
    if (g->h_tot.error & 2) {
        set_error("a decoded size lies outside [0, max_size_phenotype]");
        return PPI_E_PROGRAM;
    }
    CompileStats cs;
    cs.total_size = g->h_tot.total_size;
    cs.total_code = g->h_tot.total_code;
    cs.max_slots = g->h_tot.max_slots;
    cs.max_col = g->h_tot.max_col;
    g->last_slots = cs.max_slots;
    LaunchPlan L;
    std::string why;
    if (!compute_plan(E, cs, P, E->cfg.form, &L, &why)) {
        set_error(why);
        return PPI_E_PROGRAM;
    }
    ppi_status s;
    if (L.form == PPI_FORM_FP && (s = dgrow(&g->d_partial, &g->cap_partial, (size_t)L.nrb * P))) return s;
    if (L.form == PPI_FORM_ORDERED) {
        const int64_t tile = fp_rows_per_block(L.R);
        size_t free_b = 0, total_b = 0;
        if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) free_b = (size_t)4 << 30;
        const int64_t budget = std::min<int64_t>((int64_t)16 << 30, (int64_t)(free_b / 2));
        const int64_t per = E->cfg.ordered_pass_rows > 0 ? (E->cfg.ordered_pass_rows + tile - 1) / tile
                                                         : budget / ((int64_t)P * tile * (int64_t)sizeof(float));
        L.nrb_pass = (int)std::max<int64_t>(1, std::min<int64_t>(per, L.nrb));
        if ((s = dgrow(&g->d_rowerr, &g->cap_rowerr, (size_t)L.nrb_pass * P * tile))) return s;
    }
    EvalArgs A = make_args(E, 0.0f);
    A.meta = g->d_meta;
    A.code = g->d_code;
    A.nprog = P;
    A.chunk = L.chunk;
    A.order = nullptr;
    A.partial = g->d_partial;
    A.bad_lo = g->d_bad_lo;
    A.bad_hi = g->d_bad_hi;
    A.raw = g->d_raw;
    A.vec = g->d_vec;
    A.rb0 = 0;
    A.nrb_pass = L.nrb;
    A.rowerr = g->d_rowerr;
    A.osum = g->d_osum;
    A.ostop = g->d_ostop;
    if ((s = enqueue_with(E, A, L, false))) return s;
    GE_HIP(hipEventRecord(g->ev[4], st));
    // The reference's epilogue: fitness with alpha*size, the best individual (best_size = 1).
    GE_HIP(launch_rank(g->d_raw, g->d_meta, P, E->cfg.strategy, g->p.alpha, best_local_for(E, P), 1,
                       g->d_fit[which], g->d_gval, g->d_gidx, g->d_best_idx, st));
    GE_HIP(ge_launch_track_best(g->d_fit[which], g->d_best_idx, g->d_gen[which], S.W, g->d_best_fit,
                                g->d_best_genome, st));
    GE_HIP(hipEventRecord(g->ev[5], st));
    g->st.node_evals = (double)cs.total_size * E->nlin;
    g->st.max_slots = cs.max_slots;
    return PPI_OK;
}

void record_times(ppi_ge* g, bool bred) {
    float ms = 0.0f;
    g->st.breed_ms = (bred && hipEventElapsedTime(&ms, g->ev[0], g->ev[1]) == hipSuccess) ? ms : 0.0;
    g->st.decode_ms = hipEventElapsedTime(&ms, g->ev[1], g->ev[2]) == hipSuccess ? ms : 0.0;
    g->st.compile_ms = hipEventElapsedTime(&ms, g->ev[2], g->ev[3]) == hipSuccess ? ms : 0.0;
    g->st.evaluate_ms = hipEventElapsedTime(&ms, g->ev[3], g->ev[4]) == hipSuccess ? ms : 0.0;
    g->st.rank_ms = hipEventElapsedTime(&ms, g->ev[4], g->ev[5]) == hipSuccess ? ms : 0.0;
}

}  // namespace

extern "C" {

ppi_status ppi_ge_create(ppi_ge** out, ppi_engine* eng) {
    if (!out || !eng) { set_error("out or engine is NULL"); return PPI_E_ARG; }
    *out = nullptr;
    (void)hipSetDevice(eng->device);
    ppi_ge* g = new ppi_ge();
    g->E = eng;
    for (auto& e : g->ev)
        if (hipEventCreate(&e) != hipSuccess) {
            set_error("hipEventCreate failed");
            for (auto& f : g->ev)
                if (f) (void)hipEventDestroy(f);
            delete g;
            return PPI_E_DEVICE;
        }
    *out = g;
    return PPI_OK;
}

void ppi_ge_destroy(ppi_ge* g) {
    if (!g) return;
    (void)hipSetDevice(g->E->device);
    (void)hipStreamSynchronize(g->E->stream);
    g->release();
    for (auto& e : g->ev)
        if (e) (void)hipEventDestroy(e);
    delete g;
}

ppi_status ppi_ge_configure(ppi_ge* g, const ppi_ge_params* prm, const ppi_ge_grammar* gr) {
    if (!g || !prm || !gr) { set_error("NULL argument"); return PPI_E_ARG; }
    ppi_engine* E = g->E;
    if (!E->configured || !E->dX) { set_error("configure the engine and set its dataset first"); return PPI_E_STATE; }
    const ppi_ge_params& p = *prm;
    if (p.population_size <= 0 || p.number_of_bits <= 0 || p.bits_per_gene <= 0 || p.bits_per_gene > 31 ||
        p.bits_per_constant < 1 || p.bits_per_constant > 63 || p.tournament_size <= 0 || p.threads <= 0 ||
        p.mutation_rate < 0.0f || p.mutation_rate > 1.0f || p.max_phenotype < 0) {
        set_error("bad GE parameters");
        return PPI_E_ARG;
    }
    for (int t = 0; t < p.threads; ++t)
        if (((uint64_t)p.seed ^ (uint64_t)t) == 0) {
            set_error("seed ^ thread is 0 for thread " + std::to_string(t) + " (the reference seeds it from the clock)");
            return PPI_E_ARG;
        }
    if (gr->n_nonterminals <= 0 || !gr->n_rules || !gr->rule_len || !gr->rule_symbols || gr->max_rule <= 0 ||
        gr->max_rule > 255) {
        set_error("bad GE grammar");
        return PPI_E_ARG;
    }
    int total_rules = 0;
    for (int i = 0; i < gr->n_nonterminals; ++i) {
        if (gr->n_rules[i] <= 0) { set_error("a non-terminal without rules"); return PPI_E_ARG; }
        total_rules += gr->n_rules[i];
    }
    for (int r = 0; r < total_rules; ++r) {
        if (gr->rule_len[r] < 0 || gr->rule_len[r] > gr->max_rule) { set_error("bad rule length"); return PPI_E_ARG; }
        for (int k = 0; k < gr->rule_len[r]; ++k) {
            const int32_t s = gr->rule_symbols[(size_t)r * gr->max_rule + k];
            if (s < 10000 && (s < 0 || s >= gr->n_nonterminals)) {
                set_error("rule symbol " + std::to_string(s) + " is not a non-terminal of the grammar");
                return PPI_E_ARG;
            }
        }
    }
    (void)hipSetDevice(E->device);
    g->release();
    g->p = p;
    // ppi.cc:258-264: bits per gene raised to what the largest rule set needs
    int bg = p.bits_per_gene;
    for (int i = 0; i < gr->n_nonterminals; ++i) {
        const int need = (int)std::ceil(std::log10((double)gr->n_rules[i]) / std::log10(2.0));
        if (need > bg) bg = need;
    }
    int64_t mx = (int64_t)gr->max_rule * p.number_of_bits / bg;  // ppi.cc:307
    if (p.max_phenotype > 0 && p.max_phenotype < mx) mx = p.max_phenotype;
    if (mx <= 0 || mx > 16000) {
        set_error("max_size_phenotype " + std::to_string(mx) + " outside [1, 16000]");
        return PPI_E_ARG;
    }
    GeShape& S = g->S;
    S.P = p.population_size;
    S.nb = p.number_of_bits;
    S.W = (p.number_of_bits + 31) / 32;
    S.bg = bg;
    S.bc = p.bits_per_constant;
    S.max_size = (int32_t)mx;
    S.lo = p.min_constant > p.max_constant ? p.max_constant : p.min_constant;
    S.hi = p.min_constant > p.max_constant ? p.min_constant : p.max_constant;
    S.cp = p.crossover_rate;
    S.ts = p.tournament_size;
    S.T = p.threads;
    S.max_bits = (int)std::ceil(p.mutation_rate * (float)p.number_of_bits);  // float product (ppi.cc:743)
    const int P = S.P;
    // grammar
    std::vector<int32_t> first(gr->n_nonterminals);
    for (int i = 0, acc = 0; i < gr->n_nonterminals; ++i) { first[i] = acc; acc += gr->n_rules[i]; }
    ppi_status s;
    if ((s = dalloc(&g->d_first, first.size())) || (s = dalloc(&g->d_nrules, first.size())) ||
        (s = dalloc(&g->d_len, (size_t)total_rules)) || (s = dalloc(&g->d_sym, (size_t)total_rules * gr->max_rule)))
        return s;
    GE_HIP(hipMemcpy(g->d_first, first.data(), first.size() * 4, hipMemcpyHostToDevice));
    GE_HIP(hipMemcpy(g->d_nrules, gr->n_rules, first.size() * 4, hipMemcpyHostToDevice));
    GE_HIP(hipMemcpy(g->d_len, gr->rule_len, (size_t)total_rules * 4, hipMemcpyHostToDevice));
    GE_HIP(hipMemcpy(g->d_sym, gr->rule_symbols, (size_t)total_rules * gr->max_rule * 4, hipMemcpyHostToDevice));
    g->G = GeDevGrammar{g->d_first, g->d_nrules, g->d_len, g->d_sym, gr->max_rule, gr->t_const, gr->t_attribute};
    // the engine's symbol table, as a direct map of Symbols 10000 .. 18191
    std::vector<int32_t> symop(8192, -1);
    std::vector<float> symval(8192, 0.0f);
    for (int k = 0; k < 8192; ++k) {
        const ppi_symbol* e = E->symtab.find(10000 + k);
        if (e) { symop[k] = e->op; symval[k] = e->value; }
    }
    if ((s = dalloc(&g->d_symop, 8192)) || (s = dalloc(&g->d_symval, 8192))) return s;
    GE_HIP(hipMemcpy(g->d_symop, symop.data(), 8192 * 4, hipMemcpyHostToDevice));
    GE_HIP(hipMemcpy(g->d_symval, symval.data(), 8192 * 4, hipMemcpyHostToDevice));
    // state
    const int n_it = (P + 1) / 2;
    const int64_t NN = 2 * (int64_t)S.max_size + 2;
    g->fstride = (int64_t)S.nb / S.bg + 2;
    if ((s = dalloc(&g->d_rng, 2 * (size_t)S.T)) || (s = dalloc(&g->d_gen[0], (size_t)P * S.W)) ||
        (s = dalloc(&g->d_gen[1], (size_t)P * S.W)) || (s = dalloc(&g->d_fit[0], (size_t)P)) ||
        (s = dalloc(&g->d_fit[1], (size_t)P)) || (s = dalloc(&g->d_best_genome, (size_t)S.W)) ||
        (s = dalloc(&g->d_best_fit, 1)) || (s = dalloc(&g->d_best_idx, kRankMax)) ||
        (s = dalloc(&g->d_immigrants, (size_t)P * S.W)) || (s = dalloc(&g->d_plan, (size_t)n_it)) ||
        (s = dalloc(&g->d_flips, (size_t)n_it * 2 * std::max(1, S.max_bits))) ||
        (s = dalloc(&g->d_jump, (size_t)kJumpLevels * 256)) || (s = dalloc(&g->d_end, (size_t)S.T)) ||
        (s = dalloc(&g->d_ph, (size_t)P * S.max_size)) || (s = dalloc(&g->d_eph, (size_t)P * S.max_size)) ||
        (s = dalloc(&g->d_size, (size_t)P)) || (s = dalloc(&g->d_frames, (size_t)P * g->fstride)) ||
        (s = dalloc(&g->X.node, (size_t)P * NN * 4)) || (s = dalloc(&g->X.stk, (size_t)P * NN * 4)) ||
        (s = dalloc(&g->X.fval, (size_t)P * NN)) || (s = dalloc(&g->X.stage, (size_t)P * (S.max_size + 1))) ||
        (s = dalloc(&g->d_comp, (size_t)P)) || (s = dalloc(&g->d_tot, 1)) || (s = dalloc(&g->d_offs, (size_t)P + 1)) ||
        (s = dalloc(&g->d_meta, (size_t)P)) ||
        (s = dalloc(&g->d_code, (size_t)P * (S.max_size + 1) + kTapePad)) || (s = dalloc(&g->d_raw, (size_t)P)) ||
        (s = dalloc(&g->d_vec, (size_t)P)) || (s = dalloc(&g->d_bad_lo, (size_t)P)) ||
        (s = dalloc(&g->d_bad_hi, (size_t)P)) || (s = dalloc(&g->d_osum, (size_t)P)) ||
        (s = dalloc(&g->d_ostop, (size_t)P)) || (s = dalloc(&g->d_gval, (size_t)P)) ||
        (s = dalloc(&g->d_gidx, (size_t)P)))
        return s;
    // Random streams: a breeding step's draws per stream (its share of the iterations times the most
    // one iteration draws), the chain tables, and the block states of either phase.
    {
        const int64_t m_max = ((int64_t)n_it + S.T - 1) / S.T;
        const int64_t cap = ((m_max * ge_draws_per_iteration(S) + kDrawBlock - 1) / kDrawBlock) * kDrawBlock;
        int levels = 1;
        while (((int64_t)1 << levels) < m_max) ++levels;
        const int64_t gen_draws = (((int64_t)P + S.T - 1) / S.T) * S.nb;
        const int64_t gen_blk = (gen_draws + kDrawBlock - 1) / kDrawBlock;
        if (gen_blk >= ((int64_t)1 << kJumpLevels) || cap >= ((int64_t)1 << 31)) {
            set_error("population x number_of_bits too large for the random-stream tables");
            return PPI_E_ARG;
        }
        const int64_t blk_max = std::max<int64_t>(cap / kDrawBlock, gen_blk);
        if ((s = dalloc(&g->d_draw, (size_t)S.T * cap)) || (s = dalloc(&g->d_bstate, (size_t)S.T * (blk_max + 1) * 2)) ||
            (s = dalloc(&g->d_nxt, (size_t)levels * S.T * (cap + 1))))
            return s;
        std::vector<uint64_t> jump((size_t)kJumpLevels * 256);
        ge_jump_tables(jump.data());
        GE_HIP(hipMemcpy(g->d_jump, jump.data(), jump.size() * 8, hipMemcpyHostToDevice));
        g->breed_draws = GeDraws{g->d_jump, g->d_rng, g->d_draw, g->d_bstate, cap, (int32_t)(cap / kDrawBlock),
                                 g->d_nxt, levels, g->d_end};
        g->gen_draws = GeDraws{g->d_jump, g->d_rng, nullptr, g->d_bstate, gen_blk * kDrawBlock, (int32_t)gen_blk,
                               nullptr, 0, g->d_end};
    }
    g->X.node_stride = NN * 4;
    g->X.stk_stride = NN * 4;
    g->X.fval_stride = NN;
    g->X.stage_stride = S.max_size + 1;
    g->configured = true;
    g->initialized = false;
    g->n_immigrants = 0;
    g->st = ppi_ge_stats{};
    g->st.bits_per_gene = bg;
    g->st.max_size_phenotype = S.max_size;
    return PPI_OK;
}

ppi_status ppi_ge_initialize(ppi_ge* g) {
    if (!g) { set_error("ge is NULL"); return PPI_E_ARG; }
    if (!g->configured) { set_error("configure the GE first"); return PPI_E_STATE; }
    ppi_engine* E = g->E;
    (void)hipSetDevice(E->device);
    const double t0 = now_ms();
    // one generator per thread, Seed(seed ^ t): StoreSeed puts (s << 32) | (s & 0xffffffff) in both words
    std::vector<uint64_t> rng(2 * (size_t)g->S.T);
    for (int t = 0; t < g->S.T; ++t) {
        const uint64_t sd = (uint64_t)g->p.seed ^ (uint64_t)t;
        rng[2 * t] = rng[2 * t + 1] = (sd << 32) | (sd & 0xffffffffull);
    }
    GE_HIP(hipMemcpyAsync(g->d_rng, rng.data(), rng.size() * 8, hipMemcpyHostToDevice, E->stream));
    const float fmax = FLT_MAX;
    GE_HIP(hipMemcpyAsync(g->d_best_fit, &fmax, 4, hipMemcpyHostToDevice, E->stream));
    GE_HIP(hipMemsetAsync(g->d_best_genome, 0, (size_t)g->S.W * 4, E->stream));
    g->cur = 0;
    GE_HIP(hipEventRecord(g->ev[0], E->stream));
    GE_HIP(hipMemsetAsync(g->d_gen[0], 0, (size_t)g->S.P * g->S.W * 4, E->stream));
    GE_HIP(ge_launch_generate(g->S, g->gen_draws, g->d_gen[0], E->stream));
    ppi_status s = evaluate_population(g, 0);
    if (s) return s;
    GE_HIP(hipStreamSynchronize(E->stream));
    record_times(g, true);
    g->initialized = true;
    g->st.generation = 0;
    g->st.generation_ms = now_ms() - t0;
    g->n_immigrants = 0;
    return PPI_OK;
}

ppi_status ppi_ge_evolve(ppi_ge* g, int32_t generations) {
    if (!g || generations < 0) { set_error("bad arguments"); return PPI_E_ARG; }
    if (!g->initialized) { set_error("initialize the population first"); return PPI_E_STATE; }
    ppi_engine* E = g->E;
    hipStream_t st = E->stream;
    (void)hipSetDevice(E->device);
    const GeShape& S = g->S;
    for (int32_t k = 0; k < generations; ++k) {
        const double t0 = now_ms();
        const int nxt = 1 - g->cur;
        GE_HIP(hipEventRecord(g->ev[0], st));
        int nimm = std::min(g->n_immigrants, S.P);
        if (nimm > 0)
            GE_HIP(hipMemcpyAsync(g->d_gen[nxt], g->d_immigrants, (size_t)nimm * S.W * 4, hipMemcpyDeviceToDevice, st));
        if (g->p.elitism && nimm < S.P) {
            // ppi_clone(&best_individual, 0, &descendentes, nImmigrants); nImmigrants++ (ppi.cc:882-887)
            GE_HIP(hipMemcpyAsync(g->d_gen[nxt] + (size_t)nimm * S.W, g->d_best_genome, (size_t)S.W * 4,
                                  hipMemcpyDeviceToDevice, st));
            ++nimm;
        }
        if (nimm < S.P) {
            GE_HIP(ge_launch_plan(S, g->breed_draws, g->d_fit[g->cur], nimm, g->d_plan, g->d_flips, st));
            GE_HIP(ge_launch_offspring(S, g->d_plan, g->d_flips, g->d_gen[g->cur], g->d_gen[nxt], nimm, st));
        }
        g->n_immigrants = 0;
        ppi_status s = evaluate_population(g, nxt);
        if (s) return s;
        GE_HIP(hipStreamSynchronize(st));
        record_times(g, true);
        g->cur = nxt;
        g->st.generation++;
        g->st.generation_ms = now_ms() - t0;
    }
    return PPI_OK;
}

ppi_status ppi_ge_set_immigrants(ppi_ge* g, const uint8_t* genomes, int32_t n) {
    if (!g || n < 0 || (n > 0 && !genomes)) { set_error("bad arguments"); return PPI_E_ARG; }
    if (!g->configured) { set_error("configure the GE first"); return PPI_E_STATE; }
    if (n > g->S.P) { set_error("more immigrants than individuals"); return PPI_E_ARG; }
    const GeShape& S = g->S;
    std::vector<uint32_t> w((size_t)n * S.W, 0u);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < S.nb; ++j)
            if (genomes[(size_t)i * S.nb + j]) w[(size_t)i * S.W + (j >> 5)] |= 1u << (j & 31);
    (void)hipSetDevice(g->E->device);
    if (n > 0) GE_HIP(hipMemcpy(g->d_immigrants, w.data(), w.size() * 4, hipMemcpyHostToDevice));
    g->n_immigrants = n;
    return PPI_OK;
}

ppi_status ppi_ge_get_population(const ppi_ge* g, uint8_t* genomes, float* fitness) {
    if (!g) { set_error("ge is NULL"); return PPI_E_ARG; }
    if (!g->initialized) { set_error("initialize the population first"); return PPI_E_STATE; }
    const GeShape& S = g->S;
    (void)hipSetDevice(g->E->device);
    if (genomes) {
        std::vector<uint32_t> w((size_t)S.P * S.W);
        GE_HIP(hipMemcpy(w.data(), g->d_gen[g->cur], w.size() * 4, hipMemcpyDeviceToHost));
        for (int i = 0; i < S.P; ++i)
            for (int j = 0; j < S.nb; ++j) genomes[(size_t)i * S.nb + j] = (w[(size_t)i * S.W + (j >> 5)] >> (j & 31)) & 1u;
    }
    if (fitness) GE_HIP(hipMemcpy(fitness, g->d_fit[g->cur], (size_t)S.P * 4, hipMemcpyDeviceToHost));
    return PPI_OK;
}

ppi_status ppi_ge_best(const ppi_ge* g, float* fitness, uint8_t* genome) {
    if (!g) { set_error("ge is NULL"); return PPI_E_ARG; }
    if (!g->initialized) { set_error("initialize the population first"); return PPI_E_STATE; }
    (void)hipSetDevice(g->E->device);
    if (fitness) GE_HIP(hipMemcpy(fitness, g->d_best_fit, 4, hipMemcpyDeviceToHost));
    if (genome) {
        std::vector<uint32_t> w((size_t)g->S.W);
        GE_HIP(hipMemcpy(w.data(), g->d_best_genome, w.size() * 4, hipMemcpyDeviceToHost));
        for (int j = 0; j < g->S.nb; ++j) genome[j] = (w[(size_t)(j >> 5)] >> (j & 31)) & 1u;
    }
    return PPI_OK;
}

ppi_status ppi_ge_decoded(const ppi_ge* g, int32_t* phenotype, float* ephemeral, int32_t* size, int32_t* stride) {
    if (!g) { set_error("ge is NULL"); return PPI_E_ARG; }
    if (!g->initialized) { set_error("initialize the population first"); return PPI_E_STATE; }
    const GeShape& S = g->S;
    (void)hipSetDevice(g->E->device);
    if (stride) *stride = S.max_size;
    if (phenotype) GE_HIP(hipMemcpy(phenotype, g->d_ph, (size_t)S.P * S.max_size * 4, hipMemcpyDeviceToHost));
    if (ephemeral) GE_HIP(hipMemcpy(ephemeral, g->d_eph, (size_t)S.P * S.max_size * 4, hipMemcpyDeviceToHost));
    if (size) GE_HIP(hipMemcpy(size, g->d_size, (size_t)S.P * 4, hipMemcpyDeviceToHost));
    return PPI_OK;
}

ppi_status ppi_ge_tape(const ppi_ge* g, uint32_t* meta, uint32_t* code, int64_t* n_code) {
    if (!g || !n_code) { set_error("bad arguments"); return PPI_E_ARG; }
    if (!g->initialized) { set_error("initialize the population first"); return PPI_E_STATE; }
    (void)hipSetDevice(g->E->device);
    uint32_t total = 0;
    GE_HIP(hipMemcpy(&total, g->d_offs + g->S.P, 4, hipMemcpyDeviceToHost));
    *n_code = (int64_t)total + kTapePad;
    if (meta) GE_HIP(hipMemcpy(meta, g->d_meta, (size_t)g->S.P * sizeof(ProgMeta), hipMemcpyDeviceToHost));
    if (code) GE_HIP(hipMemcpy(code, g->d_code, (size_t)*n_code * sizeof(Ins), hipMemcpyDeviceToHost));
    return PPI_OK;
}

ppi_status ppi_ge_get_stats(const ppi_ge* g, ppi_ge_stats* out) {
    if (!g || !out) { set_error("bad arguments"); return PPI_E_ARG; }
    *out = g->st;
    out->best_fitness = FLT_MAX;
    if (g->initialized) (void)hipMemcpy(&out->best_fitness, g->d_best_fit, 4, hipMemcpyDeviceToHost);
    return PPI_OK;
}

}  // extern "C"
