// workload.cc — seeded synthetic populations and datasets (see workload.h).
//
// Random numbers: xoshiro256** seeded through splitmix64 (Blackman & Vigna,
// "Scrambled linear pseudorandom number generators", 2018); normals by the
// Box-Muller transform in double precision.
#include "workload.h"

#include <cmath>
#include <cstring>
#include <vector>

namespace {

struct Rng {
    uint64_t s[4];
    explicit Rng(uint64_t seed) {
        uint64_t x = seed;
        for (auto& w : s) {
            x += 0x9E3779B97F4A7C15ull;
            uint64_t z = x;
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
            w = z ^ (z >> 31);
        }
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next() {
        const uint64_t r = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
        s[2] ^= t; s[3] = rotl(s[3], 45);
        return r;
    }
    double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }          // [0,1)
    double uniform(double a, double b) { return a + (b - a) * uniform(); }
    int32_t below(int32_t n) { return (int32_t)(uniform() * n); }          // [0,n)
    double normal() {
        double u1 = uniform(), u2 = uniform();
        if (u1 < 1e-300) u1 = 1e-300;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};

struct Emitter {
    const ppi_wl_grammar* g;
    Rng* rng;
    int32_t* ph;
    float* eph;
    int32_t n;
    int32_t cap;
    int32_t consts;
    bool overflow;

    void leaf() {
        if (n >= cap) { overflow = true; return; }
        bool use_const = g->t_const >= 0 && rng->uniform() < g->p_const &&
                         (g->max_consts <= 0 || consts < g->max_consts);
        if (use_const) {
            ph[n] = g->t_const;
            eph[n] = (float)rng->uniform(g->const_lo, g->const_hi);
            ++consts;
        } else {
            ph[n] = g->t_attribute;
            eph[n] = (float)rng->below(g->n_attributes);
        }
        ++n;
    }
    bool op(bool* binary) {
        if (n >= cap) { overflow = true; return false; }
        int32_t k = rng->below(g->n_binary + g->n_unary);
        *binary = k < g->n_binary;
        ph[n] = *binary ? g->binary_ops[k] : g->unary_ops[k - g->n_binary];
        eph[n] = 0.0f;
        ++n;
        return true;
    }
    // Grow: explicit stack of pending depths (prefix order), no recursion.
    void grow(int32_t max_depth) {
        std::vector<int32_t> pending{0};
        while (!pending.empty() && !overflow) {
            int32_t d = pending.back();
            pending.pop_back();
            bool stop = d >= max_depth || (d > 0 && rng->uniform() < g->p_leaf);
            if (stop) { leaf(); continue; }
            bool bin;
            if (!op(&bin)) return;
            pending.push_back(d + 1);
            if (bin) pending.push_back(d + 1);
        }
    }
    // Exact length: a node of length L >= 3 is unary (child L-1) or binary with a
    // uniform split of the remaining L-1 symbols; L == 2 is unary; L == 1 a leaf.
    void exact(int32_t len) {
        std::vector<int32_t> pending{len};
        while (!pending.empty() && !overflow) {
            int32_t L = pending.back();
            pending.pop_back();
            if (L <= 1) { leaf(); continue; }
            bool want_bin = L >= 3 && g->n_binary > 0 &&
                            (g->n_unary == 0 || rng->below(g->n_binary + g->n_unary) < g->n_binary);
            if (n >= cap) { overflow = true; return; }
            if (want_bin) {
                ph[n] = g->binary_ops[rng->below(g->n_binary)];
                eph[n] = 0.0f;
                ++n;
                int32_t left = 1 + rng->below(L - 2);
                pending.push_back(L - 1 - left);  // right child, emitted after the left one
                pending.push_back(left);
            } else {
                ph[n] = g->unary_ops[rng->below(g->n_unary)];
                eph[n] = 0.0f;
                ++n;
                pending.push_back(L - 1);
            }
        }
    }
};

bool grammar_ok(const ppi_wl_grammar* g) {
    return g && g->n_attributes > 0 && (g->n_binary > 0 || g->n_unary > 0) &&
           (g->n_binary == 0 || g->binary_ops) && (g->n_unary == 0 || g->unary_ops);
}

// Random cubic polynomial with up to 12 monomials of degree 1..3, coefficients U[-1,1].
struct Poly {
    int32_t nm = 0;
    int32_t var[12][3];
    int32_t deg[12];
    double coef[12];
    Poly(Rng& r, int32_t nfeat) {
        nm = 4 + r.below(9);
        for (int32_t m = 0; m < nm; ++m) {
            deg[m] = 1 + r.below(3);
            for (int32_t k = 0; k < 3; ++k) var[m][k] = r.below(nfeat);
            coef[m] = r.uniform(-1.0, 1.0);
        }
    }
    double eval(const float* x) const {
        double y = 0.0;
        for (int32_t m = 0; m < nm; ++m) {
            double t = coef[m];
            for (int32_t k = 0; k < deg[m]; ++k) t *= (double)x[var[m][k]];
            y += t;
        }
        return y;
    }
};

}  // namespace

extern "C" int64_t ppi_wl_grow(const ppi_wl_grammar* g, uint64_t seed, int32_t n, int32_t stride,
                               int32_t max_depth, int32_t max_symbols,
                               int32_t* phenotype, float* ephemeral, int32_t* size) {
    if (!grammar_ok(g) || n < 0 || stride <= 0 || max_symbols <= 0 || max_symbols > stride) return -1;
    Rng rng(seed);
    int64_t total = 0;
    for (int32_t i = 0; i < n; ++i) {
        Emitter e{g, &rng, phenotype + (size_t)i * stride, ephemeral + (size_t)i * stride, 0, max_symbols, 0, false};
        for (int tries = 0;; ++tries) {
            e.n = 0; e.consts = 0; e.overflow = false;
            e.grow(max_depth);
            if (!e.overflow || tries > 1000) break;
        }
        if (e.overflow) return -1;
        for (int32_t k = e.n; k < stride; ++k) { e.ph[k] = 0; e.eph[k] = 0.0f; }
        size[i] = e.n;
        total += e.n;
    }
    return total;
}

extern "C" int64_t ppi_wl_exact(const ppi_wl_grammar* g, uint64_t seed, int32_t n, int32_t stride,
                                int32_t len_lo, int32_t len_hi,
                                int32_t* phenotype, float* ephemeral, int32_t* size) {
    if (!grammar_ok(g) || n < 0 || len_lo < 1 || len_hi < len_lo || len_hi > stride) return -1;
    if (g->n_unary == 0 && len_lo % 2 == 0 && len_lo == len_hi) return -1;  // binary trees have odd length
    Rng rng(seed);
    int64_t total = 0;
    for (int32_t i = 0; i < n; ++i) {
        int32_t L = len_lo + rng.below(len_hi - len_lo + 1);
        if (g->n_unary == 0 && L % 2 == 0) L = (L + 1 <= len_hi) ? L + 1 : L - 1;
        Emitter e{g, &rng, phenotype + (size_t)i * stride, ephemeral + (size_t)i * stride, 0, stride, 0, false};
        e.exact(L);
        if (e.overflow) return -1;
        for (int32_t k = e.n; k < stride; ++k) { e.ph[k] = 0; e.eph[k] = 0.0f; }
        size[i] = e.n;
        total += e.n;
    }
    return total;
}

extern "C" int32_t ppi_wl_dataset(int32_t kind, uint64_t seed, int32_t nlin, int32_t nfeat, float* out) {
    if (nlin <= 0 || nfeat <= 0 || !out) return -1;
    Rng rng(seed);
    const int32_t ncol = nfeat + 1;
    switch (kind) {
    case PPI_WL_FORMULA_SMALL:
        if (nfeat < 4) return -1;
        for (int32_t r = 0; r < nlin; ++r) {
            float* x = out + (size_t)r * ncol;
            for (int32_t f = 0; f < nfeat; ++f) x[f] = (float)rng.uniform(-1.0, 1.0);
            double y = (double)x[0] * x[0] - 0.5 * (double)x[1] * x[2] + std::sin((double)x[3]);
            x[nfeat] = (float)y;
        }
        return 0;
    case PPI_WL_POLY_NORMAL:
    case PPI_WL_POLY_UNIFORM: {
        Poly p(rng, nfeat);
        for (int32_t r = 0; r < nlin; ++r) {
            float* x = out + (size_t)r * ncol;
            for (int32_t f = 0; f < nfeat; ++f)
                x[f] = (float)(kind == PPI_WL_POLY_NORMAL ? rng.normal() : rng.uniform(-1.0, 1.0));
            double y = p.eval(x);
            if (kind == PPI_WL_POLY_NORMAL) y += 0.01 * rng.normal();
            x[nfeat] = (float)y;
        }
        return 0;
    }
    case PPI_WL_CLUSTERS: {
        const int32_t K = 4;
        std::vector<double> centre((size_t)K * nfeat);
        for (auto& c : centre) c = rng.uniform(-3.0, 3.0);
        for (int32_t r = 0; r < nlin; ++r) {
            float* x = out + (size_t)r * ncol;
            int32_t k = r % K;
            for (int32_t f = 0; f < nfeat; ++f) x[f] = (float)(centre[(size_t)k * nfeat + f] + 0.5 * rng.normal());
            x[nfeat] = (float)k;
        }
        return 0;
    }
    case PPI_WL_FORMULA_ADJOINT:
        if (nfeat < 5) return -1;
        for (int32_t r = 0; r < nlin; ++r) {
            float* x = out + (size_t)r * ncol;
            for (int32_t f = 0; f < nfeat; ++f) x[f] = (float)rng.normal();
            double y = std::exp(0.5 * x[0]) + std::sin(2.0 * x[1]) + (double)x[2] * x[3] * x[4];
            x[nfeat] = (float)y;
        }
        return 0;
    default:
        return -1;
    }
}
