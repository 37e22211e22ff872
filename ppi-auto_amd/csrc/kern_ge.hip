// kern_ge.hip — grammatical evolution on the device (ge.h): random genomes, the breeding plan and
// the offspring, the genome -> phenotype decoder, the phenotype -> tape compiler, and the
// reference's epilogue (alpha*size and the strategy's ranking).
#include "kernels_impl.h"

#include "ge.h"

namespace ppi {
namespace dev {

// ---------------------------------------------------------------------------
// random numbers: XorShift128Plus and BaseRNG::Real (util/Random.h:68-71,130-162)
// ---------------------------------------------------------------------------

__device__ __forceinline__ uint64_t rng_next(uint64_t& s0, uint64_t& s1) {
    uint64_t x = s0;
    const uint64_t y = s1;
    s0 = y;
    x ^= x << 23;
    s1 = x ^ y ^ (x >> 17) ^ (y >> 26);
    return s1 + y;
}

// double(Int()) / 2^64: the 64-bit integer rounded once to double (the high word times 2^32 is
// exact, so the sum is the only rounding), then scaled exactly.
__device__ __forceinline__ double rng_real(uint64_t& s0, uint64_t& s1) {
    const uint64_t v = rng_next(s0, s1);
    const double d = (double)(uint32_t)(v >> 32) * 4294967296.0 + (double)(uint32_t)v;
    return d * 0x1p-64;
}

// OpenMP's static schedule: thread t of T gets iterations [lo, hi) of n.
__device__ __forceinline__ void static_range(int n, int T, int t, int& lo, int& hi) {
    int q = n / T, r = n % T;
    if (t < r) {
        r = 0;
        ++q;
    }
    lo = q * t + r;
    hi = lo + q;
}

// State of stream t advanced by k blocks of draws: the product of the jump tables of k's set bits.
__device__ void jump_blocks(const uint64_t* jump, uint64_t& s0, uint64_t& s1, int64_t k) {
    for (int j = 0; j < kJumpLevels && k; ++j, k >>= 1) {
        if (!(k & 1)) continue;
        const uint64_t* col = jump + (size_t)j * 256;
        uint64_t r0 = 0, r1 = 0;
        for (int b = 0; b < 64; ++b) {
            const uint64_t m = 0 - ((s0 >> b) & 1ull);
            r0 ^= col[2 * b] & m;
            r1 ^= col[2 * b + 1] & m;
        }
        for (int b = 0; b < 64; ++b) {
            const uint64_t m = 0 - ((s1 >> b) & 1ull);
            r0 ^= col[128 + 2 * b] & m;
            r1 ^= col[128 + 2 * b + 1] & m;
        }
        s0 = r0;
        s1 = r1;
    }
}

// OpenMP's static schedule, inverted: the thread and local index of iteration `it` of n over T threads.
__device__ __forceinline__ void static_owner(int n, int T, int it, int& t, int& l) {
    const int q = n / T, r = n % T;
    if (it < r * (q + 1)) {
        t = it / (q + 1);
        l = it % (q + 1);
    } else {
        t = r + (it - r * (q + 1)) / q;
        l = (it - r * (q + 1)) % q;
    }
}

// ppi_generate_population (ppi.cc:656-672): allele j of individual i is random_number() < 0.5, each
// stream drawing its individuals' alleles in order.  One thread per (stream, block of draws); draw d of
// stream t is allele d % nb of its (d / nb)-th individual.  The genomes must be zero.
__global__ void ge_generate_kernel(GeShape S, GeDraws D, uint32_t* genomes) {
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int t = (int)(gid / (D.nblk + 1)), k = (int)(gid % (D.nblk + 1));
    if (t >= S.T) return;
    int lo, hi;
    static_range(S.P, S.T, t, lo, hi);
    const int64_t n = (int64_t)(hi - lo) * S.nb;
    uint64_t s0 = D.rng[2 * t], s1 = D.rng[2 * t + 1];
    jump_blocks(D.jump, s0, s1, k);
    uint64_t* bs = D.bstate + ((size_t)t * (D.nblk + 1) + k) * 2;
    bs[0] = s0;
    bs[1] = s1;
    if (k == 0) D.end[t] = n;
    if (k == D.nblk) return;  // the state after the last block only
    const int64_t d0 = (int64_t)k * kDrawBlock;
    uint32_t word = 0;
    int64_t wi = -1;
    for (int i = 0; i < kDrawBlock && d0 + i < n; ++i) {
        const int64_t d = d0 + i;
        const double r = rng_real(s0, s1);
        const int64_t ind = lo + d / S.nb;
        const int a = (int)(d % S.nb);
        const int64_t w = ind * S.W + (a >> 5);
        if (w != wi) {
            if (wi >= 0 && word) atomicOr(genomes + wi, word);
            wi = w;
            word = 0;
        }
        if (r < 0.5) word |= 1u << (a & 31);
    }
    if (wi >= 0 && word) atomicOr(genomes + wi, word);
}

// Draws of every stream for a breeding step: one thread per (stream, block).
__global__ void ge_draws_kernel(GeDraws D, int T) {
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int t = (int)(gid / (D.nblk + 1)), k = (int)(gid % (D.nblk + 1));
    if (t >= T) return;
    uint64_t s0 = D.rng[2 * t], s1 = D.rng[2 * t + 1];
    jump_blocks(D.jump, s0, s1, k);
    uint64_t* bs = D.bstate + ((size_t)t * (D.nblk + 1) + k) * 2;
    bs[0] = s0;
    bs[1] = s1;
    if (k == 0) D.end[t] = 0;  // a stream without iterations consumes nothing
    if (k == D.nblk) return;
    double* out = D.draw + (size_t)t * D.cap + (size_t)k * kDrawBlock;
    for (int i = 0; i < kDrawBlock; ++i) out[i] = rng_real(s0, s1);
}

// The draws of one loop iteration of ppi_evolve (ppi.cc:896-936), in the reference's order: two
// tournaments (ppi.cc:422-452), the crossover test, the crossover's own draws (ppi.cc:686-724), then
// each child's mutation (ppi.cc:738-808).  Reads draws from v (at most `cap` of them); with `q`
// non-null it fills the plan, reading fitness for the tournaments.  Returns the draws consumed, or -1
// when the iteration would run past the buffer (a start that is not on a stream's chain).
__device__ int64_t iteration(const GeShape& S, const double* v, int64_t p, int64_t cap, int nchild, const float* fit,
                             GePlan* q, int32_t* flips) {
    int64_t k = p;
    auto next = [&](double& x) -> bool {
        if (k >= cap) return false;
        x = v[k++];
        return true;
    };
    // (int)(random_number() * n), a draw that rounds to n (Int() within 2^10 of 2^64) clamped to n - 1
    auto index = [&](double x, int n) {
        const int i = (int)(x * n);
        return i < n ? i : n - 1;
    };
    double x;
    for (int parent = 0; parent < 2; ++parent) {
        if (!next(x)) return -1;
        int w = index(x, S.P);
        float fw = q ? fit[w] : 0.0f;
        for (int t = 1; t < S.ts; ++t) {
            if (!next(x)) return -1;
            const int c = index(x, S.P);
            if (q) {
                const float fc = fit[c];
                if (fc < fw) {
                    fw = fc;
                    w = c;
                }
            }
        }
        if (q) {
            if (parent == 0) q->father = w;
            else q->mother = w;
        }
    }
    int xo = 0, a = 0, b = S.nb;
    if (!next(x)) return -1;
    if (x < (double)S.cp) {
        if (!next(x)) return -1;
        if (x < 0.66) {
            double y, z;
            if (!next(y) || !next(z)) return -1;
            a = ((int)(y * S.nb)) / S.bg * S.bg;
            b = ((int)(z * S.nb)) / S.bg * S.bg;
            if (a > b) {
                const int tmp = a;
                a = b;
                b = tmp;
            }
            xo = 2;
        } else {
            if (!next(x)) return -1;
            a = (int)(x * S.nb);
            xo = 1;
        }
    }
    if (q) {
        q->xo = xo;
        q->a = a;
        q->b = b;
    }
    for (int c = 0; c < 2; ++c) {
        if (q) {
            q->mut[c] = 0;
            q->nflip[c] = 0;
            q->s0[c] = q->s1[c] = 0;
        }
        if (c >= nchild) continue;
        if (!next(x)) return -1;
        int num = (int)(x * (S.max_bits + 1));
        if (num > S.max_bits) num = S.max_bits;  // a draw that rounds to 1.0
        if (num == 0) continue;
        if (!next(x)) return -1;
        if (x < 0.75) {
            if (k + num > cap) return -1;
            if (q) {
                q->mut[c] = 1;
                q->nflip[c] = num;
                for (int j = 0; j < num; ++j) flips[(size_t)c * S.max_bits + j] = index(v[k + j], S.nb);
            }
            k += num;
        } else {
            if (!next(x)) return -1;
            if (x < 0.10) {
                if (!next(x)) return -1;
                num = (int)(x * S.nb);
            }
            if (!next(x)) return -1;
            const int shrink = ((num + (S.bg - 1)) / S.bg) * S.bg;
            const int start = ((int)(x * S.nb)) / S.bg * S.bg;
            if (q) {
                q->mut[c] = 2;
                q->s0[c] = start;
                q->s1[c] = min(start + shrink, S.nb);
            }
        }
    }
    return k - p;
}

// nxt_0[t][p] = the start of the iteration after one that starts at draw p (two children), or cap.
__global__ void ge_len_kernel(GeShape S, GeDraws D) {
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t stride = D.cap + 1;
    const int t = (int)(gid / stride);
    const int64_t p = gid % stride;
    if (t >= S.T) return;
    int64_t n = D.cap;
    if (p < D.cap) {
        const int64_t len = iteration(S, D.draw + (size_t)t * D.cap, p, D.cap, 2, nullptr, nullptr, nullptr);
        if (len >= 0) n = p + len;
    }
    D.nxt[(size_t)t * stride + p] = (int32_t)n;
}

// nxt_j = nxt_{j-1} applied twice.
__global__ void ge_double_kernel(GeShape S, GeDraws D, int j) {
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t stride = D.cap + 1;
    if (gid >= (int64_t)S.T * stride) return;
    const int t = (int)(gid / stride);
    const int32_t* prev = D.nxt + ((size_t)(j - 1) * S.T + t) * stride;
    int32_t* cur = D.nxt + ((size_t)j * S.T + t) * stride;
    cur[gid % stride] = prev[prev[gid % stride]];
}

// One thread per iteration of the breeding loop: its stream's chain gives its first draw, then the plan.
__global__ void ge_plan_kernel(GeShape S, GeDraws D, const float* fit, int32_t nimm, GePlan* plan, int32_t* flips) {
    const int n_it = (S.P - nimm + 1) / 2;
    const int it = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (it >= n_it) return;
    int t, l;
    static_owner(n_it, S.T, it, t, l);
    const int64_t stride = D.cap + 1;
    int64_t p = 0;
    for (int j = 0, ll = l; ll; ++j, ll >>= 1)
        if (ll & 1) p = D.nxt[((size_t)j * S.T + t) * stride + p];
    const int i = nimm + 2 * it;
    GePlan q;
    q.father = q.mother = 0;  // (every iteration on a stream's chain fits its buffer and fills the plan)
    q.xo = 0;
    q.a = 0;
    q.b = S.nb;
    q.mut[0] = q.mut[1] = 0;
    q.nflip[0] = q.nflip[1] = 0;
    q.s0[0] = q.s0[1] = q.s1[0] = q.s1[1] = 0;
    const int64_t len = iteration(S, D.draw + (size_t)t * D.cap, p, D.cap, i < S.P - 1 ? 2 : 1, fit, &q,
                                  flips + (size_t)it * 2 * S.max_bits);
    plan[it] = q;
    int lo, hi;
    static_range(n_it, S.T, t, lo, hi);
    if (it == hi - 1) D.end[t] = p + (len > 0 ? len : 0);
}

// Each stream's state after the draws its phase consumed: its block's start state, stepped.
__global__ void ge_state_kernel(GeDraws D, int T) {
    const int t = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (t >= T) return;
    const int64_t e = D.end[t];
    const int64_t k = e / kDrawBlock;
    const uint64_t* bs = D.bstate + ((size_t)t * (D.nblk + 1) + k) * 2;
    uint64_t s0 = bs[0], s1 = bs[1];
    for (int64_t i = 0; i < e % kDrawBlock; ++i) (void)rng_next(s0, s1);
    D.rng[2 * t] = s0;
    D.rng[2 * t + 1] = s1;
}

__device__ __forceinline__ uint32_t gbit(const uint32_t* g, int k) { return (g[k >> 5] >> (k & 31)) & 1u; }

// Offspring words: crossover (or clone) of the planned parents, then the child's mutation: shrink
// moves the alleles after a gene-aligned gap left (memmove, the tail keeps its old alleles), bit flips
// toggle (a flip drawn twice cancels).  The last individual of an odd remainder is the second
// offspring's pattern when crossed (both offspring were written to it), its father otherwise.
__global__ void ge_offspring_kernel(GeShape S, const GePlan* plan, const int32_t* flips, const uint32_t* parents,
                                    uint32_t* children, int32_t nimm) {
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= (int64_t)(S.P - nimm) * S.W) return;
    const int j = nimm + (int)(gid / S.W), w = (int)(gid % S.W);
    const int it = (j - nimm) / 2, c = (j - nimm) % 2;
    const int i = nimm + 2 * it;
    const GePlan q = plan[it];
    const uint32_t* fa = parents + (size_t)q.father * S.W;
    const uint32_t* mo = parents + (size_t)q.mother * S.W;
    const int pattern = i == S.P - 1 ? 1 : c;
    const int mut = q.mut[c];
    const int d = q.s1[c] - q.s0[c];
    uint32_t word = 0;
    const int nbits = min(32, S.nb - 32 * w);
    for (int bb = 0; bb < nbits; ++bb) {
        int k = 32 * w + bb;
        if (mut == 2 && k >= q.s0[c] && k < S.nb - d) k += d;
        uint32_t v;
        if (q.xo == 0) {
            v = gbit(c == 0 ? fa : mo, k);
        } else {
            const bool seg = q.xo == 1 ? k >= q.a : (k >= q.a && k < q.b);
            const bool from_father = pattern == 0 ? !seg : seg;
            v = gbit(from_father ? fa : mo, k);
        }
        word |= v << bb;
    }
    if (mut == 1) {
        const int32_t* f = flips + ((size_t)it * 2 + c) * S.max_bits;
        for (int k = 0; k < q.nflip[c]; ++k)
            if ((f[k] >> 5) == w) word ^= 1u << (f[k] & 31);
    }
    children[(size_t)j * S.W + w] = word;
}

// ---------------------------------------------------------------------------
// decode (ppi.cc:106-182), one thread per individual, the recursion as an explicit stack
// ---------------------------------------------------------------------------

// float(raw) rounded to nearest even, for raw up to 2^63.
__device__ __forceinline__ float u64_to_float_rn(uint64_t x) {
    if (x < (1ull << 24)) return (float)(uint32_t)x;
    const int lz = __clzll((long long)x);
    const int shift = 64 - lz - 24;
    uint64_t mant = x >> shift;
    const uint64_t rem = x & ((1ull << shift) - 1), half = 1ull << (shift - 1);
    if (rem > half || (rem == half && (mant & 1))) ++mant;
    return ldexpf((float)(uint32_t)mant, shift);
}

__global__ void ge_decode_kernel(GeShape S, GeDevGrammar G, const uint32_t* genomes, int32_t* ph, float* eph,
                                 int32_t* size, int32_t* frames, int64_t fstride) {
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= S.P) return;
    const uint32_t* g = genomes + (size_t)i * S.W;
    int32_t* p = ph + (size_t)i * S.max_size;
    float* e = eph + (size_t)i * S.max_size;
    int32_t* fr = frames + (size_t)i * fstride;
    int allele = 0, pos = 0, depth = 0;
    // decode_rule: bg alleles, least significant first, select rule raw % tamanhos[head]
    auto rule = [&](int head) -> int {
        if (allele + S.bg > S.nb) return -1;
        uint32_t raw = 0;
        for (int b = 0; b < S.bg; ++b)
            if (gbit(g, allele++)) raw += 1u << b;
        return G.first_rule[head] + (int)(raw % (uint32_t)G.n_rules[head]);
    };
    int r = rule(0);
    if (r < 0 || pos >= S.max_size) {
        size[i] = 0;
        return;
    }
    fr[depth++] = r << 8;
    bool ok = true;
    while (depth > 0) {
        const int32_t f = fr[depth - 1];
        r = f >> 8;
        const int k = f & 255;
        if (k == G.rule_len[r]) {
            --depth;
            if (depth > 0 && pos == 0) {  // `if (!pos) return 0` after a recursive decode
                ok = false;
                break;
            }
            continue;
        }
        fr[depth - 1] = (r << 8) | (k + 1);
        const int32_t s = G.rule_sym[(size_t)r * G.max_rule + k];
        if (s >= 10000) {
            if (pos >= S.max_size) {  // undefined in the reference: a kill, like the rule check
                ok = false;
                break;
            }
            p[pos] = s;
            e[pos] = 0.0f;
            if (G.t_const >= 0 && s == G.t_const) {
                // decode_real (ppi.cc:124-137): bc alleles, or 0 without consuming any
                if (allele + S.bc <= S.nb) {
                    uint64_t raw = 0;
                    for (int b = 0; b < S.bc; ++b)
                        if (gbit(g, allele++)) raw += 1ull << b;
                    const float prod = u64_to_float_rn(raw) * (S.hi - S.lo);
                    e[pos] = (float)((double)S.lo + (double)prod / ((double)(1ull << S.bc) - 1.0));
                }
            } else if (s >= 20000) {
                p[pos] = G.t_attribute;
                e[pos] = (float)(s - 20000);
            }
            ++pos;
        } else {
            const int r2 = rule(s);
            if (r2 < 0 || pos >= S.max_size || depth >= fstride) {
                ok = false;
                break;
            }
            fr[depth++] = r2 << 8;
        }
    }
    size[i] = ok ? pos : 0;
}

// ---------------------------------------------------------------------------
// the tape compiler (tape.cc: Compiler::compile and emit), one thread per program
// ---------------------------------------------------------------------------
//
// Nodes: 4 words each — arg; ch0 | ch1 << 16; ch2 | fold << 16; op (int8) | regs << 8 | leaf_attr << 16.
// The emitted instructions are those of the host compiler, in its order; a constant subtree's value
// is evaluated here, by the interpreters' own operator code, and written where the host compiler
// leaves the placeholder that its fold programs fill.

struct CNode {
    uint32_t arg;
    int ch0, ch1, ch2, fold;
    int op, regs, leaf_attr;
};

__device__ __forceinline__ CNode node_get(const uint32_t* nd, int k) {
    const uint4 v = reinterpret_cast<const uint4*>(nd)[k];
    CNode n;
    n.arg = v.x;
    n.ch0 = (int)(v.y & 0xffffu);
    n.ch1 = (int)(v.y >> 16);
    n.ch2 = (int)(v.z & 0xffffu);
    n.fold = (int)(v.z >> 16);
    n.op = (int)(int8_t)(v.w & 0xffu);
    n.regs = (int)((v.w >> 8) & 0xffu);
    n.leaf_attr = (int)((v.w >> 16) & 1u);
    return n;
}

__device__ __forceinline__ void node_put(uint32_t* nd, int k, const CNode& n) {
    uint4 v;
    v.x = n.arg;
    v.y = (uint32_t)n.ch0 | ((uint32_t)n.ch1 << 16);
    v.z = (uint32_t)n.ch2 | ((uint32_t)n.fold << 16);
    v.w = ((uint32_t)n.op & 0xffu) | ((uint32_t)n.regs << 8) | ((uint32_t)n.leaf_attr << 16);
    reinterpret_cast<uint4*>(nd)[k] = v;
}

__device__ __forceinline__ int child_of(const CNode& n, int k) { return k == 0 ? n.ch0 : (k == 1 ? n.ch1 : n.ch2); }

struct CompileArgs {
    const int32_t* ph;
    const float* eph;
    const int32_t* size;
    int32_t stride, n;
    const int32_t* sym_op;   // symbol - 10000 -> engine opcode, -1 = not in the table (8192 entries)
    const float* sym_val;
    int32_t ncol;
    int32_t fold;
    GeCompileScratch X;
    GeCompiled* out;
    GeCompileTotals* tot;
};

// Value of the constant subtree rooted at node `root` (its leaves are constants).
template <bool P, bool N>
__device__ float fold_value(const uint32_t* nd, int root, int32_t* fst, float* fv) {
    int sp = 0, vp = 0;
    fst[sp++] = root << 2;
    while (sp > 0) {
        const int32_t e = fst[sp - 1];
        const int x = e >> 2, st = e & 3;
        const CNode n = node_get(nd, x);
        if (n.op < 0) {
            fv[vp++] = __uint_as_float(n.arg);
            --sp;
            continue;
        }
        const int ar = arity(n.op);
        if (st < ar) {
            fst[sp - 1] = (x << 2) | (st + 1);
            fst[sp++] = child_of(n, st) << 2;
            continue;
        }
        float v;
        if (ar == 1) {
            float a[1] = {fv[vp - 1]};
            unary_rows<1, P, N>(unop_of(n.op), a);
            v = a[0];
            vp -= 1;
        } else if (ar == 2) {
            v = binary<P, N>(binop_of(n.op, false), fv[vp - 2], fv[vp - 1]);
            vp -= 2;
        } else {
            v = (fv[vp - 3] != 0.0f) ? fv[vp - 2] : fv[vp - 1];
            vp -= 3;
        }
        fv[vp++] = v;
        --sp;
    }
    return fv[0];
}

template <bool P, bool N>
__global__ void __launch_bounds__(256) ge_compile_kernel(CompileArgs A) {
    stage_math<N>();
    __syncthreads();
    const int p = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (p >= A.n) return;
    const GeCompileScratch& X = A.X;
    uint32_t* nd = X.node + (size_t)p * X.node_stride;
    int32_t* stk = X.stk + (size_t)p * X.stk_stride;
    float* fv = X.fval + (size_t)p * X.fval_stride;
    Ins* code = X.stage + (size_t)p * X.stage_stride;
    const int32_t* ph = A.ph + (size_t)p * A.stride;
    const float* eph = A.eph + (size_t)p * A.stride;
    const int sz = A.size[p];
    GeCompiled res;
    res.max_col = -1;
    res.end_arg = 0;
    if (sz < 0 || sz > A.stride) {
        atomicOr(&A.tot->error, 2);
        res.ncode = 0;
        res.flags = kEmpty;
        res.end_arg = kEmpty;
        A.out[p] = res;
        return;
    }
    if (sz == 0) {
        res.ncode = 0;
        res.flags = kEmpty;
        res.end_arg = kEmpty;
        A.out[p] = res;
        return;
    }
    const uint32_t kNaN = 0x7fc00000u;
    // Scan backwards (sequential.cc:143-164 order), building the expression tree.
    int nn = 0, vs = 0;
    int32_t* vstk = stk;  // the value stack of node indices, dead once the root is known
    bool valid = true;
    for (int i = sz - 1; i >= 0; --i) {
        const int32_t sym = ph[i];
        const int idx = sym - 10000;
        int op = (idx >= 0 && idx < 8192) ? A.sym_op[idx] : -1;
        float value = (idx >= 0 && idx < 8192) ? A.sym_val[idx] : 0.0f;
        if (op < 0) op = PPI_OP_UNKNOWN;
        const int ar = arity(op);
        CNode n;
        n.ch0 = n.ch1 = n.ch2 = 0;
        n.fold = 0;
        n.regs = 0;
        n.leaf_attr = 0;
        n.arg = 0;
        if (ar == 0) {
            n.op = -1;
            if (op == PPI_OP_CONST) {
                n.arg = __float_as_uint(eph[i]);
            } else if (op == PPI_OP_ATTRIBUTE) {
                const float c = eph[i];
                if (c > -1.0f && c < (float)A.ncol) {
                    n.leaf_attr = 1;
                    n.arg = (uint32_t)(int32_t)c;
                } else {
                    n.arg = kNaN;
                }
            } else if (op == PPI_OP_VALUE) {
                n.arg = __float_as_uint(value);
            } else {
                n.arg = kNaN;
            }
            node_put(nd, nn, n);
            vstk[vs++] = nn++;
            continue;
        }
        if (vs < ar) {
            valid = false;
            break;
        }
        n.op = op;
        n.ch0 = vstk[--vs];
        if (ar >= 2) n.ch1 = vstk[--vs];
        if (ar >= 3) n.ch2 = vstk[--vs];
        bool constant = A.fold != 0;
        for (int k = 0; k < ar && constant; ++k) {
            const CNode c = node_get(nd, child_of(n, k));
            constant = c.op < 0 && !c.leaf_attr;
        }
        if (constant) {
            // a folded child contributes its whole subtree
            for (int k = 0; k < ar; ++k) {
                const CNode c = node_get(nd, child_of(n, k));
                if (c.fold) {
                    if (k == 0) n.ch0 = c.fold - 1;
                    else if (k == 1) n.ch1 = c.fold - 1;
                    else n.ch2 = c.fold - 1;
                }
            }
        }
        // Sethi-Ullman register need
        if (ar == 1) {
            n.regs = node_get(nd, n.ch0).regs;
        } else if (ar == 2) {
            const CNode a = node_get(nd, n.ch0), b = node_get(nd, n.ch1);
            const bool la = a.op < 0, lb = b.op < 0;
            if (la && lb) n.regs = 0;
            else if (la) n.regs = b.regs;
            else if (lb) n.regs = a.regs;
            else n.regs = a.regs == b.regs ? a.regs + 1 : max(a.regs, b.regs);
        } else {
            int r0 = node_get(nd, n.ch0).regs, r1 = node_get(nd, n.ch1).regs, r2 = node_get(nd, n.ch2).regs;
            // descending order
            if (r0 < r1) { const int t = r0; r0 = r1; r1 = t; }
            if (r1 < r2) { const int t = r1; r1 = r2; r2 = t; }
            if (r0 < r1) { const int t = r0; r0 = r1; r1 = t; }
            n.regs = max(r0, max(1 + r1, 2 + r2));
        }
        node_put(nd, nn, n);
        if (constant) {
            CNode leaf;
            leaf.arg = 0;
            leaf.ch0 = leaf.ch1 = leaf.ch2 = 0;
            leaf.fold = nn + 1;
            leaf.op = -1;
            leaf.regs = 0;
            leaf.leaf_attr = 0;
            node_put(nd, nn + 1, leaf);
            vstk[vs++] = nn + 1;
            nn += 2;
            continue;
        }
        vstk[vs++] = nn++;
    }
    if (!valid || vs == 0) {
        // underflow (undefined in the reference): evaluates to NaN
        code[0] = Ins{H_LOAD_C, kNaN};
        res.ncode = 1;
        res.flags = kInvalid;
        A.out[p] = res;
        return;
    }
    const int root = vstk[vs - 1];
    // emit (tape.cc: Compiler::emit); frames: [node | base << 16, phase | o0 << 8 | o1 << 16 | o2 << 24]
    int32_t* fr = stk;  // overwrites the value stack, which is no longer needed
    int32_t* fst = stk + X.stk_stride / 2;  // fold-evaluation frames
    uint32_t nc = 0;
    int pending = -1, max_slot = 0, maxcol = -1;
    int nf = 0;
    auto put = [&](uint32_t c, uint32_t a) {
        code[nc++] = Ins{c, a};
        const uint32_t h = c & 0xffu;
        if ((h == H_LOAD_A || h == H_PUSH_A || (h >= H_BIN_A && h < H_BIN_S)) && (int32_t)a < A.ncol)
            maxcol = max(maxcol, (int)a);
    };
    auto leaf_arg = [&](const CNode& l) -> uint32_t {
        if (!l.fold) return l.arg;
        return __float_as_uint(fold_value<P, N>(nd, l.fold - 1, fst, fv));
    };
    auto emit_load = [&](const CNode& l) {
        const uint32_t a = leaf_arg(l);
        if (pending >= 0) {
            put((l.leaf_attr ? H_PUSH_A : H_PUSH_C) | ((uint32_t)pending << 8), a);
            max_slot = max(max_slot, pending + 1);
            pending = -1;
        } else {
            put(l.leaf_attr ? H_LOAD_A : H_LOAD_C, a);
        }
    };
    auto emit_fold = [&](uint32_t bop, const CNode& l) { put((l.leaf_attr ? H_BIN_A : H_BIN_C) + bop, leaf_arg(l)); };
    auto push = [&](int node, int base) {
        fr[2 * nf] = node | (base << 16);
        fr[2 * nf + 1] = 0;
        ++nf;
    };
    push(root, 0);
    while (nf > 0) {
        const int node = fr[2 * nf - 2] & 0xffff, base = fr[2 * nf - 2] >> 16;
        const int32_t ph_w = fr[2 * nf - 1];
        const int phase = ph_w & 0xff;
        const CNode n = node_get(nd, node);
        if (n.op < 0) {
            emit_load(n);
            --nf;
            continue;
        }
        const int ar = arity(n.op);
        if (ar == 1) {
            if (phase == 0) {
                fr[2 * nf - 1] = 1;
                push(n.ch0, base);
            } else {
                put(H_UN + unop_of(n.op), 0);
                --nf;
            }
            continue;
        }
        if (ar == 2) {
            const CNode L = node_get(nd, n.ch0), R = node_get(nd, n.ch1);
            const bool ll = L.op < 0, rl = R.op < 0;
            if (ll && rl) {
                emit_load(L);
                emit_fold(binop_of(n.op, false), R);
                --nf;
            } else if (ll || rl) {
                if (phase == 0) {
                    fr[2 * nf - 1] = 1;
                    push(ll ? n.ch1 : n.ch0, base);
                } else {
                    emit_fold(binop_of(n.op, ll), ll ? L : R);
                    --nf;
                }
            } else {
                const bool left_first = L.regs >= R.regs;
                if (phase == 0) {
                    fr[2 * nf - 1] = 1;
                    push(left_first ? n.ch0 : n.ch1, base);
                } else if (phase == 1) {
                    fr[2 * nf - 1] = 2;
                    pending = base;
                    push(left_first ? n.ch1 : n.ch0, base + 1);
                } else {
                    put(H_BIN_S + binop_of(n.op, left_first) + ((uint32_t)base << 8), 0);
                    --nf;
                }
            }
            continue;
        }
        // IF_THEN_ELSE(cond = ch0, then = ch1, else = ch2)
        if (phase == 0) {
            const int rg[3] = {node_get(nd, n.ch0).regs, node_get(nd, n.ch1).regs, node_get(nd, n.ch2).regs};
            int o[3] = {0, 1, 2};
            // stable sort by register need, descending
            for (int a = 1; a < 3; ++a) {
                const int v = o[a];
                int b = a - 1;
                while (b >= 0 && rg[o[b]] < rg[v]) {
                    o[b + 1] = o[b];
                    --b;
                }
                o[b + 1] = v;
            }
            fr[2 * nf - 1] = 1 | (o[0] << 8) | (o[1] << 16) | (o[2] << 24);
            push(child_of(n, o[0]), base);
        } else {
            const int o0 = (ph_w >> 8) & 0xff, o1 = (ph_w >> 16) & 0xff, o2 = (ph_w >> 24) & 0xff;
            if (phase == 1) {
                fr[2 * nf - 1] = (ph_w & ~0xff) | 2;
                pending = base;
                push(child_of(n, o1), base + 1);
            } else if (phase == 2) {
                fr[2 * nf - 1] = (ph_w & ~0xff) | 3;
                pending = base + 1;
                push(child_of(n, o2), base + 2);
            } else {
                uint32_t where[3];
                where[o0] = 1;
                where[o1] = 2;
                where[o2] = 0;
                const uint32_t c = H_TERN | ((uint32_t)base << 8) | ((uint32_t)(base + 1) << 13) | (where[0] << 18) |
                                   (where[1] << 20) | (where[2] << 22);
                put(c, 0);
                --nf;
            }
        }
    }
    if (max_slot > kMaxSlots) atomicOr(&A.tot->error, 1);
    res.ncode = nc;
    res.flags = (uint32_t)max_slot << 8;
    res.max_col = maxcol;
    A.out[p] = res;
}

// Totals, then the packed tape.  Scan of (ncode + 1) in one workgroup of 1024 threads.
__global__ void __launch_bounds__(1024) ge_pack_offsets_kernel(const GeCompiled* comp, const int32_t* size, int32_t n,
                                                               uint32_t* offs, GeCompileTotals* tot) {
    __shared__ uint32_t part[1024];
    __shared__ int64_t sz_part[1024];
    __shared__ int slot_part[1024], col_part[1024];
    const int t = (int)threadIdx.x;
    const int per = (n + 1023) / 1024;
    const int lo = min(n, t * per), hi = min(n, lo + per);
    uint32_t s = 0;
    int64_t zs = 0;
    int ms = 0, mc = -1;
    for (int i = lo; i < hi; ++i) {
        const GeCompiled c = comp[i];
        s += c.ncode + 1;
        zs += size[i] > 0 ? size[i] : 0;
        ms = max(ms, (int)((c.flags >> 8) & 0xffu));
        mc = max(mc, c.max_col);
    }
    part[t] = s;
    sz_part[t] = zs;
    slot_part[t] = ms;
    col_part[t] = mc;
    __syncthreads();
    if (t == 0) {
        uint32_t acc = 0;
        int64_t zacc = 0;
        int m1 = 0, m2 = -1;
        for (int k = 0; k < 1024; ++k) {
            const uint32_t v = part[k];
            part[k] = acc;
            acc += v;
            zacc += sz_part[k];
            m1 = max(m1, slot_part[k]);
            m2 = max(m2, col_part[k]);
        }
        offs[n] = acc;
        tot->total_size = zacc;
        tot->total_code = (int64_t)acc - n;
        tot->max_slots = m1;
        tot->max_col = m2;
    }
    __syncthreads();
    uint32_t o = part[t];
    for (int i = lo; i < hi; ++i) {
        offs[i] = o;
        o += comp[i].ncode + 1;
    }
}

__global__ void ge_pack_kernel(GeCompileScratch X, const GeCompiled* comp, const int32_t* size, int32_t n,
                               const uint32_t* offs, ProgMeta* meta, Ins* code) {
    const int p = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (p < n) {
        const GeCompiled c = comp[p];
        const uint32_t off = offs[p];
        ProgMeta m;
        m.code_off = off;
        m.ncode = c.ncode;
        m.size = size[p];
        m.flags = c.flags;
        meta[p] = m;
        const Ins* src = X.stage + (size_t)p * X.stage_stride;
        for (uint32_t k = 0; k < c.ncode; ++k) code[off + k] = src[k];
        code[off + c.ncode] = Ins{H_END, c.end_arg};
    }
    if (p < kTapePad) code[offs[n] + p] = Ins{H_END, kEmpty};
}

__global__ void ge_track_best_kernel(const float* fitness, const int32_t* best, const uint32_t* genomes, int32_t W,
                                     float* best_fit, uint32_t* best_genome) {
    const int b = best[0];
    const float f = fitness[b];
    if (!(f < best_fit[0])) return;  // uniform over the workgroup
    __syncthreads();
    for (int w = (int)threadIdx.x; w < W; w += (int)blockDim.x) best_genome[w] = genomes[(size_t)b * W + w];
    __syncthreads();
    if (threadIdx.x == 0) best_fit[0] = f;
}

// ---------------------------------------------------------------------------
// the reference's epilogue on the device
// ---------------------------------------------------------------------------

// vec[i]: PP / PDP add alpha*size to the kernel's value; SEQ / FP keep FLT_MAX and empty programs'
// values and add alpha*size otherwise (accelerator.cc:801,849; sequential.cc:185-189).
__global__ void rank_vec_kernel(const float* raw, const ProgMeta* meta, int32_t n, int32_t strategy, float alpha,
                                float* vec) {
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n) return;
    const float r = raw[i];
    const ProgMeta m = meta[i];
    const float as = alpha * (float)m.size;
    float v;
    if (strategy == PPI_STRATEGY_PP || strategy == PPI_STRATEGY_PDP) v = r + as;
    else v = (r == FLT_MAX || (m.flags & kEmpty)) ? r : r + as;
    vec[i] = v;
}

// best_individual (accelerator.cl:270-313): work-group g of `local2` entries; the tree keeps the lower
// value, the earlier position on ties; entries past n are (FLT_MAX, -1).
__global__ void __launch_bounds__(1024) rank_groups_kernel(const float* v, int32_t n, int32_t local2, float* gval,
                                                           int32_t* gidx) {
    __shared__ float lb[1024];
    __shared__ int32_t li[1024];
    const int l = (int)threadIdx.x;
    const int g = (int)blockIdx.x;
    const int gl = g * local2 + l;
    if (l < local2) {
        lb[l] = gl < n ? v[gl] : FLT_MAX;
        li[l] = gl < n ? gl : -1;
    }
    __syncthreads();
    int np2 = 1;
    while (np2 < local2) np2 <<= 1;
    for (int s = np2 / 2; s > 0; s >>= 1) {
        if (l < s && l + s < local2 && lb[l] > lb[l + s]) {
            lb[l] = lb[l + s];
            li[l] = li[l + s];
        }
        __syncthreads();
    }
    if (l == 0) {
        gval[g] = lb[0];
        gidx[g] = li[0];
    }
}

// util::PickNBest (Util.h:158-173) over m candidates: the k lowest values, on equal values the larger
// position first; index[j] = idx ? idx[position] : position.  One workgroup, k rounds of arg-min.
__global__ void __launch_bounds__(1024) pick_n_best_kernel(const float* val, const int32_t* idx, int32_t m, int32_t k,
                                                           int32_t* index) {
    __shared__ float bv[1024];
    __shared__ int32_t bp[1024];
    __shared__ int32_t chosen[kRankMax];
    const int t = (int)threadIdx.x;
    for (int round = 0; round < k; ++round) {
        float best = 0.0f;
        int pos = -1;
        for (int i = t; i < m; i += (int)blockDim.x) {
            bool taken = false;
            for (int c = 0; c < round; ++c) taken |= chosen[c] == i;
            if (taken) continue;
            const float x = val[i];
            if (pos < 0 || x < best || (!(best < x) && i > pos)) {
                best = x;
                pos = i;
            }
        }
        bv[t] = best;
        bp[t] = pos;
        __syncthreads();
        for (int s = 512; s > 0; s >>= 1) {
            if (t < s) {
                const int po = bp[t + s];
                const float xo = bv[t + s];
                const int pm = bp[t];
                const float xm = bv[t];
                if (po >= 0 && (pm < 0 || xo < xm || (!(xm < xo) && po > pm))) {
                    bv[t] = xo;
                    bp[t] = po;
                }
            }
            __syncthreads();
        }
        if (t == 0) {
            chosen[round] = bp[0];
            index[round] = bp[0] < 0 ? -1 : (idx ? idx[bp[0]] : bp[0]);
        }
        __syncthreads();
    }
}

}  // namespace dev

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------

hipError_t ge_launch_generate(const GeShape& S, const GeDraws& D, uint32_t* genomes, hipStream_t st) {
    const int64_t n = (int64_t)S.T * (D.nblk + 1);
    hipLaunchKernelGGL(dev::ge_generate_kernel, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, st, S, D, genomes);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(dev::ge_state_kernel, dim3((unsigned)((S.T + 63) / 64)), dim3(64), 0, st, D, S.T);
    return hipGetLastError();
}

hipError_t ge_launch_plan(const GeShape& S, const GeDraws& D, const float* fitness, int32_t nimm, GePlan* plan,
                          int32_t* flips, hipStream_t st) {
    const int n_it = (S.P - nimm + 1) / 2;
    const int64_t nb = (int64_t)S.T * (D.nblk + 1);
    hipLaunchKernelGGL(dev::ge_draws_kernel, dim3((unsigned)((nb + 63) / 64)), dim3(64), 0, st, D, S.T);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const int64_t np = (int64_t)S.T * (D.cap + 1);
    hipLaunchKernelGGL(dev::ge_len_kernel, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, st, S, D);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    for (int j = 1; j < D.levels; ++j) {
        hipLaunchKernelGGL(dev::ge_double_kernel, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, st, S, D, j);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    hipLaunchKernelGGL(dev::ge_plan_kernel, dim3((unsigned)((n_it + 255) / 256)), dim3(256), 0, st, S, D, fitness,
                       nimm, plan, flips);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(dev::ge_state_kernel, dim3((unsigned)((S.T + 63) / 64)), dim3(64), 0, st, D, S.T);
    return hipGetLastError();
}

hipError_t ge_launch_offspring(const GeShape& S, const GePlan* plan, const int32_t* flips, const uint32_t* parents,
                               uint32_t* children, int32_t nimm, hipStream_t st) {
    const int64_t n = (int64_t)(S.P - nimm) * S.W;
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(dev::ge_offspring_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, S, plan, flips,
                       parents, children, nimm);
    return hipGetLastError();
}

hipError_t ge_launch_decode(const GeShape& S, const GeDevGrammar& G, const uint32_t* genomes, int32_t* ph,
                            float* eph, int32_t* size, int32_t* frames, int64_t frame_stride, hipStream_t st) {
    hipLaunchKernelGGL(dev::ge_decode_kernel, dim3((unsigned)((S.P + 127) / 128)), dim3(128), 0, st, S, G, genomes,
                       ph, eph, size, frames, frame_stride);
    return hipGetLastError();
}

hipError_t ge_launch_compile(const int32_t* ph, const float* eph, const int32_t* size, int32_t stride, int32_t n,
                             const int32_t* sym_op, const float* sym_val, int32_t ncol, bool fold, bool P, bool N,
                             const GeCompileScratch& X, GeCompiled* out, GeCompileTotals* tot, hipStream_t st) {
    dev::CompileArgs A{ph, eph, size, stride, n, sym_op, sym_val, ncol, fold ? 1 : 0, X, out, tot};
    dim3 grid((unsigned)((n + 255) / 256));
    if (P && N) hipLaunchKernelGGL((dev::ge_compile_kernel<true, true>), grid, dim3(256), 0, st, A);
    else if (P) hipLaunchKernelGGL((dev::ge_compile_kernel<true, false>), grid, dim3(256), 0, st, A);
    else if (N) hipLaunchKernelGGL((dev::ge_compile_kernel<false, true>), grid, dim3(256), 0, st, A);
    else hipLaunchKernelGGL((dev::ge_compile_kernel<false, false>), grid, dim3(256), 0, st, A);
    return hipGetLastError();
}

hipError_t ge_launch_pack(const GeCompileScratch& X, const GeCompiled* comp, const int32_t* size, int32_t n,
                          uint32_t* offs, ProgMeta* meta, Ins* code, GeCompileTotals* tot, hipStream_t st) {
    hipLaunchKernelGGL(dev::ge_pack_offsets_kernel, dim3(1), dim3(1024), 0, st, comp, size, n, offs, tot);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const int threads = n > kTapePad ? n : kTapePad;
    hipLaunchKernelGGL(dev::ge_pack_kernel, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, st, X, comp, size,
                       n, offs, meta, code);
    return hipGetLastError();
}

hipError_t ge_launch_track_best(const float* fitness, const int32_t* best, const uint32_t* genomes, int32_t W,
                                float* best_fit, uint32_t* best_genome, hipStream_t st) {
    hipLaunchKernelGGL(dev::ge_track_best_kernel, dim3(1), dim3(256), 0, st, fitness, best, genomes, W, best_fit,
                       best_genome);
    return hipGetLastError();
}

hipError_t launch_rank(const float* raw, const ProgMeta* meta, int32_t n, int32_t strategy, float alpha,
                       int32_t best_local, int32_t k, float* vec, float* group_val, int32_t* group_idx, int32_t* index,
                       hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(dev::rank_vec_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, raw, meta, n,
                       strategy, alpha, vec);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || k <= 0) return e;
    if (strategy == PPI_STRATEGY_SEQ) {
        hipLaunchKernelGGL(dev::pick_n_best_kernel, dim3(1), dim3(1024), 0, st, vec, (const int32_t*)nullptr, n, k,
                           index);
        return hipGetLastError();
    }
    // DP ranks the alpha-inclusive vector, PP and PDP the kernel's values (accelerator.cc:866-883)
    const float* v = strategy == PPI_STRATEGY_FP ? vec : raw;
    const int ng = (n + best_local - 1) / best_local;
    hipLaunchKernelGGL(dev::rank_groups_kernel, dim3((unsigned)ng), dim3(1024), 0, st, v, n, best_local, group_val,
                       group_idx);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(dev::pick_n_best_kernel, dim3(1), dim3(1024), 0, st, group_val, group_idx, ng, k, index);
    return hipGetLastError();
}

}  // namespace ppi
