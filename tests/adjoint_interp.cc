// Test-only host interpreter of the adjoint tape (no GPU involved).
//
// Executes ppi::compile_adjoint's output the way the device kernel does — per
// job and row: for each segment, the forward tape storing that segment's
// partials, then the segment's reverse instructions — in float.  Comparing its
// gradients with the oracle's FP64 forward-mode (dual number) gradient pins the
// compiler's active-path analysis, partial records, adjoint slot assignment and
// segmentation on the CPU; running it with different partial-store sizes B pins
// that segmentation (forward recomputation) does not change the result.
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../ppi-auto_amd/csrc/tape.h"

using namespace ppi;

namespace {

float bf(bool b) { return b ? 1.0f : 0.0f; }
const float PI_2 = 1.57079637050629f, PI = 3.14159274101257f, E = 2.71828174591064f;

float sq(float x, int P) { return (P && x < 0.0f) ? 1.0f : (float)std::sqrt((double)x); }

float un(uint32_t u, float a, int P) {
    switch (u) {
    case U_NOT: return bf(!(a != 0.0f));
    case U_ABS: return std::fabs(a);
    case U_POW2: return a * a;
    case U_POW3: return a * a * a;
    case U_POW4: return a * a * a * a;
    case U_POW5: return a * a * a * a * a;
    case U_NEG: return -a;
    case U_ROUND: return std::round(a);
    case U_CEIL: return std::ceil(a);
    case U_FLOOR: return std::floor(a);
    case U_EXP: return (float)std::exp((double)a);
    case U_EXP10: return (float)::exp10((double)a);
    case U_EXP2: return (float)std::exp2((double)a);
    case U_SIN: return (float)std::sin((double)a);
    case U_COS: return (float)std::cos((double)a);
    case U_STEP: return bf(a >= 0.0f);
    case U_SIGN: return a > 0.0f ? 1.0f : (a < 0.0f ? -1.0f : 0.0f);
    case U_LOGISTIC: return 1.0f / (1.0f + (float)std::exp((double)-a));
    case U_GAUSS: return (float)std::exp((double)(-a * a));
    case U_SQRT: return sq(a, P);
    case U_LOG: return (P && a < 1.0f) ? 1.0f : (float)std::log((double)a);
    case U_LOG10: return (P && a < 1.0f) ? 1.0f : (float)std::log10((double)a);
    case U_LOG2: return (P && a < 1.0f) ? 1.0f : (float)std::log2((double)a);
    case U_TAN:
        if (P && (std::fabs(a) == PI_2 || std::fabs(a) == 3.0f * PI_2)) return 1.0f;
        return (float)std::tan((double)a);
    default: return NAN;  // GAMMA is not used by these tests
    }
}

float bin(uint32_t op, float a, float b, int P) {
    switch (op) {
    case B_ADD: return a + b;
    case B_SUB: return a - b;
    case B_RSUB: return b - a;
    case B_MULT: return a * b;
    case B_DIV: return (P && b == 0.0f) ? 1.0f : a / b;
    case B_RDIV: return (P && a == 0.0f) ? 1.0f : b / a;
    case B_MEAN: return (a + b) / 2.0f;
    case B_MAX: return std::fmax(a, b);
    case B_RMAX: return std::fmax(b, a);
    case B_MIN: return std::fmin(a, b);
    case B_RMIN: return std::fmin(b, a);
    case B_MOD: return (float)std::fmod((double)a, (double)b);
    case B_RMOD: return (float)std::fmod((double)b, (double)a);
    case B_POW: return (float)std::pow((double)a, (double)b);
    case B_RPOW: return (float)std::pow((double)b, (double)a);
    case B_AND: return bf(a != 0.0f && b != 0.0f);
    case B_OR: return bf(a != 0.0f || b != 0.0f);
    case B_XOR: return bf((a != 0.0f) != (b != 0.0f));
    case B_GT: return bf(a > b);
    case B_GE: return bf(a >= b);
    case B_LT: return bf(a < b);
    case B_LE: return bf(a <= b);
    case B_EQ: return bf(a == b);
    default: return bf(a != b);
    }
}

// Local derivatives (calculus of each operator; protected branches are constant).
float dun(uint32_t u, float a, float z, int P) {
    switch (u) {
    case U_ABS: return a > 0.0f ? 1.0f : (a < 0.0f ? -1.0f : 0.0f);
    case U_POW2: return 2.0f * a;
    case U_POW3: return 3.0f * a * a;
    case U_POW4: return 4.0f * a * a * a;
    case U_POW5: return 5.0f * a * a * a * a;
    case U_NEG: return -1.0f;
    case U_EXP: return z;
    case U_EXP10: return z * 2.30258509f;
    case U_EXP2: return z * 0.693147181f;
    case U_SIN: return std::cos(a);
    case U_COS: return -std::sin(a);
    case U_LOGISTIC: return z * (1.0f - z);
    case U_GAUSS: return -2.0f * a * z;
    case U_SQRT: return (P && a < 0.0f) ? 0.0f : 0.5f / z;
    case U_LOG: return (P && a < 1.0f) ? 0.0f : 1.0f / a;
    case U_LOG10: return (P && a < 1.0f) ? 0.0f : 1.0f / (a * 2.30258509f);
    case U_LOG2: return (P && a < 1.0f) ? 0.0f : 1.0f / (a * 0.693147181f);
    case U_TAN: return (P && (std::fabs(a) == PI_2 || std::fabs(a) == 3.0f * PI_2)) ? 0.0f : 1.0f + z * z;
    default: return 0.0f;
    }
}

void dbin(uint32_t op, float a, float b, float z, int P, float* da, float* db) {
    *da = *db = 0.0f;
    switch (op) {
    case B_ADD: *da = 1; *db = 1; break;
    case B_SUB: *da = 1; *db = -1; break;
    case B_RSUB: *da = -1; *db = 1; break;
    case B_MULT: *da = b; *db = a; break;
    case B_MEAN: *da = 0.5f; *db = 0.5f; break;
    case B_DIV: if (!(P && b == 0.0f)) { *da = 1.0f / b; *db = -z / b; } break;
    case B_RDIV: if (!(P && a == 0.0f)) { *db = 1.0f / a; *da = -z / a; } break;
    case B_MAX: *da = (std::isnan(b) || a >= b) ? 1.0f : 0.0f; *db = 1.0f - *da; break;
    case B_RMAX: *db = (std::isnan(a) || b >= a) ? 1.0f : 0.0f; *da = 1.0f - *db; break;
    case B_MIN: *da = (std::isnan(b) || a <= b) ? 1.0f : 0.0f; *db = 1.0f - *da; break;
    case B_RMIN: *db = (std::isnan(a) || b <= a) ? 1.0f : 0.0f; *da = 1.0f - *db; break;
    case B_MOD: *da = 1; *db = -std::trunc(a / b); break;
    case B_RMOD: *db = 1; *da = -std::trunc(b / a); break;
    case B_POW: *da = b * std::pow(a, b - 1.0f); *db = a > 0.0f ? z * std::log(a) : 0.0f; break;
    case B_RPOW: *db = a * std::pow(b, a - 1.0f); *da = b > 0.0f ? z * std::log(b) : 0.0f; break;
    default: break;
    }
}

}  // namespace

// grad: n*stride (ephemeral layout), fit: n (sum of errors / nlin, NaN if a row's error is not finite).
// info[0] = jobs, info[1] = most segments of a job, info[2] = most adjoint slots, info[3] = constants.
extern "C" int adj_eval(const ppi_symbol* table, int ntable, const int32_t* ph, const float* eph, const int32_t* size,
                        int stride, int n, const float* data, int nlin, int ncol, int P, int ef, int B, float* grad,
                        float* fit, int32_t* info) {
    SymbolTable st;
    st.build(table, ntable);
    AdjTape T;
    std::string err;
    if (!compile_adjoint(st, ph, eph, size, stride, n, ncol, B, &T, &err)) return -1;
    std::memset(grad, 0, sizeof(float) * (size_t)n * stride);
    for (int i = 0; i < n; ++i) fit[i] = 0.0f;
    std::vector<double> fsum((size_t)n, 0.0);
    std::vector<char> bad((size_t)n, 0);
    std::vector<double> gsum(T.consts.size(), 0.0);
    std::vector<float> ps((size_t)B);
    for (const AdjJob& J : T.jobs) {
        if (J.flags & kEmpty) { bad[(size_t)J.prog] = 1; continue; }
        for (int r = 0; r < nlin; ++r) {
            const float* row = data + (size_t)r * ncol;
            const float y = row[ncol - 1];
            float adj[kAdjSlots] = {0};
            float g[kAdjMaxConsts] = {0};
            for (uint32_t s = 0; s < J.nseg; ++s) {
                const AdjSeg& sg = T.segs[J.seg_off + s];
                float acc = NAN, reg[kMaxSlots];
                for (int k = 0; k < kMaxSlots; ++k) reg[k] = NAN;
                uint32_t next = sg.st_begin;
                for (uint32_t pc = 0; pc < sg.fwd_end; ++pc) {
                    const Ins in = T.fwd.code[J.code_off + pc];
                    int sa = -1, sb = -1;
                    if (next < sg.st_end && T.stores[next].pc == pc) {
                        const uint32_t sl = T.stores[next++].slots;
                        sa = (int)(sl & 0xffffu) - 1;
                        sb = (int)(sl >> 16) - 1;
                    }
                    const uint32_t h = in.code & 0xffu, s1 = (in.code >> 8) & 31u, s2 = (in.code >> 13) & 31u;
                    float cv;
                    std::memcpy(&cv, &in.arg, 4);
                    auto term = [&](bool attr) { return attr ? row[in.arg] : cv; };
                    if (h >= H_BIN_C) {
                        uint32_t op;
                        float b;
                        if (h >= H_BIN_S) { op = h - H_BIN_S; b = reg[s1]; }
                        else if (h >= H_BIN_A) { op = h - H_BIN_A; b = term(true); }
                        else { op = h - H_BIN_C; b = term(false); }
                        const float z = bin(op, acc, b, P);
                        float da, db;
                        dbin(op, acc, b, z, P, &da, &db);
                        if (sa >= 0) ps[(size_t)sa] = da;
                        if (sb >= 0) ps[(size_t)sb] = db;
                        acc = z;
                    } else if (h >= H_UN) {
                        const float z = un(h - H_UN, acc, P);
                        if (sa >= 0) ps[(size_t)sa] = dun(h - H_UN, acc, z, P);
                        acc = z;
                    } else if (h == H_LOAD_C || h == H_LOAD_A) {
                        acc = term(h == H_LOAD_A);
                    } else if (h == H_PUSH_C || h == H_PUSH_A) {
                        reg[s1] = acc;
                        acc = term(h == H_PUSH_A);
                    } else {
                        const float v[3] = {acc, reg[s1], reg[s2]};
                        const float c = v[(in.code >> 18) & 3u], th = v[(in.code >> 20) & 3u], el = v[(in.code >> 22) & 3u];
                        if (sa >= 0) ps[(size_t)sa] = c != 0.0f ? 1.0f : 0.0f;
                        if (sb >= 0) ps[(size_t)sb] = c != 0.0f ? 0.0f : 1.0f;
                        acc = (c != 0.0f) ? th : el;
                    }
                }
                if (next != sg.st_end) return -2;  // a store outside [0, fwd_end) or out of order
                if (s == 0) {
                    const float d = acc - y;
                    const float e = ef == PPI_EF_SQUARED ? d * d : std::fabs(d);
                    if (J.flags & kAdjFitness) {
                        if (!std::isfinite(e)) bad[(size_t)J.prog] = 1;
                        else fsum[(size_t)J.prog] += e;
                    }
                    adj[0] = ef == PPI_EF_SQUARED ? 2.0f * d : (d > 0.0f ? 1.0f : (d < 0.0f ? -1.0f : 0.0f));
                }
                for (uint32_t k = sg.rev_begin; k < sg.rev_end; ++k) {
                    const RevIns ri = T.rev[k];
                    const uint32_t src = (ri.w >> 4) & 15u, dst = (ri.w >> 8) & 31u;
                    float p;
                    if (ri.w & kRevStored) p = ps[(size_t)(((ri.w >> 16) & 0xffffu) % (uint32_t)B)];
                    else std::memcpy(&p, &ri.p, 4);
                    const float v = adj[src] == 0.0f ? 0.0f : adj[src] * p;
                    if ((ri.w & 3u) == kRevAdj) adj[dst] = v;
                    else g[dst] += v;
                }
            }
            for (uint32_t k = 0; k < J.nconst; ++k) gsum[J.const_base + k] += g[k];
        }
    }
    for (size_t k = 0; k < T.consts.size(); ++k) {
        const AdjConst& c = T.consts[k];
        grad[(size_t)c.prog * stride + c.pos] = bad[(size_t)c.prog] ? 0.0f : (float)(gsum[k] / nlin);
    }
    for (int i = 0; i < n; ++i) fit[i] = bad[(size_t)i] ? NAN : (float)(fsum[(size_t)i] / nlin);
    info[0] = (int32_t)T.jobs.size();
    info[1] = T.max_seg;
    info[2] = T.max_adj_slots;
    info[3] = (int32_t)T.consts.size();
    return 0;
}
