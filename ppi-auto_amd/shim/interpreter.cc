// interpreter.cc — PPI's interpreter entry points on the MI355X engine.
//
// Drop-in replacement for src/interpreter/accelerator.cc and sequential.cc:
// the same functions, signatures, argv options and per-mode results, built the
// way the reference builds libinterpreter — against the grammar's generated
// `symbol` header and the CMake-configured `definitions.h` / `costmatrix`.  The
// grammar tool (shim/grammar.py) writes `ppi_symbol_table.inc` next to `symbol`;
// it maps every Symbol of the grammar to an engine opcode.
//
// Every evaluation runs on the GPU through the C ABI of include/ppi_engine.h; if
// the engine cannot be created the init call fails (there is no CPU fallback).
// The headers sequential.cc sees (sequential.cc:16-22), first: ERROR(X,Y) expands here with
// the same overloads (libstdc++'s <stdlib.h> puts std::abs's float overloads in the global
// namespace; <cmath> alone leaves the C library's double routines there).
#include <stdio.h>
#include <stdlib.h>
#include <cmath>
#include <limits>
#include <string>

#include <cstring>
#include <stdexcept>
#include <vector>

#include <definitions.h>

#include "../../include/ppi_engine.h"
#include "../../include/ppi_interpreter.h"

namespace {

// Symbol -> engine opcode for this grammar.
const ppi_symbol kSymbolTable[] = {
#include "ppi_symbol_table.inc"
};

// Fallback COST, shadowed by the `costmatrix` definition when the build gives one.
const float COST[1][1] = {{0.0f}};

#define PPI_XSTR(x) PPI_STR(x)
#define PPI_STR(x) #x

struct ErrorFnInfo {
    int ef = -1;
    int cost_n = 0;
    std::vector<float> cost;
    std::string text;
    std::string why;  // why a user expression was not accepted
};

std::string squeeze(const char* s) {
    std::string out;
    for (; *s; ++s)
        if (*s != ' ' && *s != '\t' && *s != '\n') out += *s;
    return out;
}

// Identifies the compile-time ERROR(X,Y) (definitions.h:18) among the engine's
// built-in error functions: by its expansion text for the COST family, then by
// values on probe points for every candidate; any other expression is handed to the
// engine as text (PPI_EF_EXPR) once its compiled form has matched the macro.
ErrorFnInfo detect_error_fn() {
#include "costmatrix"
    ErrorFnInfo info;
    info.text = PPI_XSTR(ERROR(X, Y));
    const std::string t = squeeze(info.text.c_str());
    const bool uses_cost = t.find("COST") != std::string::npos;
    if (uses_cost) {
        const int n = (int)(sizeof(COST) / sizeof(COST[0]));
        const int m = (int)(sizeof(COST[0]) / sizeof(COST[0][0]));
        if (n != m) return info;
        info.cost_n = n;
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) info.cost.push_back(COST[i][j]);
        const bool clamp = t.find("fmax") != std::string::npos && t.find("fmin") != std::string::npos;
        // In-range probes only: the unclamped form indexes COST directly.
        bool ok = true;
        for (int i = 0; i < n && ok; ++i)
            for (int j = 0; j < n && ok; ++j) {
                const float X = (float)i + 0.25f, Y = (float)j;
                (void)X; (void)Y;
                const float e = ERROR(X, Y);
                ok = e == COST[i][j];
            }
        if (ok) info.ef = clamp ? PPI_EF_COST_CLAMP : PPI_EF_COST;
        return info;
    }
    const float px[] = {0.5f, -1.5f, 3.0f, 2.0f, 0.0f, 1e3f, -7.25f};
    const float py[] = {2.25f, 0.25f, 3.0f, 1.0f, 3.0f, -1e3f, -7.0f};
    bool is_abs = true, is_sq = true, is_neq = true;
    for (int k = 0; k < 7; ++k) {
        const float X = px[k], Y = py[k];
        const float e = ERROR(X, Y);
        is_abs &= e == std::fabs(X - Y);
        is_sq &= e == (X - Y) * (X - Y);
        is_neq &= e == (float)(X != Y);
    }
    info.ef = is_abs ? PPI_EF_ABS : is_sq ? PPI_EF_SQUARED : is_neq ? PPI_EF_NEQ : -1;
    if (info.ef >= 0) return info;
    // Any other expression: compiled by the engine (PPI_EF_EXPR), accepted only if the compiled
    // form gives this build's own ERROR(X,Y) bit for bit on a probe set (specials, small
    // integers and halves, and a spread of magnitudes).
    std::vector<float> vals = {0.0f, -0.0f, 0.5f, -0.5f, 1.0f, -1.0f, 1.5f, 2.0f, -2.5f, 3.0f, 7.0f, 1e-30f,
                               1e-45f, 3.0e38f, -3.0e38f, 1e30f, std::numeric_limits<float>::infinity(),
                               -std::numeric_limits<float>::infinity(), std::numeric_limits<float>::quiet_NaN()};
    uint32_t r = 2463534242u;
    for (int k = 0; k < 40; ++k) {
        r ^= r << 13; r ^= r >> 17; r ^= r << 5;
        vals.push_back(((float)(int32_t)r) * 0x1p-31f * (float)(1 << (k % 20)));
    }
    std::vector<float> xs, ys, want;
    for (float X : vals)
        for (float Y : vals) {
            xs.push_back(X);
            ys.push_back(Y);
            want.push_back(ERROR(X, Y));
        }
    std::vector<float> got(xs.size());
    if (ppi_ef_evaluate(info.text.c_str(), xs.data(), ys.data(), (int32_t)xs.size(), got.data()) != PPI_OK) {
        info.why = ppi_last_error();
        return info;
    }
    for (size_t k = 0; k < xs.size(); ++k) {
        uint32_t a, b;
        std::memcpy(&a, &want[k], 4);
        std::memcpy(&b, &got[k], 4);
        if (a != b && !(want[k] != want[k] && got[k] != got[k])) {
            char buf[160];
            snprintf(buf, sizeof buf, "the compiled expression gives %a at X = %a, Y = %a where the build gives %a",
                     (double)got[k], (double)xs[k], (double)ys[k], (double)want[k]);
            info.why = buf;
            return info;
        }
    }
    info.ef = PPI_EF_EXPR;
    return info;
}

struct Options {
    std::string strategy = "PDP";  // accelerator.cc:571 default
    int mls = -1, device = -1, platform = -1;
    bool verbose = false, transpose = false;
    std::string type;
    int form = PPI_FORM_AUTO, rows_per_lane = 0;
    std::string error;
};

// The options acc_interpret_init reads (accelerator.cc:565-571); everything else
// in argv belongs to other modules and is ignored, as CmdLine::Parser does.  A
// "--" stops the scan.  Two engine extras: -acc-form PP|FP|PDP|ORDERED and -acc-rows N.
Options parse_options(int argc, char** argv) {
    Options o;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i] ? argv[i] : "";
        if (a == "--") break;
        auto value = [&](std::string* out) {
            if (i + 1 < argc && argv[i + 1]) { *out = argv[++i]; return true; }
            o.error = "option " + a + " needs a value";
            return false;
        };
        std::string v;
        if (a == "-strategy") {
            if (!value(&v)) break;
            if (v == "pp" || v == "PP") o.strategy = "PP";
            else if (v == "dp" || v == "DP") o.strategy = "DP";
            else if (v == "pdp" || v == "PDP") o.strategy = "PDP";
            else { o.error = "Valid strategy: PP, DP and PDP."; break; }
        } else if (a == "-cl-mls" || a == "--cl-max-local-size") {
            if (!value(&v)) break;
            o.mls = std::atoi(v.c_str());
        } else if (a == "-cl-d" || a == "--cl-device-id") {
            if (!value(&v)) break;
            o.device = std::atoi(v.c_str());
        } else if (a == "-cl-p" || a == "--cl-platform-id") {
            if (!value(&v)) break;
            o.platform = std::atoi(v.c_str());
        } else if (a == "-type") {
            if (!value(&v)) break;
            o.type = v;
            if (v != "GPU" && v != "gpu") { o.error = "Not a single compatible device found."; break; }
        } else if (a == "-v" || a == "--verbose") {
            o.verbose = true;
        } else if (a == "-transpose" || a == "--transpose") {
            o.transpose = true;  // the engine always stores the dataset column-major
        } else if (a == "-acc-form") {
            if (!value(&v)) break;
            o.form = v == "PP" ? PPI_FORM_PP : v == "FP" ? PPI_FORM_FP : v == "PDP" ? PPI_FORM_PDP
                   : v == "ORDERED" ? PPI_FORM_ORDERED : PPI_FORM_AUTO;
        } else if (a == "-acc-rows") {
            if (!value(&v)) break;
            o.rows_per_lane = std::atoi(v.c_str());
        }
    }
    return o;
}

struct Interp {
    ppi_engine* eng = nullptr;
    unsigned max_size = 0;
    int nlin = 0;
    double dataset_ms = 0.0;
    // PROFILING accumulators (accelerator.cc:52 / sequential.cc:30 field meanings)
    double t_kernel1_gen = 0.0, t_kernel1_total = 0.0;
    double t_kernel2_gen = 0.0, t_kernel2_total = 0.0;
    double t_send1_gen = 0.0, t_send1_total = 0.0;
    double t_recv1_gen = 0.0, t_recv1_total = 0.0;
    double gpops_kernel = 0.0, gpops_comm = 0.0;
};

Interp g_acc, g_seq;

[[noreturn]] void fail(const char* where, ppi_status s) {
    throw std::runtime_error(std::string(where) + ": " + ppi_last_error() + " (status " + std::to_string((int)s) + ")");
}

// `native`: the accelerator entry points honour -DNATIVE (the OpenCL native_* built-ins of
// functions.h:59-193); the sequential ones never do, because their native_* shims are the
// double routines themselves (sequential.cc:32-74), so a NATIVE build of seq_interpret gives
// the same bits as one without it.
int setup(Interp& I, int device, int strategy, int form, int mls, int rows_per_lane, const unsigned size,
          float** input, int nlin, int ncol, bool verbose, bool native) {
    ErrorFnInfo ef = detect_error_fn();
    if (ef.ef < 0) {
        fprintf(stderr, "Unsupported error function ERROR(X,Y) = %s: %s\n", ef.text.c_str(),
                ef.why.empty() ? "COST is supported as COST[(int)X][(int)Y] and its clamped form"
                               : ef.why.c_str());
        return 1;
    }
    ppi_status s = ppi_engine_create(&I.eng, device < 0 ? 0 : device);
    if (s) {
        fprintf(stderr, "Error in device initialization: %s\n", ppi_last_error());
        return 1;
    }
    ppi_config cfg;
    std::memset(&cfg, 0, sizeof(cfg));
    cfg.strategy = strategy;
    cfg.form = form;
    cfg.error_fn = ef.ef;
#ifdef REDUCEMAX
    cfg.reducemax = 1;
#endif
#ifdef PROTECTED
    cfg.protected_ops = 1;
#endif
    cfg.native_math = native ? 1 : 0;
    cfg.cost_n = ef.cost_n;
    cfg.cost = ef.cost.empty() ? nullptr : ef.cost.data();
    cfg.ef_expr = ef.ef == PPI_EF_EXPR ? ef.text.c_str() : nullptr;
    cfg.max_local_size = mls > 0 ? mls : 0;
    cfg.rows_per_lane = rows_per_lane;
    s = ppi_engine_configure(I.eng, &cfg, kSymbolTable, (int32_t)(sizeof(kSymbolTable) / sizeof(kSymbolTable[0])));
    if (!s) s = ppi_engine_set_dataset(I.eng, input, nlin, ncol);
    if (s) {
        fprintf(stderr, "Error in initialization phase: %s\n", ppi_last_error());
        ppi_engine_destroy(I.eng);
        I.eng = nullptr;
        return 1;
    }
    I.max_size = size;
    I.nlin = nlin;
    if (verbose) fprintf(stdout, "\nEngine: HIP device %d, error function %s\n", device < 0 ? 0 : device, ef.text.c_str());
    return 0;
}

void evaluate(Interp& I, const char* who, Symbol* phenotype, float* ephemeral, int* size, float* vector, int nInd,
              void (*send)(Population*), int (*receive)(GENOME_TYPE**), Population* migrants, int* nImmigrants,
              int* index, int* best_size, int ppp_mode, int prediction_mode, float alpha,
              unsigned long sum_size_gen) {
    if (!I.eng) throw std::runtime_error(std::string(who) + ": interpreter not initialised");
    static_assert(sizeof(Symbol) == sizeof(int32_t), "Symbol must be a 32-bit enum");
    ppi_status s;
    if (ppp_mode && prediction_mode) {
        // ppp -pred: one value per row (accelerator.cl:72-75, sequential.cc:165-167).
        s = ppi_engine_predict(I.eng, reinterpret_cast<const int32_t*>(phenotype), ephemeral, size[0], vector);
        if (s) fail(who, s);
        return;
    }
    s = ppi_engine_upload(I.eng, reinterpret_cast<const int32_t*>(phenotype), ephemeral, size, (int32_t)I.max_size, nInd);
    if (s) fail(who, s);
    s = ppi_engine_launch(I.eng);
    if (s) fail(who, s);
    // Island exchange overlaps the device work, as in accelerator.cc:736-740.
    if (!ppp_mode && send && receive) {
        send(migrants);
        *nImmigrants = receive(migrants->genome);
    }
    int32_t bs = best_size ? *best_size : 0;
    s = ppi_engine_wait(I.eng, alpha, vector, (!ppp_mode && index) ? index : nullptr,
                        (!ppp_mode && best_size) ? &bs : nullptr);
    if (s) fail(who, s);
    if (!ppp_mode && best_size) *best_size = bs;
    ppi_stats st;
    if (ppi_engine_get_stats(I.eng, &st) == PPI_OK) {
        I.t_kernel1_gen = st.kernel_ms * 1e-3;
        I.t_kernel2_gen = st.download_ms * 1e-3;
        I.t_send1_gen = st.upload_ms * 1e-3;
        I.t_recv1_gen = 0.0;
        I.t_kernel1_total += I.t_kernel1_gen;
        I.t_kernel2_total += I.t_kernel2_gen;
        I.t_send1_total += I.t_send1_gen;
        I.t_recv1_total += I.t_recv1_gen;
        const double w = (double)sum_size_gen * I.nlin;
        I.gpops_kernel = I.t_kernel1_gen > 0 ? w / I.t_kernel1_gen : 0.0;
        const double tc = I.t_kernel1_gen + I.t_send1_gen + I.t_kernel2_gen;
        I.gpops_comm = tc > 0 ? w / tc : 0.0;
    }
}

void cleanup() {
    if (g_acc.eng) ppi_engine_destroy(g_acc.eng);
    g_acc.eng = nullptr;
    if (g_seq.eng) ppi_engine_destroy(g_seq.eng);
    g_seq.eng = nullptr;
}

}  // namespace

int acc_interpret_init(int argc, char** argv, const unsigned size, const unsigned max_arity,
                       const unsigned population_size, float** input, int nlin, int ncol, int ppp_mode,
                       int prediction_mode) {
    (void)max_arity;  // the engine sizes stacks per program at upload time
    (void)population_size;
    (void)ppp_mode;
    (void)prediction_mode;
    Options o = parse_options(argc, argv);
    if (!o.error.empty()) {
        fprintf(stderr, "%s\n", o.error.c_str());
        return 1;
    }
    if (ncol < 0) {
        fprintf(stderr, "Missing number of columns of dataset.\n");
        return 1;
    }
    if (g_acc.eng) {
        ppi_engine_destroy(g_acc.eng);
        g_acc.eng = nullptr;
    }
    const int strategy = o.strategy == "PP" ? PPI_STRATEGY_PP : o.strategy == "DP" ? PPI_STRATEGY_FP : PPI_STRATEGY_PDP;
    // evaluate_pp adds each program's rows in order (accelerator.cl:22-84): the ordered form gives
    // the same sums bit for bit.  -acc-form picks another form (FP: faster, sums in tile order).
    int form = o.form;
    if (form == PPI_FORM_AUTO && strategy == PPI_STRATEGY_PP) form = PPI_FORM_ORDERED;
#ifdef NATIVE
    const bool native = true;
#else
    const bool native = false;
#endif
    if (setup(g_acc, o.device, strategy, form, o.mls, o.rows_per_lane, size, input, nlin, ncol, o.verbose, native))
        return 1;
    static bool registered = false;
    if (!registered) { std::atexit(cleanup); registered = true; }
    return 0;
}

void acc_interpret(Symbol* phenotype, float* ephemeral, int* size,
#ifdef PROFILING
                   unsigned long sum_size_gen,
#endif
                   float* vector, int nInd, void (*send)(Population*), int (*receive)(GENOME_TYPE**),
                   Population* migrants, int* nImmigrants, int* index, int* best_size, int ppp_mode,
                   int prediction_mode, float alpha) {
#ifndef PROFILING
    const unsigned long sum_size_gen = 0;
#endif
    evaluate(g_acc, "acc_interpret", phenotype, ephemeral, size, vector, nInd, send, receive, migrants, nImmigrants,
             index, best_size, ppp_mode, prediction_mode, alpha, sum_size_gen);
}

#ifdef PROFILING
void acc_print_time(bool total, unsigned long long sum_size) {
    // Same fields as accelerator.cc:945-958: kernel[1] = device evaluation,
    // kernel[2] = download + ranking, send1 = program upload (pack + H2D).
    const Interp& I = g_acc;
    double k1 = total ? I.t_kernel1_total : I.t_kernel1_gen;
    double k2 = total ? I.t_kernel2_total : I.t_kernel2_gen;
    double s1 = total ? I.t_send1_total : I.t_send1_gen;
    double r1 = total ? I.t_recv1_total : I.t_recv1_gen;
    double gk = total ? (k1 > 0 ? sum_size * (double)I.nlin / k1 : 0.0) : I.gpops_kernel;
    double gc = total ? ((k1 + s1 + k2) > 0 ? sum_size * (double)I.nlin / (k1 + s1 + k2) : 0.0) : I.gpops_comm;
    printf(", time_kernel[1]: %lf, time_kernel[2]: %lf, time_communication_dataset: %lf, time_communication_send1: %lf, "
           "time_communication_receive1: %lf, time_communication_send2: %lf, time_communication_receive2: %lf",
           k1, k2, I.dataset_ms * 1e-3, s1, r1, 0.0, 0.0);
    printf("; gpops_kernel: %lf, gpops_kernel_communication: %lf", gk, gc);
}
#endif

void seq_interpret_init(const unsigned size, float** input, int nlin, int ncol) {
    if (g_seq.eng) {
        ppi_engine_destroy(g_seq.eng);
        g_seq.eng = nullptr;
    }
    // sequential.cc has no options; PPI_DEVICE selects the GPU (default 0).  The loop of
    // sequential.cc:133-190 adds each program's rows in order: the ordered form gives the same sums
    // bit for bit.  PPI_SEQ_FORM=FP selects the faster FP form (sums in tile order, within a relative
    // 1e-5 of the sequential ones).
    const char* dev = std::getenv("PPI_DEVICE");
    const char* sf = std::getenv("PPI_SEQ_FORM");
    const int form = (sf && std::strcmp(sf, "FP") == 0) ? PPI_FORM_FP : PPI_FORM_ORDERED;
    if (setup(g_seq, dev ? std::atoi(dev) : 0, PPI_STRATEGY_SEQ, form, 0, 0, size, input, nlin, ncol, false, false))
        throw std::runtime_error("seq_interpret_init: engine initialisation failed");
    static bool registered = false;
    if (!registered) { std::atexit(cleanup); registered = true; }
}

void seq_interpret(Symbol* phenotype, float* ephemeral, int* size,
#ifdef PROFILING
                   unsigned long sum_size_gen,
#endif
                   float* vector, int nInd, int* index, int* best_size, int ppp_mode, int prediction_mode,
                   float alpha) {
#ifndef PROFILING
    const unsigned long sum_size_gen = 0;
#endif
    evaluate(g_seq, "seq_interpret", phenotype, ephemeral, size, vector, nInd, nullptr, nullptr, nullptr, nullptr,
             index, best_size, ppp_mode, prediction_mode, alpha, sum_size_gen);
}

void seq_interpret_destroy() {
    if (g_seq.eng) ppi_engine_destroy(g_seq.eng);
    g_seq.eng = nullptr;
}

#ifdef PROFILING
void seq_print_time(bool total, unsigned long long sum_size) {
    // sequential.cc:219-227 fields.
    const Interp& I = g_seq;
    double k1 = total ? I.t_kernel1_total : I.t_kernel1_gen;
    double k2 = total ? I.t_kernel2_total : I.t_kernel2_gen;
    double gk = total ? (k1 > 0 ? sum_size * (double)I.nlin / k1 : 0.0) : I.gpops_kernel;
    printf(", time_kernel[1]: %lf, time_kernel[2]: %lf", k1, k2);
    printf("; gpops_kernel: %lf", gk);
}
#endif
