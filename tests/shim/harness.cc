// Test harness for the drop-in interpreter library: replays the call sequences
// of PPI's ppi.cc (GE evaluation, ppi.cc:604-634) and ppp.cc (single solution,
// ppp.cc:98-140) against libppi_interpreter and writes the results to a file.
//
//   harness <acc|seq|acc-ppp|seq-ppp|acc-pred|seq-pred> <in> <out> [reference argv ...]
//
// in:  int32 nlin, ncol, n, stride, best_size; float alpha; float data[nlin*ncol];
//      int32 phenotype[n*stride]; float ephemeral[n*stride]; int32 size[n]
// out: int32 count; float vector[count]; int32 best_size; int32 index[best_size];
//      int32 send_calls, receive_calls, nImmigrants
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "ppi_interpreter.h"

static int g_send = 0, g_recv = 0;
static void send_cb(Population* p) { (void)p; ++g_send; }
static int receive_cb(GENOME_TYPE** g) { (void)g; ++g_recv; return 3; }

template <typename T>
static void rd(FILE* f, T* p, size_t n) {
    if (fread(p, sizeof(T), n, f) != n) { fprintf(stderr, "short read\n"); std::exit(3); }
}

int main(int argc, char** argv) {
    if (argc < 4) { fprintf(stderr, "usage: harness mode in out [argv...]\n"); return 2; }
    const std::string mode = argv[1];
    FILE* f = fopen(argv[2], "rb");
    if (!f) return 3;
    int32_t hdr[5];
    float alpha;
    rd(f, hdr, 5);
    rd(f, &alpha, 1);
    const int nlin = hdr[0], ncol = hdr[1], n = hdr[2], stride = hdr[3];
    int best_size = hdr[4];
    std::vector<float> data((size_t)nlin * ncol);
    std::vector<int32_t> ph((size_t)n * stride), sz((size_t)n);
    std::vector<float> eph((size_t)n * stride);
    rd(f, data.data(), data.size());
    rd(f, ph.data(), ph.size());
    rd(f, eph.data(), eph.size());
    rd(f, sz.data(), sz.size());
    fclose(f);
    std::vector<float*> rows((size_t)nlin);
    for (int r = 0; r < nlin; ++r) rows[r] = data.data() + (size_t)r * ncol;

    // argv handed to the interpreter: the program name followed by the extra options.
    std::vector<char*> av;
    av.push_back(argv[0]);
    for (int i = 4; i < argc; ++i) av.push_back(argv[i]);

    const bool acc = mode.rfind("acc", 0) == 0;
    const bool ppp = mode.find("ppp") != std::string::npos || mode.find("pred") != std::string::npos;
    const bool pred = mode.find("pred") != std::string::npos;
    std::vector<float> vec(pred ? (size_t)nlin : (size_t)n);
    std::vector<int32_t> index((size_t)(best_size > 0 ? best_size : 1));
    Population migrants{nullptr, nullptr};
    int nImmigrants = -1;
    Symbol* sph = reinterpret_cast<Symbol*>(ph.data());
    try {
        if (acc) {
            if (acc_interpret_init((int)av.size(), av.data(), (unsigned)stride, ppp ? (unsigned)-1 : 4u,
                                   (unsigned)n, rows.data(), nlin, ncol, ppp ? 1 : 0, pred ? 1 : 0))
                return 4;
            acc_interpret(sph, eph.data(), sz.data(),
#ifdef PROFILING
                          0ul,
#endif
                          vec.data(), n, ppp ? nullptr : &send_cb, ppp ? nullptr : &receive_cb,
                          ppp ? nullptr : &migrants, ppp ? nullptr : &nImmigrants, ppp ? nullptr : index.data(),
                          ppp ? nullptr : &best_size, ppp ? 1 : 0, pred ? 1 : 0, ppp ? 0.0f : alpha);
#ifdef PROFILING
            acc_print_time(false, 0);
            printf("\n");
#endif
        } else {
            seq_interpret_init((unsigned)stride, rows.data(), nlin, ncol);
            seq_interpret(sph, eph.data(), sz.data(),
#ifdef PROFILING
                          0ul,
#endif
                          vec.data(), n, ppp ? nullptr : index.data(), ppp ? nullptr : &best_size, ppp ? 1 : 0,
                          pred ? 1 : 0, ppp ? 0.0f : alpha);
#ifdef PROFILING
            seq_print_time(false, 0);
            printf("\n");
#endif
            seq_interpret_destroy();
        }
    } catch (const std::exception& e) {
        fprintf(stderr, "> Error: %s\n", e.what());
        return 8;  // main.cc:201-203
    }
    FILE* o = fopen(argv[3], "wb");
    if (!o) return 3;
    int32_t cnt = (int32_t)vec.size();
    fwrite(&cnt, 4, 1, o);
    fwrite(vec.data(), 4, vec.size(), o);
    int32_t bs = ppp ? 0 : best_size;
    fwrite(&bs, 4, 1, o);
    fwrite(index.data(), 4, (size_t)bs, o);
    int32_t cb[3] = {g_send, g_recv, nImmigrants};
    fwrite(cb, 4, 3, o);
    fclose(o);
    return 0;
}
