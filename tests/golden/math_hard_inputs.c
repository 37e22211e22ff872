/* Float inputs x whose glibc result f((double) x) lies within 256 double ulps of a float rounding
 * boundary (bits 0-28 of the double significand within 256 of 0x10000000): the inputs where a
 * double routine within a few ulps of glibc's could round to a different float, and where the
 * engine's exact fast paths (which accept a result more than 64 ulps from a boundary) fall back
 * to the device library's double routine.  Scans all 2^32 float bit patterns for each function;
 * writes hard_<fn>.u32 (the input bit patterns, increasing) into the directory given.
 * Built and run by make_math_hard_inputs.py. */
#define _GNU_SOURCE 1
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef double (*Fn)(double);
static const struct { const char* name; Fn f; } kFns[] = {
    {"exp", exp}, {"log", log}, {"sin", sin}, {"cos", cos}, {"tan", tan},
    {"exp10", exp10}, {"exp2", exp2}, {"log10", log10}, {"log2", log2},
};

static int hard(double y) {
    const double a = fabs(y);
    if (!(a >= 0x1p-126 && a < 0x1p128)) return 0;
    uint64_t b;
    memcpy(&b, &y, 8);
    const int32_t d = (int32_t)(uint32_t)(b & 0x1fffffffu) - 0x10000000;
    return d >= -256 && d <= 256;
}

int main(int argc, char** argv) {
    if (argc != 2) { fprintf(stderr, "usage: %s outdir\n", argv[0]); return 2; }
    for (size_t k = 0; k < sizeof kFns / sizeof kFns[0]; ++k) {
        enum { kChunks = 256 };
        uint32_t* found[kChunks];
        size_t nfound[kChunks];
#pragma omp parallel for schedule(dynamic)
        for (int c = 0; c < kChunks; ++c) {
            size_t cap = 1024, n = 0;
            uint32_t* out = malloc(cap * sizeof(uint32_t));
            const uint64_t lo = (uint64_t)c << 24, hi = lo + (1ull << 24);
            for (uint64_t u = lo; u < hi; ++u) {
                const uint32_t bits = (uint32_t)u;
                float x;
                memcpy(&x, &bits, 4);
                if (!isfinite(x)) continue;
                if (hard(kFns[k].f((double)x))) {
                    if (n == cap) out = realloc(out, (cap *= 2) * sizeof(uint32_t));
                    out[n++] = bits;
                }
            }
            found[c] = out;
            nfound[c] = n;
        }
        char path[4096];
        snprintf(path, sizeof path, "%s/hard_%s.u32", argv[1], kFns[k].name);
        FILE* f = fopen(path, "wb");
        if (!f) { perror(path); return 1; }
        size_t total = 0;
        for (int c = 0; c < kChunks; ++c) {
            fwrite(found[c], sizeof(uint32_t), nfound[c], f);
            total += nfound[c];
            free(found[c]);
        }
        fclose(f);
        printf("%s: %zu hard inputs\n", kFns[k].name, total);
        fflush(stdout);
    }
    return 0;
}
