// Draws from the reference's own random number generator (src/util/Random.h, XorShift128Plus and
// BaseRNG::Real / Probability), compiled from that header where it lies in the reference tree.
// Test infrastructure: its output pins the oracle's restatement (oracle/ppi_ge_oracle.c).
//
//   ref_rng <seed> <n>   ->  n lines "Int() Real() Probability(0.66)" from three generators
//                            seeded `seed` (each consumes its own stream), printed exactly.
#include "Random.h"

#include <cstdio>
#include <cstdlib>

int main(int argc, char** argv) {
    if (argc != 3) {
        std::fprintf(stderr, "usage: ref_rng seed n\n");
        return 2;
    }
    const unsigned long seed = std::strtoul(argv[1], nullptr, 10);
    const long n = std::strtol(argv[2], nullptr, 10);
    XorShift128Plus a, b, c;
    a.Seed(seed);
    b.Seed(seed);
    c.Seed(seed);
    for (long i = 0; i < n; ++i) {
        const unsigned long x = a.Int();
        const double r = b.Real();
        const bool p = c.Probability(0.66);
        std::printf("%lu %.17g %d\n", x, r, p ? 1 : 0);
    }
    return 0;
}
