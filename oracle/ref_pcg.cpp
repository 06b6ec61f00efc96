// Reference probe (test infrastructure): prints DeterministicRNG streams from the
// reference's own header (include/axiom/math/random.hpp:19-68), compiled where
// the reference tree is mounted.  Output: one line per seed,
// "seed u32_0 .. u32_{n-1} | float_0 .. float_{n-1}" with floats as hex bits.
#include <cinttypes>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "axiom/math/random.hpp"

int main(int argc, char** argv) {
    const int n = argc > 1 ? std::atoi(argv[1]) : 16;
    for (int a = 2; a < argc; ++a) {
        const uint64_t seed = std::strtoull(argv[a], nullptr, 0);
        axiom::math::DeterministicRNG r1(seed), r2(seed);
        std::printf("%" PRIu64, seed);
        for (int i = 0; i < n; ++i) std::printf(" %u", r1.next());
        std::printf(" |");
        for (int i = 0; i < n; ++i) {
            const float f = r2.nextFloat(-2.5f, 7.25f);
            uint32_t u;
            std::memcpy(&u, &f, 4);
            std::printf(" %08x", u);
        }
        std::printf("\n");
    }
    return 0;
}
