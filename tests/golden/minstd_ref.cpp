// Prints values drawn with libstdc++'s std::default_random_engine and
// std::uniform_real_distribution<float>(-1, 1) -- the generator of the
// reference's gen_diagdom (tests/genmat.hxx:72-73). Used by the CPU tests to
// pin the oracle's generator to the library itself.
// usage: minstd_ref <seed|-1 for default> <count>
#include <cstdio>
#include <cstdlib>
#include <random>

int main(int argc, char **argv) {
    long seed = argc > 1 ? std::atol(argv[1]) : -1;
    long count = argc > 2 ? std::atol(argv[2]) : 16;
    std::default_random_engine gen;
    if (seed >= 0) gen.seed(static_cast<std::default_random_engine::result_type>(seed));
    std::uniform_real_distribution<float> dist(-1.0, 1.0);
    for (long i = 0; i < count; ++i) {
        float v = dist(gen);
        unsigned u;
        static_assert(sizeof(u) == sizeof(v), "");
        __builtin_memcpy(&u, &v, 4);
        std::printf("%08x\n", u);
    }
    return 0;
}
