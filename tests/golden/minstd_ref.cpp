// Prints values drawn with libstdc++'s std::default_random_engine and
// std::uniform_real_distribution<float> -- the generators of the reference's
// harness (tests/genmat.hxx:72-73, tests/common.hxx:298-319, 391-455). Used by
// the CPU tests to pin the oracle's generators to the library itself.
// usage:
//   minstd_ref <seed|-1 for default> <count>     uniform(-1,1) draws, hex bits
//   minstd_ref uniform <lo> <hi> <count>         uniform(lo,hi), default seed
//   minstd_ref arrowhead <n> <lo> <hi>           FNV-1a of the column-major bits
//   minstd_ref randmat <m> <n> <seed>            FNV-1a of the column-major bits
// The arrowhead and randmat loops below restate the reference's generators
// (k = 256 for the arrowhead) with the real std:: engine and distribution.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

static unsigned bits(float v) {
    unsigned u;
    static_assert(sizeof(u) == sizeof(v), "");
    std::memcpy(&u, &v, 4);
    return u;
}

static uint64_t fnv1a(const std::vector<float>& a) {
    uint64_t h = 1469598103934665603ULL;
    for (float v : a) {
        unsigned u = bits(v);
        for (int b = 0; b < 4; ++b) {
            h ^= (u >> (8 * b)) & 0xffu;
            h *= 1099511628211ULL;
        }
    }
    return h;
}

int main(int argc, char **argv) {
    if (argc > 1 && !std::strcmp(argv[1], "uniform")) {
        float lo = std::atof(argv[2]), hi = std::atof(argv[3]);
        long count = std::atol(argv[4]);
        std::default_random_engine gen;
        std::uniform_real_distribution<float> dist(lo, hi);
        for (long i = 0; i < count; ++i) std::printf("%08x\n", bits(dist(gen)));
        return 0;
    }
    if (argc > 1 && !std::strcmp(argv[1], "arrowhead")) {
        const int n = std::atoi(argv[2]), m = n, k = 256;
        float lo = std::atof(argv[3]), hi = std::atof(argv[4]);
        std::vector<float> a((size_t)n * n, 0.0f);
        std::default_random_engine gen;
        std::uniform_real_distribution<float> dist(lo, hi);
        int nc = 1 + (n - 1) / k;
        for (int c = 0; c < nc; ++c)
            for (int j = c * k; j < std::min(n, (c + 1) * k); ++j)
                for (int i = c * k; i < std::min(m, (c + 1) * k); ++i) {
                    a[(size_t)j * m + i] = dist(gen);
                    a[(size_t)i * m + j] = dist(gen);
                }
        for (int j = 0; j < n; ++j) {
            for (int i = std::max(m - k, j); i < m; ++i) {
                a[(size_t)j * m + i] = dist(gen);
                a[(size_t)i * m + j] = dist(gen);
            }
            a[(size_t)j * m + j] = static_cast<float>(2 * k);
        }
        std::printf("%016llx\n", (unsigned long long)fnv1a(a));
        return 0;
    }
    if (argc > 1 && !std::strcmp(argv[1], "randmat")) {
        const int m = std::atoi(argv[2]), n = std::atoi(argv[3]);
        std::default_random_engine gen(static_cast<std::default_random_engine::result_type>(
            std::atol(argv[4])));
        std::uniform_real_distribution<float> dist(-1.0, 1.0);
        std::vector<float> a((size_t)m * n);
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < m; ++i) a[(size_t)j * m + i] = dist(gen);
        std::printf("%016llx\n", (unsigned long long)fnv1a(a));
        return 0;
    }
    long seed = argc > 1 ? std::atol(argv[1]) : -1;
    long count = argc > 2 ? std::atol(argv[2]) : 16;
    std::default_random_engine gen;
    if (seed >= 0) gen.seed(static_cast<std::default_random_engine::result_type>(seed));
    std::uniform_real_distribution<float> dist(-1.0, 1.0);
    for (long i = 0; i < count; ++i) std::printf("%08x\n", bits(dist(gen)));
    return 0;
}
