// Known-answer vectors of the engine and distributions under ompl::RNG
// (util/RandomNumbers.h:66-76, 190-191): std::mt19937, and
// std::uniform_real_distribution<> (libstdc++ generate_canonical<double,53>).
// Built and run by make_golden.py; writes JSON to stdout.
#include <cmath>
#include <cstdio>
#include <random>

int main()
{
    std::printf("{\n");
    {
        std::mt19937 g; // default seed 5489
        std::printf("\"mt_default_first\": [");
        for (int i = 0; i < 32; ++i) std::printf("%s%u", i ? "," : "", (unsigned)g());
        std::printf("],\n");
        std::mt19937 h;
        unsigned v = 0;
        for (int i = 0; i < 10000; ++i) v = (unsigned)h();
        std::printf("\"mt_default_10000th\": %u,\n", v);
    }
    const unsigned seeds[] = {1u, 42u, 1000000000u};
    std::printf("\"uniform01\": {");
    for (int s = 0; s < 3; ++s) {
        std::mt19937 g(seeds[s]);
        std::uniform_real_distribution<> u{0, 1};
        std::printf("%s\"%u\": [", s ? "," : "", seeds[s]);
        for (int i = 0; i < 64; ++i) std::printf("%s%.17g", i ? "," : "", u(g));
        std::printf("]");
    }
    std::printf("},\n");
    // RNG::uniformReal(lo, hi) = (hi - lo) * uniDist_(generator_) + lo over 7 joints in [-pi, pi]
    {
        const double pi = 3.141592653589793238462643383279502884;
        std::mt19937 g(7u);
        std::uniform_real_distribution<> u{0, 1};
        std::printf("\"sample_uniform_seed7_pi\": [");
        for (int i = 0; i < 8 * 7; ++i) std::printf("%s%.17g", i ? "," : "", (pi - -pi) * u(g) + -pi);
        std::printf("],\n");
    }
    // a draw pair whose canonical value rounds to 1: generate_canonical clamps it
    std::printf("\"canonical_max_words\": %.17g\n", std::nextafter(1.0, 0.0));
    std::printf("}\n");
    return 0;
}
