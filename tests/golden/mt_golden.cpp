// Writes raw streams of libstdc++'s std::mt19937_64 (C++ standard engine,
// [rand.predef]) as little-endian uint64 to stdout: for each seed in argv,
// `count` draws. Used only to make committed test fixtures.
#include <cstdio>
#include <cstdlib>
#include <cstdint>
#include <random>

int main(int argc, char** argv) {
    if (argc < 3) return 2;
    long long count = std::atoll(argv[1]);
    for (int a = 2; a < argc; ++a) {
        std::mt19937_64 g(std::strtoull(argv[a], nullptr, 10));
        for (long long i = 0; i < count; ++i) {
            uint64_t v = g();
            std::fwrite(&v, 8, 1, stdout);
        }
    }
    return 0;
}
