// Uses the C++ front end (include/eigenrand_amd.hpp) the way EigenRand code
// uses Eigen::Rand, and writes each result as raw bytes into the directory
// given as argv[1] (one file per result) for the GPU test to check.
#include <cstdio>
#include <string>
#include <vector>

#include "eigenrand_amd.hpp"

namespace er = erd::rand;

template <class T>
static void dump(const std::string& dir, const char* name, const std::vector<T>& v) {
    const std::string path = dir + "/" + name + ".bin";
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) throw std::runtime_error("cannot write " + path);
    std::fwrite(v.data(), sizeof(T), v.size(), f);
    std::fclose(f);
}

int main(int argc, char** argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s out_dir\n", argv[0]);
        return 2;
    }
    const std::string dir = argv[1];
    try {
        er::Vmt19937_64 urng(42);                           // Vmt19937_64 urng{42};
        dump(dir, "raw", er::randBits(256, urng).values.to_host());
        dump(dir, "uniform", er::uniformReal(32, 32, urng).values.to_host());
        dump(dir, "normal", er::normal(64, 64, urng).values.to_host());
        dump(dir, "gamma", er::gamma(128, 1, urng, 3.7f, 2.0f).values.to_host());
        dump(dir, "poisson", er::poisson(128, 1, urng, 20.0).values.to_host());

        er::Vmt19937_64 pair(std::vector<uint64_t>{5, 9});
        er::GammaGen gen(0.4f, 1.5f, pair.size());          // reused generator object
        std::vector<float> g1 = gen.generate(8, 1, pair).values.to_host();
        std::vector<float> g2 = gen.generate(24, 1, pair).values.to_host();
        g1.insert(g1.end(), g2.begin(), g2.end());
        dump(dir, "gamma_gen", g1);

        er::P8_mt19937_64 p8(std::vector<uint64_t>{0});
        er::DeviceArray<uint64_t> d(64);
        er::check(erd_p8_fill_u64(p8.handle(), d.data(), 64, nullptr), "erd_p8_fill_u64");
        dump(dir, "p8", d.to_host());
    } catch (const er::Error& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    std::printf("shim example ok\n");
    return 0;
}
