// Code written the way EigenRand code is written (Eigen::Rand::normal<T>(rows,
// cols, urng), generator objects, vectorize-over-params overloads, the
// multivariate generators), compiled against include/eigenrand_amd/EigenRand
// with a small Eigen-like array type in place of Eigen (not installed here).
// Each result is written as raw bytes into the directory given as argv[1]
// for the GPU test to compare with the Python binding.
#include <cstdio>
#include <string>
#include <vector>

#include "eigenrand_amd/EigenRand"

// The parts of Eigen's dense-object interface the front end relies on.
template <class T>
class Array2 {
public:
    using Scalar = T;
    using PlainObject = Array2;
    Array2() = default;
    Array2(std::ptrdiff_t r, std::ptrdiff_t c) { resize(r, c); }
    void resize(std::ptrdiff_t r, std::ptrdiff_t c) {
        v_.assign((size_t)(r * c), T());
        r_ = r, c_ = c;
    }
    std::ptrdiff_t rows() const { return r_; }
    std::ptrdiff_t cols() const { return c_; }
    std::ptrdiff_t size() const { return r_ * c_; }
    T* data() { return v_.data(); }
    const T* data() const { return v_.data(); }
    T& operator()(std::ptrdiff_t i, std::ptrdiff_t j) { return v_[(size_t)(j * r_ + i)]; }
    const T& operator()(std::ptrdiff_t i, std::ptrdiff_t j) const { return v_[(size_t)(j * r_ + i)]; }
    T& operator()(std::ptrdiff_t i) { return v_[(size_t)i]; }
    const T& operator()(std::ptrdiff_t i) const { return v_[(size_t)i]; }

private:
    std::vector<T> v_;
    std::ptrdiff_t r_ = 0, c_ = 0;
};
using ArrayXXf = Array2<float>;
using ArrayXXd = Array2<double>;
using ArrayXXi = Array2<int32_t>;

template <class M>
static void dump(const std::string& dir, const char* name, const M& m) {
    const std::string path = dir + "/" + name + ".bin";
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) throw std::runtime_error("cannot write " + path);
    std::fwrite(m.data(), sizeof(typename M::Scalar), (size_t)m.size(), f);
    std::fclose(f);
}

int main(int argc, char** argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s out_dir\n", argv[0]);
        return 2;
    }
    const std::string dir = argv[1];
    namespace Rand = Eigen::Rand;
    try {
        Rand::Vmt19937_64 urng{42};
        ArrayXXf x = Rand::normal<ArrayXXf>(64, 32, urng);
        dump(dir, "normal", x);
        dump(dir, "normal_like", Rand::normalLike(x, urng, 2.0f, 0.5f));
        dump(dir, "uniform", Rand::uniformReal<ArrayXXf>(32, 32, urng, -1.0f, 2.0f));
        dump(dir, "gamma", Rand::gamma<ArrayXXf>(128, 1, urng, 3.7f, 2.0f));
        dump(dir, "poisson", Rand::poisson<ArrayXXi>(128, 1, urng, 20.0));
        dump(dir, "binomial", Rand::binomial<ArrayXXi>(8, 8, urng, 100, 0.3));
        dump(dir, "discrete", Rand::discrete<ArrayXXi>(64, 1, urng, {1.0, 2.0, 3.0, 4.0}));
        dump(dir, "normal_f64", Rand::normal<ArrayXXd>(16, 4, urng));
        dump(dir, "chisq", Rand::chiSquared<ArrayXXf>(64, 1, urng, 3.0f));

        // vectorize over parameters: a mean array and a scalar stdev
        ArrayXXf mean(64, 1);
        for (int i = 0; i < 64; ++i) mean(i) = (float)i - 32.0f;
        dump(dir, "normal_vs", Rand::normal(urng, mean, 2.0f));
        ArrayXXi trials(64, 1);
        ArrayXXf p(64, 1);
        for (int i = 0; i < 64; ++i) trials(i) = 10 * i + 1, p(i) = (float)(i + 1) / 65.0f;
        dump(dir, "binomial_vv", Rand::binomial(urng, trials, p));

        // generator objects on two engines: two calls equal one longer fill
        Rand::Vmt19937_64 pair{std::vector<uint64_t>{5, 9}};
        Rand::GammaGen<float> gen{0.4f, 1.5f};
        ArrayXXf g1 = gen.generate<ArrayXXf>(16, 1, pair);
        ArrayXXf g2 = gen.generate<ArrayXXf>(48, 1, pair);
        dump(dir, "gamma_gen1", g1);
        dump(dir, "gamma_gen2", g2);

        // a P8 adaptor as the engine, as the reference's tests draw
        Rand::P8_mt19937_64 p8{42};
        dump(dir, "p8_normal", Rand::normal<ArrayXXf>(24, 1, p8));
        dump(dir, "p8_poisson", Rand::poisson<ArrayXXi>(12, 1, p8, 3.0));

        // a device-resident destination is filled in place
        Rand::Vmt19937_64 dev{7};
        Rand::DeviceMatrix<float> d = Rand::uniformReal<Rand::DeviceMatrix<float>>(16, 16, dev);
        Rand::HostMatrix<float> dh(16, 16);
        std::vector<float> hv = d.to_host();
        std::copy(hv.begin(), hv.end(), dh.data());
        dump(dir, "device_uniform", dh);

        // multivariate generators on engine 0
        Rand::Vmt19937_64 mvu{11};
        ArrayXXf mu(4, 1), cov(4, 4);
        for (int i = 0; i < 4; ++i) {
            mu(i) = (float)i;
            for (int j = 0; j < 4; ++j) cov(i, j) = i == j ? 2.0f : 0.5f;
        }
        auto mv = Rand::makeMvNormalGen(mu, cov);
        dump(dir, "mvnormal", mv.generate(mvu, 64));
        auto wi = Rand::makeWishartGen(6, cov);
        dump(dir, "wishart", wi.generate(mvu, 8));
        ArrayXXf w(5, 1);
        for (int i = 0; i < 5; ++i) w(i) = (float)(i + 1);
        auto mn = Rand::makeMultinomialGen(20, w);
        dump(dir, "multinomial", mn.generate(mvu, 16));
        auto di = Rand::makeDirichletGen(w);
        dump(dir, "dirichlet", di.generate(mvu, 16));
        if (urng.status() != 0) throw Rand::Error(ERD_ERR_ITER_CAP, "status");
    } catch (const Rand::Error& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    std::printf("eigen front end example ok\n");
    return 0;
}
