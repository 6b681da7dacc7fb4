/*
 * eigenrand_amd.hpp: an Eigen::Rand-style C++ front end over the C ABI of
 * eigenrand_amd.h, header-only, for C++ callers that do not use the Python
 * package. Names and argument order follow the reference's free functions
 * and generator classes (EigenRand/Dists/ and EigenRand/MvDists/ headers):
 *
 *   Rand::normal<ArrayXXf>(rows, cols, urng)  ->  erd::rand::normal(rows, cols, urng)
 *   Rand::GammaGen<float> gen(a, b); gen.generate(rows, cols, urng)
 *                                             ->  erd::rand::GammaGen gen(a, b, urng.size());
 *                                                 gen.generate(rows, cols, urng)
 *
 * Results live in device memory (DeviceMatrix, column-major like Eigen's
 * default); with an engine set of n engines every call fills n independent
 * rows x cols blocks, block g from engine g. Errors are thrown as
 * erd::rand::Error carrying the erd_status. Streams are passed as void*
 * (a hipStream_t, or nullptr for the default stream); calls are asynchronous
 * on that stream except where they return host data.
 */
#ifndef EIGENRAND_AMD_HPP
#define EIGENRAND_AMD_HPP

#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "eigenrand_amd.h"

namespace erd {
namespace rand {

class Error : public std::runtime_error {
public:
    Error(int status, const char* what)
        : std::runtime_error(std::string(what) + ": " + erd_status_string(status)), status_(status) {}
    int status() const { return status_; }

private:
    int status_;
};

inline void check(int rc, const char* what) {
    if (rc != ERD_OK) throw Error(rc, what);
}

// Owned device buffer of n elements.
template <class T>
class DeviceArray {
public:
    DeviceArray() = default;
    explicit DeviceArray(int64_t n) : n_(n) {
        if (n > 0) check(erd_device_alloc(reinterpret_cast<void**>(&p_), sizeof(T) * (size_t)n), "erd_device_alloc");
    }
    ~DeviceArray() { reset(); }
    DeviceArray(const DeviceArray&) = delete;
    DeviceArray& operator=(const DeviceArray&) = delete;
    DeviceArray(DeviceArray&& o) noexcept : p_(o.p_), n_(o.n_) {
        o.p_ = nullptr;
        o.n_ = 0;
    }
    DeviceArray& operator=(DeviceArray&& o) noexcept {
        if (this != &o) {
            reset();
            p_ = o.p_;
            n_ = o.n_;
            o.p_ = nullptr;
            o.n_ = 0;
        }
        return *this;
    }
    T* data() const { return p_; }
    int64_t size() const { return n_; }
    // Waits for `stream` and copies the buffer to the host.
    std::vector<T> to_host(void* stream = nullptr) const {
        std::vector<T> h((size_t)n_);
        if (n_ > 0) check(erd_memcpy_d2h(h.data(), p_, sizeof(T) * (size_t)n_, stream), "erd_memcpy_d2h");
        return h;
    }
    void reset() {
        if (p_) (void)erd_device_free(p_);
        p_ = nullptr;
        n_ = 0;
    }

private:
    T* p_ = nullptr;
    int64_t n_ = 0;
};

// Result of a fill: `engines` column-major rows x cols blocks, back to back.
template <class T>
struct DeviceMatrix {
    DeviceArray<T> values;
    int64_t rows = 0, cols = 0, engines = 0;
    T* block(int64_t g) const { return values.data() + g * rows * cols; }
};

// A set of Vmt19937_64 engines resident on the device
// (PacketRandomEngine.h:585-599). Engine g is seeded seed + g*stride.
class Vmt19937_64 {
public:
    static constexpr uint64_t default_seed = 5489;

    explicit Vmt19937_64(uint64_t seed = default_seed, int64_t n = 1, uint64_t stride = 4,
                         void* stream = nullptr) {
        check(erd_engines_create_range(&h_, seed, stride, n, stream), "erd_engines_create_range");
    }
    explicit Vmt19937_64(const std::vector<uint64_t>& seeds, void* stream = nullptr) {
        check(erd_engines_create(&h_, seeds.data(), (int64_t)seeds.size(), stream), "erd_engines_create");
    }
    ~Vmt19937_64() {
        if (h_) (void)erd_engines_destroy(h_);
    }
    Vmt19937_64(const Vmt19937_64&) = delete;
    Vmt19937_64& operator=(const Vmt19937_64&) = delete;
    Vmt19937_64(Vmt19937_64&& o) noexcept : h_(o.h_) { o.h_ = nullptr; }

    // `auto c = urng;` of the reference: an independent set in the same state.
    Vmt19937_64 copy(void* stream = nullptr) const {
        Vmt19937_64 c(0, size(), 4, stream);
        check(erd_engines_copy(c.h_, h_, stream), "erd_engines_copy");
        return c;
    }
    void discard(int64_t calls, void* stream = nullptr) {
        check(erd_engines_discard(h_, calls, stream), "erd_engines_discard");
    }
    std::vector<uint8_t> snapshot(void* stream = nullptr) const {
        std::vector<uint8_t> b((size_t)size() * erd_engine_state_bytes());
        check(erd_engines_snapshot(h_, b.data(), stream), "erd_engines_snapshot");
        return b;
    }
    void restore(const std::vector<uint8_t>& b, void* stream = nullptr) {
        if (b.size() != (size_t)size() * erd_engine_state_bytes()) throw Error(ERD_ERR_INVALID, "restore");
        check(erd_engines_restore(h_, b.data(), stream), "erd_engines_restore");
    }
    int64_t size() const { return erd_engines_count(h_); }
    erd_engines* handle() const { return h_; }

protected:
    Vmt19937_64(erd_engines* h, int) : h_(h) {}
    erd_engines* h_ = nullptr;
};

// P8_mt19937_64 adaptors (PacketRandomEngine.h:429-517, 600-609): adaptor k
// is the Vmt pair (seed_k, seed_k + 4).
class P8_mt19937_64 : public Vmt19937_64 {
public:
    explicit P8_mt19937_64(const std::vector<uint64_t>& seeds, void* stream = nullptr)
        : Vmt19937_64(create(seeds, stream), 0) {}
    int64_t adaptors() const { return size() / 2; }

private:
    static erd_engines* create(const std::vector<uint64_t>& seeds, void* stream) {
        erd_engines* h = nullptr;
        check(erd_engines_create_p8(&h, seeds.data(), (int64_t)seeds.size(), stream), "erd_engines_create_p8");
        return h;
    }
};

namespace detail {

template <class T, class F>
DeviceMatrix<T> fill(int64_t rows, int64_t cols, int64_t engines, F&& f, const char* what) {
    DeviceMatrix<T> m;
    m.rows = rows;
    m.cols = cols;
    m.engines = engines;
    m.values = DeviceArray<T>(rows * cols * engines);
    check(f(m.values.data(), rows * cols), what);
    return m;
}

}  // namespace detail

// ------------------------------------------------ free functions (fresh generators)
inline DeviceMatrix<uint64_t> randBits(int64_t draws, Vmt19937_64& urng, void* stream = nullptr) {
    return detail::fill<uint64_t>(4, draws, urng.size(),
                                  [&](uint64_t* o, int64_t n) { return erd_fill_u64(urng.handle(), o, n, stream); },
                                  "erd_fill_u64");
}
inline DeviceMatrix<float> uniformReal(int64_t rows, int64_t cols, Vmt19937_64& urng, float min = 0.f,
                                       float max = 1.f, void* stream = nullptr) {
    return detail::fill<float>(rows, cols, urng.size(), [&](float* o, int64_t n) {
        return erd_fill_uniform_f32(urng.handle(), o, n, min, max, stream);
    }, "erd_fill_uniform_f32");
}
inline DeviceMatrix<double> uniformRealDouble(int64_t rows, int64_t cols, Vmt19937_64& urng, double min = 0.0,
                                              double max = 1.0, void* stream = nullptr) {
    return detail::fill<double>(rows, cols, urng.size(), [&](double* o, int64_t n) {
        return erd_fill_uniform_f64(urng.handle(), o, n, min, max, stream);
    }, "erd_fill_uniform_f64");
}
inline DeviceMatrix<float> normal(int64_t rows, int64_t cols, Vmt19937_64& urng, void* stream = nullptr) {
    return detail::fill<float>(rows, cols, urng.size(), [&](float* o, int64_t n) {
        return erd_fill_std_normal_f32(urng.handle(), o, n, stream);
    }, "erd_fill_std_normal_f32");
}
inline DeviceMatrix<float> normal(int64_t rows, int64_t cols, Vmt19937_64& urng, float mean, float stdev,
                                  void* stream = nullptr) {
    return detail::fill<float>(rows, cols, urng.size(), [&](float* o, int64_t n) {
        return erd_fill_normal_f32(urng.handle(), o, n, mean, stdev, stream);
    }, "erd_fill_normal_f32");
}
// Rand::normal<MatrixXd> / exponential<MatrixXd> / cauchy<MatrixXd>: the
// Packet4d forms (four doubles per engine draw).
inline DeviceMatrix<double> normalDouble(int64_t rows, int64_t cols, Vmt19937_64& urng, void* stream = nullptr) {
    return detail::fill<double>(rows, cols, urng.size(), [&](double* o, int64_t n) {
        return erd_fill_std_normal_f64(urng.handle(), o, n, stream);
    }, "erd_fill_std_normal_f64");
}
inline DeviceMatrix<double> normalDouble(int64_t rows, int64_t cols, Vmt19937_64& urng, double mean, double stdev,
                                         void* stream = nullptr) {
    return detail::fill<double>(rows, cols, urng.size(), [&](double* o, int64_t n) {
        return erd_fill_normal_f64(urng.handle(), o, n, mean, stdev, stream);
    }, "erd_fill_normal_f64");
}
inline DeviceMatrix<double> exponentialDouble(int64_t rows, int64_t cols, Vmt19937_64& urng, double lambda = 1.0,
                                              void* stream = nullptr) {
    return detail::fill<double>(rows, cols, urng.size(), [&](double* o, int64_t n) {
        return erd_fill_exponential_f64(urng.handle(), o, n, lambda, stream);
    }, "erd_fill_exponential_f64");
}
inline DeviceMatrix<double> cauchyDouble(int64_t rows, int64_t cols, Vmt19937_64& urng, double a = 0.0,
                                         double b = 1.0, void* stream = nullptr) {
    return detail::fill<double>(rows, cols, urng.size(), [&](double* o, int64_t n) {
        return erd_fill_cauchy_f64(urng.handle(), o, n, a, b, stream);
    }, "erd_fill_cauchy_f64");
}
inline DeviceMatrix<float> lognormal(int64_t rows, int64_t cols, Vmt19937_64& urng, float mean = 0.f,
                                     float stdev = 1.f, void* stream = nullptr) {
    return detail::fill<float>(rows, cols, urng.size(), [&](float* o, int64_t n) {
        return erd_fill_lognormal_f32(urng.handle(), o, n, mean, stdev, stream);
    }, "erd_fill_lognormal_f32");
}
inline DeviceMatrix<float> weibull(int64_t rows, int64_t cols, Vmt19937_64& urng, float a = 1.f, float b = 1.f,
                                   void* stream = nullptr) {
    return detail::fill<float>(rows, cols, urng.size(), [&](float* o, int64_t n) {
        return erd_fill_weibull_f32(urng.handle(), o, n, a, b, stream);
    }, "erd_fill_weibull_f32");
}
inline DeviceMatrix<float> extremeValue(int64_t rows, int64_t cols, Vmt19937_64& urng, float a = 0.f,
                                        float b = 1.f, void* stream = nullptr) {
    return detail::fill<float>(rows, cols, urng.size(), [&](float* o, int64_t n) {
        return erd_fill_extreme_value_f32(urng.handle(), o, n, a, b, stream);
    }, "erd_fill_extreme_value_f32");
}
inline DeviceMatrix<float> exponential(int64_t rows, int64_t cols, Vmt19937_64& urng, float lambda = 1.f,
                                       void* stream = nullptr) {
    return detail::fill<float>(rows, cols, urng.size(), [&](float* o, int64_t n) {
        return erd_fill_exponential_f32(urng.handle(), o, n, lambda, stream);
    }, "erd_fill_exponential_f32");
}
inline DeviceMatrix<float> cauchy(int64_t rows, int64_t cols, Vmt19937_64& urng, float a = 0.f, float b = 1.f,
                                  void* stream = nullptr) {
    return detail::fill<float>(rows, cols, urng.size(), [&](float* o, int64_t n) {
        return erd_fill_cauchy_f32(urng.handle(), o, n, a, b, stream);
    }, "erd_fill_cauchy_f32");
}
inline DeviceMatrix<float> studentT(int64_t rows, int64_t cols, Vmt19937_64& urng, float n = 1.f,
                                    void* stream = nullptr) {
    return detail::fill<float>(rows, cols, urng.size(), [&](float* o, int64_t k) {
        return erd_fill_studentt_f32(urng.handle(), o, k, n, stream);
    }, "erd_fill_studentt_f32");
}
inline DeviceMatrix<float> gamma(int64_t rows, int64_t cols, Vmt19937_64& urng, float alpha = 1.f,
                                 float beta = 1.f, void* stream = nullptr) {
    return detail::fill<float>(rows, cols, urng.size(), [&](float* o, int64_t n) {
        return erd_fill_gamma_f32(urng.handle(), o, n, alpha, beta, stream);
    }, "erd_fill_gamma_f32");
}
inline DeviceMatrix<float> chiSquared(int64_t rows, int64_t cols, Vmt19937_64& urng, float n = 1.f,
                                      void* stream = nullptr) {
    return detail::fill<float>(rows, cols, urng.size(), [&](float* o, int64_t k) {
        return erd_fill_chisquared_f32(urng.handle(), o, k, n, stream);
    }, "erd_fill_chisquared_f32");
}
inline DeviceMatrix<int32_t> poisson(int64_t rows, int64_t cols, Vmt19937_64& urng, double mean = 1.0,
                                     void* stream = nullptr) {
    return detail::fill<int32_t>(rows, cols, urng.size(), [&](int32_t* o, int64_t n) {
        return erd_fill_poisson_i32(urng.handle(), o, n, mean, stream);
    }, "erd_fill_poisson_i32");
}
inline DeviceMatrix<int32_t> binomial(int64_t rows, int64_t cols, Vmt19937_64& urng, int32_t trials = 1,
                                      double p = 0.5, void* stream = nullptr) {
    return detail::fill<int32_t>(rows, cols, urng.size(), [&](int32_t* o, int64_t n) {
        return erd_fill_binomial_i32(urng.handle(), o, n, trials, p, stream);
    }, "erd_fill_binomial_i32");
}
// Vectorize-over-params forms: per-element parameters in device memory, one
// element per output (n engines x rows*cols).
inline DeviceMatrix<float> normal(Vmt19937_64& urng, const DeviceMatrix<float>& mean,
                                  const DeviceMatrix<float>& stdev, void* stream = nullptr) {
    return detail::fill<float>(mean.rows, mean.cols, urng.size(), [&](float* o, int64_t n) {
        return erd_fill_normal_vop_f32(urng.handle(), o, n, mean.values.data(), stdev.values.data(), stream);
    }, "erd_fill_normal_vop_f32");
}
inline DeviceMatrix<int32_t> binomial(Vmt19937_64& urng, const DeviceMatrix<int32_t>& trials,
                                      const DeviceMatrix<float>& p, void* stream = nullptr) {
    return detail::fill<int32_t>(trials.rows, trials.cols, urng.size(), [&](int32_t* o, int64_t n) {
        return erd_fill_binomial_vop_i32(urng.handle(), o, n, trials.values.data(), p.values.data(), stream);
    }, "erd_fill_binomial_vop_i32");
}

// Raw-bit and small integer generators.
inline DeviceMatrix<int32_t> randBitsInt(int64_t rows, int64_t cols, Vmt19937_64& urng, void* stream = nullptr) {
    return detail::fill<int32_t>(rows, cols, urng.size(), [&](int32_t* o, int64_t n) {
        return erd_fill_randbits_i32(urng.handle(), o, n, stream);
    }, "erd_fill_randbits_i32");
}
inline DeviceMatrix<int32_t> uniformInt(int64_t rows, int64_t cols, Vmt19937_64& urng, int32_t min, int32_t max,
                                        void* stream = nullptr) {
    return detail::fill<int32_t>(rows, cols, urng.size(), [&](int32_t* o, int64_t n) {
        return erd_fill_uniform_int_i32(urng.handle(), o, n, min, max, stream);
    }, "erd_fill_uniform_int_i32");
}
// discreteF<int32>: DiscreteGen<int32_t, float>; weights on the host.
inline DeviceMatrix<int32_t> discreteF(int64_t rows, int64_t cols, Vmt19937_64& urng, const std::vector<double>& w,
                                       void* stream = nullptr) {
    return detail::fill<int32_t>(rows, cols, urng.size(), [&](int32_t* o, int64_t n) {
        return erd_fill_discrete_f_i32(urng.handle(), o, n, w.data(), (int64_t)w.size(), stream);
    }, "erd_fill_discrete_f_i32");
}
inline DeviceMatrix<float> bernoulli(int64_t rows, int64_t cols, Vmt19937_64& urng, double p = 0.5,
                                     void* stream = nullptr) {
    return detail::fill<float>(rows, cols, urng.size(), [&](float* o, int64_t n) {
        return erd_fill_bernoulli_f32(urng.handle(), o, n, p, stream);
    }, "erd_fill_bernoulli_f32");
}
inline DeviceMatrix<int32_t> bernoulliInt(int64_t rows, int64_t cols, Vmt19937_64& urng, double p = 0.5,
                                          void* stream = nullptr) {
    return detail::fill<int32_t>(rows, cols, urng.size(), [&](int32_t* o, int64_t n) {
        return erd_fill_bernoulli_i32(urng.handle(), o, n, p, stream);
    }, "erd_fill_bernoulli_i32");
}
inline DeviceMatrix<float> balanced(int64_t rows, int64_t cols, Vmt19937_64& urng, float a = -1.f, float b = 1.f,
                                    void* stream = nullptr) {
    return detail::fill<float>(rows, cols, urng.size(), [&](float* o, int64_t n) {
        return erd_fill_balanced_f32(urng.handle(), o, n, a, b, stream);
    }, "erd_fill_balanced_f32");
}
inline DeviceMatrix<float> beta(int64_t rows, int64_t cols, Vmt19937_64& urng, float a = 1.f, float b = 1.f,
                                void* stream = nullptr) {
    return detail::fill<float>(rows, cols, urng.size(), [&](float* o, int64_t n) {
        return erd_fill_beta_f32(urng.handle(), o, n, a, b, stream);
    }, "erd_fill_beta_f32");
}
inline DeviceMatrix<float> fisherF(int64_t rows, int64_t cols, Vmt19937_64& urng, float m = 1.f, float n_ = 1.f,
                                   void* stream = nullptr) {
    return detail::fill<float>(rows, cols, urng.size(), [&](float* o, int64_t n) {
        return erd_fill_fisher_f_f32(urng.handle(), o, n, m, n_, stream);
    }, "erd_fill_fisher_f_f32");
}
inline DeviceMatrix<int32_t> negativeBinomial(int64_t rows, int64_t cols, Vmt19937_64& urng, int32_t trials = 1,
                                              double p = 0.5, void* stream = nullptr) {
    return detail::fill<int32_t>(rows, cols, urng.size(), [&](int32_t* o, int64_t n) {
        return erd_fill_negative_binomial_i32(urng.handle(), o, n, trials, p, stream);
    }, "erd_fill_negative_binomial_i32");
}
inline DeviceMatrix<int32_t> geometric(int64_t rows, int64_t cols, Vmt19937_64& urng, double p = 0.5,
                                       void* stream = nullptr) {
    return detail::fill<int32_t>(rows, cols, urng.size(), [&](int32_t* o, int64_t n) {
        return erd_fill_geometric_i32(urng.handle(), o, n, p, stream);
    }, "erd_fill_geometric_i32");
}

// Vectorize-over-params overloads: parameters are device matrices shaped
// like the result (engine g takes block g).
namespace detail {
template <class F>
DeviceMatrix<float> vop(const DeviceMatrix<float>& like, Vmt19937_64& urng, F&& f, const char* what) {
    return fill<float>(like.rows, like.cols, urng.size(), [&](float* o, int64_t n) { return f(o, n); }, what);
}
}  // namespace detail
inline DeviceMatrix<float> uniformReal(Vmt19937_64& urng, const DeviceMatrix<float>& a, const DeviceMatrix<float>& b,
                                       void* stream = nullptr) {
    return detail::vop(a, urng, [&](float* o, int64_t n) {
        return erd_fill_uniform_vop_f32(urng.handle(), o, n, a.values.data(), b.values.data(), stream);
    }, "erd_fill_uniform_vop_f32");
}
inline DeviceMatrix<float> balanced(Vmt19937_64& urng, const DeviceMatrix<float>& a, const DeviceMatrix<float>& b,
                                    void* stream = nullptr) {
    return detail::vop(a, urng, [&](float* o, int64_t n) {
        return erd_fill_balanced_vop_f32(urng.handle(), o, n, a.values.data(), b.values.data(), stream);
    }, "erd_fill_balanced_vop_f32");
}
inline DeviceMatrix<float> bernoulli(Vmt19937_64& urng, const DeviceMatrix<float>& p, void* stream = nullptr) {
    return detail::vop(p, urng, [&](float* o, int64_t n) {
        return erd_fill_bernoulli_vop_f32(urng.handle(), o, n, p.values.data(), stream);
    }, "erd_fill_bernoulli_vop_f32");
}
inline DeviceMatrix<float> exponential(Vmt19937_64& urng, const DeviceMatrix<float>& lambda, void* stream = nullptr) {
    return detail::vop(lambda, urng, [&](float* o, int64_t n) {
        return erd_fill_exponential_vop_f32(urng.handle(), o, n, lambda.values.data(), stream);
    }, "erd_fill_exponential_vop_f32");
}
inline DeviceMatrix<float> cauchy(Vmt19937_64& urng, const DeviceMatrix<float>& a, const DeviceMatrix<float>& b,
                                  void* stream = nullptr) {
    return detail::vop(a, urng, [&](float* o, int64_t n) {
        return erd_fill_cauchy_vop_f32(urng.handle(), o, n, a.values.data(), b.values.data(), stream);
    }, "erd_fill_cauchy_vop_f32");
}
inline DeviceMatrix<float> lognormal(Vmt19937_64& urng, const DeviceMatrix<float>& mean,
                                     const DeviceMatrix<float>& stdev, void* stream = nullptr) {
    return detail::vop(mean, urng, [&](float* o, int64_t n) {
        return erd_fill_lognormal_vop_f32(urng.handle(), o, n, mean.values.data(), stdev.values.data(), stream);
    }, "erd_fill_lognormal_vop_f32");
}
inline DeviceMatrix<float> studentT(Vmt19937_64& urng, const DeviceMatrix<float>& n_, void* stream = nullptr) {
    return detail::vop(n_, urng, [&](float* o, int64_t n) {
        return erd_fill_studentt_vop_f32(urng.handle(), o, n, n_.values.data(), stream);
    }, "erd_fill_studentt_vop_f32");
}
inline DeviceMatrix<float> weibull(Vmt19937_64& urng, const DeviceMatrix<float>& a, const DeviceMatrix<float>& b,
                                   void* stream = nullptr) {
    return detail::vop(a, urng, [&](float* o, int64_t n) {
        return erd_fill_weibull_vop_f32(urng.handle(), o, n, a.values.data(), b.values.data(), stream);
    }, "erd_fill_weibull_vop_f32");
}
inline DeviceMatrix<float> extremeValue(Vmt19937_64& urng, const DeviceMatrix<float>& a,
                                        const DeviceMatrix<float>& b, void* stream = nullptr) {
    return detail::vop(a, urng, [&](float* o, int64_t n) {
        return erd_fill_extreme_value_vop_f32(urng.handle(), o, n, a.values.data(), b.values.data(), stream);
    }, "erd_fill_extreme_value_vop_f32");
}

// --------------------------------------------------- reusable generator objects
// Rand::XxxGen objects: consecutive generate() calls keep the caches the
// reference keeps (normal cos packet, CompressMask leftovers).
template <int Dist, class T>
class GenBase {
public:
    GenBase(double a, double b, int64_t engines, void* stream = nullptr) {
        check(erd_generator_create(&g_, Dist, a, b, engines, stream), "erd_generator_create");
    }
    ~GenBase() {
        if (g_) (void)erd_generator_destroy(g_);
    }
    GenBase(const GenBase&) = delete;
    GenBase& operator=(const GenBase&) = delete;
    DeviceMatrix<T> generate(int64_t rows, int64_t cols, Vmt19937_64& urng, void* stream = nullptr) {
        return detail::fill<T>(rows, cols, urng.size(), [&](T* o, int64_t n) {
            return erd_generator_fill(g_, urng.handle(), o, n, stream);
        }, "erd_generator_fill");
    }
    void reset(void* stream = nullptr) { check(erd_generator_reset(g_, stream), "erd_generator_reset"); }

private:
    erd_generator* g_ = nullptr;
};

struct StdNormalGen : GenBase<ERD_DIST_STD_NORMAL_F32, float> {
    explicit StdNormalGen(int64_t engines = 1) : GenBase(0, 0, engines) {}
};
struct NormalGen : GenBase<ERD_DIST_NORMAL_F32, float> {
    NormalGen(float mean, float stdev, int64_t engines = 1) : GenBase(mean, stdev, engines) {}
};
struct LognormalGen : GenBase<ERD_DIST_LOGNORMAL_F32, float> {
    LognormalGen(float mean, float stdev, int64_t engines = 1) : GenBase(mean, stdev, engines) {}
};
struct GammaGen : GenBase<ERD_DIST_GAMMA_F32, float> {
    GammaGen(float alpha, float beta, int64_t engines = 1) : GenBase(alpha, beta, engines) {}
};
struct PoissonGen : GenBase<ERD_DIST_POISSON_I32, int32_t> {
    explicit PoissonGen(double mean, int64_t engines = 1) : GenBase(mean, 0, engines) {}
};
struct BinomialGen : GenBase<ERD_DIST_BINOMIAL_I32, int32_t> {
    BinomialGen(int32_t trials, double p, int64_t engines = 1) : GenBase(trials, p, engines) {}
};

// ------------------------------------------------------------- multivariate
// Many MvNormalGen<float>(mean_i, cov_i) at once: the FP64 factorisation runs
// at construction (MvNormal.h:86-98); generate() uses instance i on engine i.
// cov: device doubles (column-major d_i x d_i back to back); mean: device
// floats or nullptr.
class MvNormalBatch {
public:
    MvNormalBatch(const std::vector<int32_t>& dims, const double* cov_device, const float* mean_device = nullptr,
                  void* stream = nullptr)
        : dims_(dims), mean_(mean_device) {
        check(erd_mv_factors_create(&f_, (int64_t)dims.size(), dims.data(), cov_device, nullptr, stream),
              "erd_mv_factors_create");
    }
    ~MvNormalBatch() {
        if (f_) (void)erd_mv_factors_destroy(f_);
    }
    MvNormalBatch(const MvNormalBatch&) = delete;
    MvNormalBatch& operator=(const MvNormalBatch&) = delete;
    // X_i (d_i x samples_i) back to back.
    DeviceArray<float> generate(Vmt19937_64& urng, const std::vector<int64_t>& samples, void* stream = nullptr) {
        int64_t total = 0;
        for (size_t i = 0; i < dims_.size(); ++i) total += (int64_t)dims_[i] * samples[i];
        DeviceArray<float> out(total);
        check(erd_mvnormal_generate(urng.handle(), f_, samples.data(), mean_, out.data(), nullptr, stream),
              "erd_mvnormal_generate");
        return out;
    }

private:
    std::vector<int32_t> dims_;
    const float* mean_;
    erd_mv_factors* f_ = nullptr;
};

// Many WishartGen<float>(df_i, scale_i) at once (MvNormal.h:184-302).
// InvWishartGen with inverse = true (factors of the inverse scales).
class WishartBatch {
public:
    WishartBatch(const std::vector<int32_t>& dims, const std::vector<int64_t>& df, const double* scale_device,
                 void* stream = nullptr, bool inverse = false)
        : dims_(dims), df_(df), inverse_(inverse) {
        if (inverse)
            check(erd_mv_factors_create_inverse(&f_, (int64_t)dims.size(), dims.data(), scale_device, nullptr, stream),
                  "erd_mv_factors_create_inverse");
        else
            check(erd_mv_factors_create(&f_, (int64_t)dims.size(), dims.data(), scale_device, nullptr, stream),
                  "erd_mv_factors_create");
    }
    ~WishartBatch() {
        if (f_) (void)erd_mv_factors_destroy(f_);
    }
    WishartBatch(const WishartBatch&) = delete;
    WishartBatch& operator=(const WishartBatch&) = delete;
    DeviceArray<float> generate(Vmt19937_64& urng, const std::vector<int64_t>& samples, void* stream = nullptr) {
        int64_t total = 0;
        for (size_t i = 0; i < dims_.size(); ++i) total += (int64_t)dims_[i] * dims_[i] * samples[i];
        DeviceArray<float> out(total);
        if (inverse_)
            check(erd_invwishart_generate(urng.handle(), f_, df_.data(), samples.data(), out.data(), stream),
                  "erd_invwishart_generate");
        else
            check(erd_wishart_generate(urng.handle(), f_, df_.data(), samples.data(), out.data(), stream),
                  "erd_wishart_generate");
        return out;
    }

private:
    std::vector<int32_t> dims_;
    std::vector<int64_t> df_;
    bool inverse_ = false;
    erd_mv_factors* f_ = nullptr;
};

// DirichletGen<float>(alpha_i).generate for many instances (Multinomial.h:159-217).
inline DeviceArray<float> dirichlet(Vmt19937_64& urng, const std::vector<int32_t>& dims,
                                    const std::vector<int64_t>& samples, const std::vector<float>& alpha_host,
                                    void* stream = nullptr) {
    int64_t total = 0;
    for (size_t i = 0; i < dims.size(); ++i) total += (int64_t)dims[i] * samples[i];
    DeviceArray<float> out(total);
    check(erd_dirichlet_batch(urng.handle(), (int64_t)dims.size(), dims.data(), samples.data(), alpha_host.data(),
                              out.data(), stream),
          "erd_dirichlet_batch");
    return out;
}

// MultinomialGen<int32_t>(trials_i, weights_i).generate for many instances
// (Multinomial.h:25-134); weights are host doubles, K_i each, back to back.
inline DeviceArray<int32_t> multinomial(Vmt19937_64& urng, const std::vector<int64_t>& dims,
                                        const std::vector<int32_t>& trials, const std::vector<int64_t>& samples,
                                        const std::vector<double>& weights_host, void* stream = nullptr) {
    int64_t total = 0;
    for (size_t i = 0; i < dims.size(); ++i) total += dims[i] * samples[i];
    DeviceArray<int32_t> out(total);
    check(erd_multinomial_batch(urng.handle(), (int64_t)dims.size(), dims.data(), trials.data(), samples.data(),
                                weights_host.data(), out.data(), stream),
          "erd_multinomial_batch");
    return out;
}

}  // namespace rand
}  // namespace erd

#endif  // EIGENRAND_AMD_HPP
