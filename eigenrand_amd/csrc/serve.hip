// Serving runtime: continuous batching of sample requests from many clients.
//
// Every client owns E Vmt19937_64 engines that stay resident in device
// memory for the life of the server (seeded from the client id). Requests are
// queued as they arrive; a launcher thread repeatedly takes the oldest
// admissible requests (their client has no batch in flight, their output fits
// the HBM budget) and turns them into one segment program launch (+ the
// batched L*Z transform for MvNormal requests) on a free stream, while
// earlier batches are still running on other streams. A completion thread
// polls the in-flight batches' events, retires each batch as soon as it has
// finished (in any order), stamps its requests done and frees its stream.
//
// Engines: a request is spread over e consecutive engines of its client
// (about 2^16 normal-equivalents of work per engine, so heavier samplers
// spread wider), taken from a per-client cursor that rotates through the
// client's E engines in submission order. A request may start once no batch
// in flight holds any of its engines; requests on disjoint engines of one
// client therefore run concurrently, and requests sharing an engine run in
// submission order (in one batch, in segment order; across batches, the later
// one waits for the earlier batch to finish).
//
// A request's result depends only on its client's request sequence: which
// engines it uses and their states are fixed by the earlier requests of the
// same client, and clients never share engines. A result is the same whether
// the request runs alone or in a mix.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>
#include "eigenrand_amd.h"
#include "erd_capi_internal.hpp"
#include "erd_device.hpp"
#include "erd_internal.hpp"
#include "erd_kernels.hpp"
#include "erd_mv.hpp"
#include "erd_segment.hpp"

namespace erd {
namespace {

constexpr double kEngineWork = 65536.0;  // default normal-equivalent samples per engine

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

uint64_t splitmix64(uint64_t x) {
    x += 0x9e3779b97f4a7c15ull;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
    return x ^ (x >> 31);
}

int64_t round_up(int64_t v, int64_t m) { return (v + m - 1) / m * m; }

int packet_width(int kind) {
    switch (kind) {
        case ERD_DIST_U64: case ERD_DIST_UNIFORM_F64: case ERD_DIST_POISSON_I32:
        case ERD_DIST_BINOMIAL_I32: return 4;
        default: return 8;
    }
}

size_t elem_bytes(int kind) {
    switch (kind) {
        case ERD_DIST_U64: case ERD_DIST_UNIFORM_F64: return 8;
        default: return 4;
    }
}

// Outputs (samples for MvNormal/multinomial) one engine takes for a request,
// from the relative cost of each sampler: full-chip rates of the fills, as a
// multiple of the normal fill's time per sample (gamma: GS 5, tangent 6;
// Poisson: Knuth 10 + 6 * mean, tangent 7; binomial: n < 25 1 + 1.2 n,
// Knuth 13, tangent 12).
int64_t chunk_for(const erd_serve_request& r, double work) {
    double cost = 1.0;
    switch (r.kind) {
        case ERD_DIST_STUDENT_T_F32: cost = 2.0; break;
        case ERD_DIST_GAMMA_F32: {
            const double a = r.a;
            if (a == 1.0) cost = 1.0;
            else if (a == std::floor(a) && a < 20) cost = a / 2;
            else cost = a < 1 ? 5.0 : 6.0;
            break;
        }
        case ERD_DIST_POISSON_I32: cost = r.a < 12 ? 10.0 + 6.0 * r.a : 7.0; break;
        case ERD_DIST_BINOMIAL_I32: {
            const double n = r.a, m = n * std::min(r.b, 1 - r.b);
            cost = n < 25 ? 1.0 + 1.2 * n : (m < 1 ? 13.0 : 12.0);
            break;
        }
        case ERD_SERVE_MVNORMAL: cost = (double)r.dim; break;
        case ERD_SERVE_MULTINOMIAL: {
            const double K = r.dim, T = r.a;
            cost = T < std::max(30.0 * (K - 1), 100.0) ? 1.0 + T / 8 + K / 8 : 12.0 * K;
            break;
        }
        default: break;
    }
    return std::max<int64_t>(1, (int64_t)(work / std::max(cost, 1e-3)));
}

// An MvNormal model of a client: its covariance (and mean), factorised in
// FP64 on its first request and kept as bf16 hi/lo planes. Registered models
// are named by id; a request that passes the matrix itself uses the model of
// the same (client, contents).
struct Model {
    uint64_t client = 0;
    int d = 0;
    std::vector<double> cov;
    std::vector<float> mean;
    uint16_t* hi = nullptr;  // device planes, set by the launcher thread
    uint16_t* lo = nullptr;
};

struct Req;
void spread(const erd_serve_request& r, int E, int width, double work, Req* q);

struct Req {
    uint64_t ticket = 0;
    erd_serve_request r{};
    std::vector<double> param;  // multinomial weights
    std::vector<float> mean;    // MvNormal mean given with the request
    Model* model = nullptr;     // MvNormal model
    int e_req = 1;              // engines the request is spread over
    int eng0 = 0;               // first of them (client-local, wraps modulo E)
    int64_t per = 0;            // outputs (or columns) per engine
    int64_t out_elems = 0;
    size_t out_bytes = 0;
    void* out = nullptr;
    double t_submit = 0, t_done = 0;
    int state = 0;              // 0 queued, 1 running, 2 done, 3 failed
    int status = ERD_OK;
    int slot = -1;
    bool seed_first = false;    // first request of its client: seeds the engines
    int wclass = 0;             // latency class: batches hold requests of one class
};

struct Batch {
    hipStream_t stream = nullptr;
    int stream_idx = 0;
    hipEvent_t ev = nullptr;
    std::vector<Req*> reqs;
    std::unique_ptr<Scratch> scratch;
    int status = ERD_OK;
    uint32_t serial = 0;
};

// e = ceil(count / chunk) engines (at most E), each taking an equal,
// packet-aligned share.
void spread(const erd_serve_request& r, int E, int width, double work, Req* q) {
    const int64_t chunk = std::max<int64_t>(chunk_for(r, work), width);  // at least one packet
    q->e_req = (int)std::min<int64_t>(E, (r.count + chunk - 1) / chunk);
    q->per = round_up((r.count + q->e_req - 1) / q->e_req, width);
    // Time on one engine, in units of `work` (engines run in parallel, so
    // this is the request's time in a launch): up to 1 when the request has
    // enough engines; more when it needs more than E of them, or when one
    // output alone costs more than `work` (large multinomial rows). Class
    // 0 for up to 1, then one class per factor of 4.
    const double t = (double)q->per / (double)std::max<int64_t>(chunk_for(r, work), 1);
    int c = 0;
    for (double u = 1.0; t > u && c < 16; u *= 4.0) ++c;
    q->wclass = c;
}

// Device blocks for request outputs and batch scratch, recycled by the server
// rather than allocated and freed through the driver for every request (those
// calls serialise the launcher, the completer and the callers' releases on the
// runtime's locks). Sizes are rounded up to 2^k * {4,5,6,7}/4 (at most 25%
// slack). A block returns to its class's free list once its last user is
// done: an output at release(), scratch when its batch has completed. Freed
// blocks beyond `max_cached` bytes go back to the driver.
class BlockPool final : public Recycler {
public:
    explicit BlockPool(size_t max_cached) : max_cached_(max_cached) {}
    ~BlockPool() { drop_cached(); }

    void* get(size_t bytes) override {
        size_t size;
        const int c = size_class(bytes, &size);
        {
            std::lock_guard<std::mutex> lk(mu_);
            if ((size_t)c < free_.size() && !free_[c].empty()) {
                void* p = free_[c].back();
                free_[c].pop_back();
                cached_ -= size;
                return p;
            }
        }
        void* p = nullptr;
        if (hipMalloc(&p, size) != hipSuccess) {
            drop_cached();  // fragmentation: give the cached blocks back and retry once
            if (hipMalloc(&p, size) != hipSuccess) return nullptr;
        }
        std::lock_guard<std::mutex> lk(mu_);
        class_of_[p] = c;
        return p;
    }

    void put(void* p) override {
        if (!p) return;
        std::unique_lock<std::mutex> lk(mu_);
        auto it = class_of_.find(p);
        if (it == class_of_.end()) return;
        const int c = it->second;
        const size_t size = class_size(c);
        if (cached_ + size > max_cached_) {
            class_of_.erase(it);
            lk.unlock();
            (void)hipFree(p);
            return;
        }
        if ((size_t)c >= free_.size()) free_.resize(c + 1);
        free_[c].push_back(p);
        cached_ += size;
    }

private:
    static size_t class_size(int c) { return ((size_t)4 + (c & 3)) << (c >> 2) >> 2; }
    static int size_class(size_t bytes, size_t* size) {
        bytes = std::max<size_t>(bytes, 256);
        int k = 63 - __builtin_clzll(bytes);  // 2^k <= bytes < 2^(k+1)
        for (int m = 0; m < 4; ++m) {
            const int c = 4 * k + m;
            if (class_size(c) >= bytes) {
                *size = class_size(c);
                return c;
            }
        }
        *size = class_size(4 * (k + 1));
        return 4 * (k + 1);
    }
    void drop_cached() {
        std::vector<void*> drop;
        {
            std::lock_guard<std::mutex> lk(mu_);
            for (auto& f : free_) {
                for (void* p : f) {
                    class_of_.erase(p);
                    drop.push_back(p);
                }
                f.clear();
            }
            cached_ = 0;
        }
        for (void* p : drop) (void)hipFree(p);
    }

    std::mutex mu_;
    std::unordered_map<void*, int> class_of_;  // every block handed out or cached
    std::vector<std::vector<void*>> free_;     // cached blocks per class
    size_t cached_ = 0;
    size_t max_cached_;
};

}  // namespace

struct Server {
    erd_server_config cfg{};
    erd_engines* pool = nullptr;
    std::unique_ptr<BlockPool> blocks;  // outputs and batch scratch
    std::mutex mu;
    std::condition_variable cv;
    std::deque<Req*> pending;
    std::unordered_map<uint64_t, std::unique_ptr<Req>> reqs;
    std::unordered_map<uint64_t, int> slot_of;  // client -> engine slot
    std::vector<uint64_t> slot_client;
    std::vector<char> slot_seeded;
    std::vector<int> slot_cursor;                // next engine of the client
    std::vector<uint32_t> owner;                 // pool engine -> batch holding it (0: none)
    uint32_t next_serial = 1;
    std::vector<hipStream_t> streams;
    std::vector<char> stream_free;
    std::vector<void*> pin;        // per stream: pinned staging buffer
    std::vector<size_t> pin_cap;
    uint32_t* pin_status = nullptr;  // per stream: the last batch's device status bits (pinned)
    std::deque<Batch*> inflight;
    std::unordered_map<uint64_t, std::unique_ptr<Model>> models;     // id -> model
    std::map<std::pair<uint64_t, uint64_t>, uint64_t> model_by_hash;  // (client, contents) -> id
    uint64_t next_model = 1;
    double prep_s = 0;  // launcher time spent building and launching batches
    int64_t budget_used = 0;
    uint64_t next_ticket = 1;
    bool stop = false;
    // statistics
    std::vector<double> latencies;
    int64_t done_requests = 0, done_elems = 0, batches = 0;
    double t_first_submit = 0, t_last_done = 0;
    std::thread launcher, completer;

    int slot_for(uint64_t client) {
        auto it = slot_of.find(client);
        if (it != slot_of.end()) return it->second;
        if ((int64_t)slot_client.size() >= cfg.max_clients) return -1;
        const int s = (int)slot_client.size();
        slot_client.push_back(client);
        slot_seeded.push_back(0);
        slot_cursor.push_back(0);
        slot_of[client] = s;
        return s;
    }

    // Pool index of engine g of request q.
    int32_t engine_of(const Req* q, int g) const {
        const int E = cfg.engines_per_client;
        return q->slot * E + (q->eng0 + g) % E;
    }
    // Engines a request holds while its batch runs: all of the client's when
    // it seeds them, else its own e_req.
    template <class F>
    void for_held(const Req* q, F f) const {
        const int E = cfg.engines_per_client;
        if (q->seed_first) {
            for (int g = 0; g < E; ++g) f(q->slot * E + g);
        } else {
            for (int g = 0; g < q->e_req; ++g) f(engine_of(q, g));
        }
    }
    int prepare_and_launch(Batch* b);
    void launcher_loop();
    void completer_loop();
};

namespace {

uint64_t hash_params(const std::vector<double>& v) {
    uint64_t h = 0x12345678ull;
    for (double x : v) {
        uint64_t u;
        std::memcpy(&u, &x, 8);
        h = splitmix64(h ^ u);
    }
    return h;
}

}  // namespace

// Host data of one batch (seeds, discrete tables, means, the segment program
// and the L*Z tile queues), packed into one pinned buffer and uploaded with a
// single copy. Pointers into it are patched once the device block exists.
struct Stage {
    std::vector<char> host;
    // 256-byte aligned space for `bytes`, zero-filled.
    size_t reserve(size_t bytes) {
        const size_t off = (host.size() + 255) & ~(size_t)255;
        host.resize(off + bytes);
        return off;
    }
    size_t put(const void* p, size_t bytes) {
        const size_t off = reserve(bytes);
        if (bytes) std::memcpy(host.data() + off, p, bytes);
        return off;
    }
    template <class T>
    size_t put(const std::vector<T>& v) {
        return v.empty() ? reserve(0) : put(v.data(), sizeof(T) * v.size());
    }
};

// A staged pointer to patch: segment `seg`'s in0/in1 (field 0/1) or MvNormal
// instance `seg`'s mean (field 2) = device block + off.
struct Reloc {
    size_t seg;
    int field;
    size_t off;
};

// Builds and launches one batch (called without the server lock). Only the
// first use of a covariance synchronises (its FP64 factorisation must be
// checked); everything else is enqueued on the batch's stream.
int Server::prepare_and_launch(Batch* b) {
    hipStream_t s = b->stream;
    b->scratch.reset(new Scratch(s, blocks.get()));
    Scratch& sc = *b->scratch;
    const int E = cfg.engines_per_client;
    Stage st;
    std::vector<Reloc> relocs;
    // This batch's own status word (zeroed by the upload), so that a cap hit
    // is traced to its batch.
    const size_t off_status = st.reserve(sizeof(uint32_t));
    // Engines of clients seen for the first time:
    // engine g of client c = Vmt19937_64(splitmix64(c ^ salt) + 4g).
    std::vector<std::pair<int, size_t>> seed_jobs;  // (slot, staged seeds)
    for (Req* q : b->reqs) {
        if (!q->seed_first) continue;
        std::vector<uint64_t> seeds(E);
        const uint64_t base = splitmix64(q->r.client ^ cfg.seed_salt);
        for (int g = 0; g < E; ++g) seeds[g] = base + 4ull * (uint64_t)g;
        seed_jobs.emplace_back(q->slot, st.put(seeds));
    }
    std::vector<Segment> segs;
    std::vector<int32_t> seg_engine;
    std::vector<MvInstance> mv;
    for (Req* q : b->reqs) {
        const erd_serve_request& r = q->r;
        if (!(q->out = blocks->get(q->out_bytes))) return ERD_ERR_OOM;
        if (r.kind == ERD_SERVE_MVNORMAL) {
            Model& m = *q->model;
            const int d = m.d;
            if (!m.hi) {  // first request of the model: FP64 factorisation
                double* a;
                if (sc.alloc(&a, (size_t)d * d) != hipSuccess) return ERD_ERR_OOM;
                if (hipMemcpyAsync(a, m.cov.data(), sizeof(double) * d * d, hipMemcpyHostToDevice, s) !=
                        hipSuccess ||
                    hipStreamSynchronize(s) != hipSuccess)
                    return ERD_ERR_HIP;
                uint16_t *hi, *lo;
                float* l32;
                std::vector<int64_t> so;
                int rc = factor_and_split({d}, a, {0, (int64_t)d * d}, &hi, &lo, &l32, so, nullptr, s, sc);
                if (rc) return rc;
                const int dp = (d + 15) & ~15;
                uint16_t *fhi = nullptr, *flo = nullptr;
                if (hipMalloc(&fhi, sizeof(uint16_t) * dp * dp) != hipSuccess) return ERD_ERR_OOM;
                if (hipMalloc(&flo, sizeof(uint16_t) * dp * dp) != hipSuccess) {
                    (void)hipFree(fhi);
                    return ERD_ERR_OOM;
                }
                // The planes are complete before they are published: a later
                // request of the model may run on another stream.
                if (hipMemcpyAsync(fhi, hi, sizeof(uint16_t) * dp * dp, hipMemcpyDeviceToDevice, s) != hipSuccess ||
                    hipMemcpyAsync(flo, lo, sizeof(uint16_t) * dp * dp, hipMemcpyDeviceToDevice, s) != hipSuccess ||
                    hipStreamSynchronize(s) != hipSuccess) {
                    (void)hipFree(fhi);
                    (void)hipFree(flo);
                    return ERD_ERR_HIP;
                }
                m.hi = fhi;
                m.lo = flo;
            }
            float* z;
            if (sc.alloc(&z, (size_t)q->out_elems) != hipSuccess) return ERD_ERR_OOM;
            for (int g = 0; g < q->e_req; ++g) {
                Segment sg{};
                sg.kind = SEG_STD_NORMAL_F32;
                sg.count = (int64_t)d * q->per;
                sg.out = z + (int64_t)g * d * q->per;
                segs.push_back(sg);
                seg_engine.push_back(engine_of(q, g));
            }
            const std::vector<float>& mean = q->mean.empty() ? m.mean : q->mean;
            if (!mean.empty()) relocs.push_back({mv.size(), 2, st.put(mean)});
            mv.push_back(MvInstance{d, 0, q->per * q->e_req, m.hi, m.lo, z, (float*)q->out, nullptr});
        } else if (r.kind == ERD_SERVE_MULTINOMIAL) {
            const int64_t K = r.dim;
            const int32_t T = (int32_t)r.a;
            int32_t* o = (int32_t*)q->out;
            const bool small = T < std::max<int64_t>(30 * (K - 1), 100);
            if (small) {
                HostTable t;
                int rc = build_table(q->param.data(), K, t);
                if (rc) return rc;
                if (hipMemsetAsync(o, 0, q->out_bytes, s) != hipSuccess) return ERD_ERR_HIP;
                Segment proto{};
                proto.kind = SEG_MULTINOMIAL_SMALL;
                proto.i0 = T;
                proto.d[0] = (double)K;
                size_t in0, in1 = 0;
                if (!t.cdf.empty()) {
                    in0 = st.put(t.cdf);
                    proto.d[1] = (double)t.cdf.size();
                } else {
                    in0 = st.put(t.prob);
                    in1 = st.put(t.alias);
                    proto.d[2] = (double)t.bitmask;
                }
                for (int g = 0; g < q->e_req; ++g) {
                    Segment sg = proto;
                    sg.count = q->per;
                    sg.out = o + (int64_t)g * K * q->per;
                    relocs.push_back({segs.size(), 0, in0});
                    if (t.cdf.empty()) relocs.push_back({segs.size(), 1, in1});
                    segs.push_back(sg);
                    seg_engine.push_back(engine_of(q, g));
                }
                continue;
            }
            int32_t* rr;  // remaining trials per sample, written by the first row
            if (sc.alloc(&rr, (size_t)q->per * q->e_req) != hipSuccess) return ERD_ERR_OOM;
            const size_t in1 = st.put(multinomial_row_p(q->param.data(), K));
            for (int g = 0; g < q->e_req; ++g) {
                Segment sg{};
                sg.kind = SEG_MULTINOMIAL_ROWS;
                sg.i0 = T;
                sg.count = q->per;
                sg.out = o + (int64_t)g * K * q->per;
                sg.in0 = rr + (int64_t)g * q->per;
                sg.d[1] = (double)K;
                relocs.push_back({segs.size(), 1, in1});
                segs.push_back(sg);
                seg_engine.push_back(engine_of(q, g));
            }
        } else {
            for (int g = 0; g < q->e_req; ++g) {
                erd_request er{};
                er.dist = r.kind;
                er.engine = engine_of(q, g);
                er.count = q->per;
                er.out = (char*)q->out + (size_t)g * q->per * elem_bytes(r.kind);
                er.a = r.a;
                er.b = r.b;
                Segment sg;
                int rc = make_segment(er, sg);
                if (rc) return rc;
                segs.push_back(sg);
                seg_engine.push_back(er.engine);
            }
        }
    }
    // L*Z plan (tuned variants); an untuned dimension bucket falls back to
    // run_lz, which tunes it with synchronised trial launches.
    std::vector<LzLaunch> plan;
    const bool staged_lz = mv.empty() || plan_lz(mv, plan) == ERD_OK;
    // Layout of the rest of the block: program, instances, tile queues.
    std::vector<size_t> order(segs.size());
    for (size_t i = 0; i < order.size(); ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(),
                     [&](size_t x, size_t y) { return seg_engine[x] < seg_engine[y]; });
    std::vector<int32_t> engine_list, begin;
    for (size_t k = 0; k < order.size(); ++k) {
        const int32_t eg = seg_engine[order[k]];
        if (engine_list.empty() || engine_list.back() != eg) {
            engine_list.push_back(eg);
            begin.push_back((int32_t)k);
        }
    }
    begin.push_back((int32_t)order.size());
    const size_t off_segs = st.reserve(sizeof(Segment) * segs.size());
    const size_t off_eng = st.put(engine_list);
    const size_t off_begin = st.put(begin);
    const size_t off_mv = st.reserve(sizeof(MvInstance) * mv.size());
    std::vector<size_t> off_tiles, off_qbeg, off_cnt;
    if (staged_lz) {
        const int32_t zero[kLzQueues] = {};
        for (const LzLaunch& l : plan) {
            off_tiles.push_back(st.put(l.tiles));
            off_qbeg.push_back(st.put(l.qbeg));
            off_cnt.push_back(st.put(zero, sizeof(zero)));
        }
    }
    char* dev;
    if (sc.alloc(&dev, st.host.size()) != hipSuccess) return ERD_ERR_OOM;
    for (const Reloc& rl : relocs) {
        void* ptr = dev + rl.off;
        if (rl.field == 0) segs[rl.seg].in0 = ptr;
        else if (rl.field == 1) segs[rl.seg].in1 = ptr;
        else mv[rl.seg].mean = (const float*)ptr;
    }
    Segment* sorted = reinterpret_cast<Segment*>(st.host.data() + off_segs);
    for (size_t k = 0; k < order.size(); ++k) sorted[k] = segs[order[k]];
    if (!mv.empty()) std::memcpy(st.host.data() + off_mv, mv.data(), sizeof(MvInstance) * mv.size());
    // Pinned copy of the block (this stream slot's buffer; the slot is idle).
    const int si = b->stream_idx;
    if (pin_cap[si] < st.host.size()) {
        if (pin[si]) (void)hipHostFree(pin[si]);
        pin[si] = nullptr;
        pin_cap[si] = 0;
        const size_t cap = std::max(st.host.size(), (size_t)1 << 20) * 2;
        if (hipHostMalloc(&pin[si], cap, hipHostMallocDefault) != hipSuccess) return ERD_ERR_OOM;
        pin_cap[si] = cap;
    }
    std::memcpy(pin[si], st.host.data(), st.host.size());
    if (hipMemcpyAsync(dev, pin[si], st.host.size(), hipMemcpyHostToDevice, s) != hipSuccess) return ERD_ERR_HIP;
    for (auto& j : seed_jobs)
        if (launch_seed(pool->states + (int64_t)j.first * E, (const uint64_t*)(dev + j.second), E, s) !=
            hipSuccess)
            return ERD_ERR_HIP;
    if (!segs.empty() &&
        launch_program(pool->states, (int32_t)engine_list.size(), (const int32_t*)(dev + off_eng),
                       (const int32_t*)(dev + off_begin), (const Segment*)(dev + off_segs),
                       Guard{(uint32_t*)(dev + off_status),
                             cfg.reject_cap > 0 ? cfg.reject_cap : kDefaultRejectCap,
                             cfg.lockstep_cap > 0 ? cfg.lockstep_cap : kDefaultLockstepCap},
                       s) != hipSuccess)
        return ERD_ERR_HIP;
    if (!mv.empty()) {
        MvInstance* d_inst = (MvInstance*)(dev + off_mv);
        if (staged_lz) {
            for (size_t k = 0; k < plan.size(); ++k)
                if (launch_lz_gemm(plan[k].variant, plan[k].grid, (const MvTile*)(dev + off_tiles[k]),
                                   (const int32_t*)(dev + off_qbeg[k]), (int32_t*)(dev + off_cnt[k]), d_inst,
                                   s) != hipSuccess)
                    return ERD_ERR_HIP;
        } else {
            const int rc = run_lz(mv, d_inst, s, sc);
            if (rc) return rc;
        }
    }
    // the status word comes back with the batch (read once its event is done)
    if (hipMemcpyAsync(pin_status + 16 * si, dev + off_status, sizeof(uint32_t), hipMemcpyDeviceToHost, s) !=
        hipSuccess)
        return ERD_ERR_HIP;
    if (hipEventRecord(b->ev, s) != hipSuccess) return ERD_ERR_HIP;
    return ERD_OK;
}

void Server::launcher_loop() {
    for (;;) {
        Batch* b = nullptr;
        {
            std::unique_lock<std::mutex> lk(mu);
            for (;;) {
                if (stop) return;
                int free_idx = -1;
                for (size_t i = 0; i < stream_free.size(); ++i)
                    if (stream_free[i]) { free_idx = (int)i; break; }
                if (free_idx >= 0 && !pending.empty()) {
                    // Oldest admissible requests; a client blocked once in this
                    // pass stays blocked so that its order is kept.
                    // A batch holds requests of one latency class (that of its
                    // oldest admissible request): a request completes with
                    // its batch, so a short request must not wait for a long
                    // one; the other classes go to the next launch.
                    std::vector<char> blocked(slot_client.size() + 1, 0);
                    std::vector<Req*> take;
                    int64_t take_elems = 0;
                    const uint32_t serial = next_serial;
                    for (auto it = pending.begin(); it != pending.end();) {
                        Req* q = *it;
                        if ((int64_t)take.size() >= cfg.max_batch_requests) break;
                        const int sl = q->slot;
                        if (!take.empty() && q->wclass != take.front()->wclass) {
                            blocked[sl] = 1;  // the client's later requests stay behind it
                            ++it;
                            continue;
                        }
                        if (!take.empty() && cfg.max_batch_elements > 0 &&
                            take_elems + q->out_elems > cfg.max_batch_elements)
                            break;
                        bool ok = !blocked[sl] && budget_used + (int64_t)q->out_bytes <= cfg.hbm_budget_bytes;
                        if (ok) {
                            q->seed_first = !slot_seeded[sl];
                            for_held(q, [&](int32_t g) { ok = ok && (owner[g] == 0 || owner[g] == serial); });
                        }
                        if (!ok) {
                            q->seed_first = false;
                            blocked[sl] = 1;  // keeps the client's later requests behind it
                            ++it;
                            continue;
                        }
                        slot_seeded[sl] = 1;
                        for_held(q, [&](int32_t g) { owner[g] = serial; });
                        budget_used += (int64_t)q->out_bytes;
                        take_elems += q->out_elems;
                        q->state = 1;
                        take.push_back(q);
                        it = pending.erase(it);
                    }
                    if (!take.empty()) {
                        ++next_serial;
                        b = new Batch();
                        b->serial = serial;
                        b->reqs = take;
                        b->stream_idx = free_idx;
                        b->stream = streams[free_idx];
                        stream_free[free_idx] = 0;
                        break;
                    }
                }
                cv.wait(lk);
            }
        }
        const double t_prep = now_s();
        b->status = hipEventCreateWithFlags(&b->ev, hipEventDisableTiming) == hipSuccess ? prepare_and_launch(b)
                                                                                      : ERD_ERR_HIP;
        const double dt_prep = now_s() - t_prep;
        {
            std::lock_guard<std::mutex> lk(mu);
            inflight.push_back(b);
            ++batches;
            prep_s += dt_prep;
        }
        cv.notify_all();
    }
}

// Batches on different streams finish in any order: each is retired as soon
// as its event has completed, so a long batch holds only its own stream and
// engines, not the retirement of the batches launched after it.
void Server::completer_loop() {
    std::vector<Batch*> snap;
    for (;;) {
        Batch* b = nullptr;
        {
            std::unique_lock<std::mutex> lk(mu);
            cv.wait(lk, [&] { return stop || !inflight.empty(); });
            if (inflight.empty()) return;
            snap.assign(inflight.begin(), inflight.end());
        }
        for (Batch* c : snap) {
            if (c->status != ERD_OK) {
                (void)hipStreamSynchronize(c->stream);
                b = c;
                break;
            }
            const hipError_t q = hipEventQuery(c->ev);
            if (q == hipSuccess) {
                // kernel-raised errors of this batch (its own status word)
                const uint32_t bits = reinterpret_cast<volatile uint32_t*>(pin_status)[16 * c->stream_idx];
                if (bits & kStatusIterCap) c->status = ERD_ERR_ITER_CAP;
                else if (bits) c->status = ERD_ERR_INVALID;
                b = c;
                break;
            }
            if (q != hipErrorNotReady) {
                c->status = ERD_ERR_HIP;
                b = c;
                break;
            }
        }
        if (!b) {
            std::this_thread::sleep_for(std::chrono::microseconds(10));
            continue;
        }
        const double t = now_s();
        b->scratch.reset();  // stream-ordered frees
        if (b->ev) (void)hipEventDestroy(b->ev);
        {
            std::lock_guard<std::mutex> lk(mu);
            inflight.erase(std::find(inflight.begin(), inflight.end(), b));
            for (Req* q : b->reqs) {
                q->t_done = t;
                q->state = b->status == ERD_OK ? 2 : 3;
                q->status = b->status;
                if (b->status != ERD_OK) {
                    // A failed request returns no output: its block (which may
                    // hold another request's earlier data where the sampler did
                    // not write) goes back to the pool. Engines it was to seed
                    // are seeded again by the client's next request.
                    blocks->put(q->out);
                    q->out = nullptr;
                    if (q->seed_first) slot_seeded[q->slot] = 0;
                }
                for_held(q, [&](int32_t g) {
                    if (owner[g] == b->serial) owner[g] = 0;
                });
                if (b->status == ERD_OK) {
                    latencies.push_back(t - q->t_submit);
                    ++done_requests;
                    done_elems += q->r.kind == ERD_SERVE_MVNORMAL || q->r.kind == ERD_SERVE_MULTINOMIAL
                                      ? (int64_t)q->r.dim * q->r.count
                                      : q->r.count;
                }
            }
            t_last_done = t;
            stream_free[b->stream_idx] = 1;
        }
        delete b;
        cv.notify_all();
    }
}

}  // namespace erd

using erd::Req;
using erd::Server;

struct erd_server {
    Server s;
};

extern "C" {

int erd_server_create(erd_server** out, const erd_server_config* cfg) {
    if (!out || !cfg || cfg->max_clients <= 0 || cfg->engines_per_client <= 0 || cfg->max_inflight <= 0 ||
        cfg->hbm_budget_bytes <= 0 || cfg->max_batch_requests <= 0)
        return ERD_ERR_INVALID;
    erd_server* srv = new (std::nothrow) erd_server();
    if (!srv) return ERD_ERR_OOM;
    Server& S = srv->s;
    S.cfg = *cfg;
    S.owner.assign((size_t)(cfg->max_clients * cfg->engines_per_client), 0);
    // Outputs and scratch are recycled blocks; up to the HBM budget of freed
    // blocks stays cached, so steady-state requests allocate nothing.
    S.blocks.reset(new erd::BlockPool((size_t)std::max<int64_t>(cfg->hbm_budget_bytes, 0)));
    int rc = erd_engines_create_range(&S.pool, 0, 4, cfg->max_clients * cfg->engines_per_client, nullptr);
    if (rc) {
        delete srv;
        return rc;
    }
    for (int i = 0; i < cfg->max_inflight; ++i) {
        hipStream_t st;
        if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
            erd_engines_destroy(S.pool);
            delete srv;
            return ERD_ERR_HIP;
        }
        S.streams.push_back(st);
        S.stream_free.push_back(1);
        S.pin.push_back(nullptr);
        S.pin_cap.push_back(0);
    }
    if (hipHostMalloc((void**)&S.pin_status, sizeof(uint32_t) * 16 * (size_t)cfg->max_inflight,
                      hipHostMallocDefault) != hipSuccess) {
        for (hipStream_t st : S.streams) (void)hipStreamDestroy(st);
        erd_engines_destroy(S.pool);
        delete srv;
        return ERD_ERR_OOM;
    }
    S.launcher = std::thread([&S] { S.launcher_loop(); });
    S.completer = std::thread([&S] { S.completer_loop(); });
    *out = srv;
    return ERD_OK;
}

int erd_server_register_mvnormal(erd_server* srv, uint64_t client, int32_t d, const double* cov,
                                 const float* mean, uint64_t* model) {
    if (!srv || !cov || !model || d <= 0) return ERD_ERR_INVALID;
    Server& S = srv->s;
    std::unique_ptr<erd::Model> m(new erd::Model());
    m->client = client;
    m->d = d;
    m->cov.assign(cov, cov + (size_t)d * d);
    for (double v : m->cov)
        if (!std::isfinite(v)) return ERD_ERR_NONFINITE;
    if (mean) m->mean.assign(mean, mean + d);
    // A matrix passed without a mean is looked up by contents, so that
    // repeated requests with the same covariance share one factorisation.
    const uint64_t key = mean ? 0 : erd::hash_params(m->cov) ^ (uint64_t)d;
    std::lock_guard<std::mutex> lk(S.mu);
    if (!mean) {
        auto it = S.model_by_hash.find(std::make_pair(client, key));
        if (it != S.model_by_hash.end()) {
            *model = it->second;
            return ERD_OK;
        }
    }
    const uint64_t id = S.next_model++;
    S.models[id] = std::move(m);
    if (!mean) S.model_by_hash[std::make_pair(client, key)] = id;
    *model = id;
    return ERD_OK;
}

int erd_server_submit(erd_server* srv, const erd_serve_request* r, uint64_t* ticket) {
    if (!srv || !r || !ticket || r->count <= 0) return ERD_ERR_INVALID;
    Server& S = srv->s;
    std::unique_ptr<Req> q(new Req());
    q->r = *r;
    const int E = S.cfg.engines_per_client;
    const double work = S.cfg.engine_work > 0 ? (double)S.cfg.engine_work : erd::kEngineWork;
    if (r->kind == ERD_SERVE_MVNORMAL) {
        int d = r->dim;
        if (r->model) {
            std::lock_guard<std::mutex> lk(S.mu);
            auto it = S.models.find(r->model);
            if (it == S.models.end() || it->second->client != r->client) return ERD_ERR_INVALID;
            q->model = it->second.get();
            if (d == 0) d = q->model->d;
            if (d != q->model->d) return ERD_ERR_INVALID;
        } else {
            if (d <= 0 || !r->host_param) return ERD_ERR_INVALID;
            uint64_t id = 0;
            int rc = erd_server_register_mvnormal(srv, r->client, d, r->host_param, nullptr, &id);
            if (rc) return rc;
            std::lock_guard<std::mutex> lk(S.mu);
            q->model = S.models[id].get();
        }
        q->r.dim = d;
        if (r->host_mean) q->mean.assign(r->host_mean, r->host_mean + d);
        erd::spread(q->r, E, 8, work, q.get());
        q->out_elems = (int64_t)d * q->per * q->e_req;
        q->out_bytes = sizeof(float) * (size_t)q->out_elems;
    } else if (r->kind == ERD_SERVE_MULTINOMIAL) {
        // trials: an integer in [0, INT32_MAX] (it is cast to int32 at launch)
        if (r->dim <= 0 || !r->host_param || !(r->a >= 0.0 && r->a <= 2147483647.0) ||
            std::floor(r->a) != r->a)
            return ERD_ERR_INVALID;
        q->param.assign(r->host_param, r->host_param + r->dim);
        double sum = 0;
        for (double v : q->param) {
            if (v < 0) return ERD_ERR_INVALID;
            sum += v;
        }
        if (!std::isfinite(sum)) return ERD_ERR_NONFINITE;
        if (!(sum > 0)) return ERD_ERR_INVALID;
        erd::spread(*r, E, 8, work, q.get());
        q->out_elems = (int64_t)r->dim * q->per * q->e_req;
        q->out_bytes = sizeof(int32_t) * (size_t)q->out_elems;
    } else {
        const bool scalar = (r->kind >= 0 && r->kind <= ERD_DIST_BINOMIAL_I32 && r->kind != ERD_DIST_NORMAL_VOP_F32) ||
                            r->kind == ERD_DIST_LOGNORMAL_F32 || r->kind == ERD_DIST_WEIBULL_F32 ||
                            r->kind == ERD_DIST_EXTREME_VALUE_F32;
        if (!scalar) return ERD_ERR_INVALID;  // per-element parameter kinds are not served
        erd_request probe{};
        probe.dist = r->kind;
        probe.count = erd::packet_width(r->kind);
        probe.out = (void*)16;
        probe.a = r->a;
        probe.b = r->b;
        erd::Segment tmp;
        int rc = erd::make_segment(probe, tmp);
        if (rc) return rc;
        erd::spread(*r, E, erd::packet_width(r->kind), work, q.get());
        q->out_elems = q->per * q->e_req;
        q->out_bytes = erd::elem_bytes(r->kind) * (size_t)q->out_elems;
    }
    if ((int64_t)q->out_bytes > S.cfg.hbm_budget_bytes) return ERD_ERR_OOM;
    {
        std::lock_guard<std::mutex> lk(S.mu);
        const int sl = S.slot_for(r->client);
        if (sl < 0) return ERD_ERR_UNSUPPORTED;  // engine slots exhausted
        q->slot = sl;
        q->eng0 = S.slot_cursor[sl];
        S.slot_cursor[sl] = (q->eng0 + q->e_req) % E;
        q->ticket = S.next_ticket++;
        q->t_submit = erd::now_s();
        if (S.t_first_submit == 0) S.t_first_submit = q->t_submit;
        *ticket = q->ticket;
        S.pending.push_back(q.get());
        S.reqs[q->ticket] = std::move(q);
    }
    S.cv.notify_all();
    return ERD_OK;
}

int erd_server_wait(erd_server* srv, uint64_t ticket, double timeout_s, void** out_ptr, int64_t* out_elems) {
    if (!srv) return ERD_ERR_INVALID;
    Server& S = srv->s;
    std::unique_lock<std::mutex> lk(S.mu);
    auto it = S.reqs.find(ticket);
    if (it == S.reqs.end()) return ERD_ERR_INVALID;
    Req* q = it->second.get();
    auto pred = [&] { return q->state >= 2; };
    if (timeout_s < 0) {
        S.cv.wait(lk, pred);
    } else if (!S.cv.wait_for(lk, std::chrono::duration<double>(timeout_s), pred)) {
        return ERD_ERR_UNSUPPORTED;  // not done yet
    }
    if (out_ptr) *out_ptr = q->out;
    if (out_elems) *out_elems = q->out_elems;
    return q->status;
}

int erd_server_release(erd_server* srv, uint64_t ticket) {
    if (!srv) return ERD_ERR_INVALID;
    Server& S = srv->s;
    void* p = nullptr;
    {
        std::lock_guard<std::mutex> lk(S.mu);
        auto it = S.reqs.find(ticket);
        if (it == S.reqs.end() || it->second->state < 2) return ERD_ERR_INVALID;
        p = it->second->out;
        S.budget_used -= (int64_t)it->second->out_bytes;
        S.reqs.erase(it);
    }
    S.cv.notify_all();
    // The request is complete and the caller is done with its output.
    S.blocks->put(p);
    return ERD_OK;
}

int erd_server_stats(erd_server* srv, erd_server_stats_t* st) {
    if (!srv || !st) return ERD_ERR_INVALID;
    Server& S = srv->s;
    std::lock_guard<std::mutex> lk(S.mu);
    std::vector<double> l = S.latencies;
    std::sort(l.begin(), l.end());
    auto pct = [&](double p) {
        if (l.empty()) return 0.0;
        size_t i = (size_t)std::ceil(p * (double)l.size()) - 1;
        return l[std::min(i, l.size() - 1)] * 1e3;
    };
    st->completed = S.done_requests;
    st->elements = S.done_elems;
    st->batches = S.batches;
    st->p50_ms = pct(0.50);
    st->p99_ms = pct(0.99);
    st->max_ms = l.empty() ? 0.0 : l.back() * 1e3;
    st->span_s = S.t_last_done > S.t_first_submit ? S.t_last_done - S.t_first_submit : 0.0;
    st->budget_in_use = S.budget_used;
    st->prep_s = S.prep_s;
    return ERD_OK;
}

int erd_server_reset_stats(erd_server* srv) {
    if (!srv) return ERD_ERR_INVALID;
    Server& S = srv->s;
    std::lock_guard<std::mutex> lk(S.mu);
    S.latencies.clear();
    S.done_requests = S.done_elems = S.batches = 0;
    S.t_first_submit = S.t_last_done = 0;
    S.prep_s = 0;
    return ERD_OK;
}

int erd_server_destroy(erd_server* srv) {
    if (!srv) return ERD_OK;
    Server& S = srv->s;
    {
        // Requests never launched fail with ERD_ERR_INVALID, so waiters return.
        std::lock_guard<std::mutex> lk(S.mu);
        S.stop = true;
        for (Req* q : S.pending) {
            q->state = 3;
            q->status = ERD_ERR_INVALID;
        }
        S.pending.clear();
    }
    S.cv.notify_all();
    S.launcher.join();
    // drain what is in flight
    for (;;) {
        {
            std::lock_guard<std::mutex> lk(S.mu);
            if (S.inflight.empty()) break;
        }
        S.cv.notify_all();
        std::this_thread::sleep_for(std::chrono::milliseconds(1));
    }
    S.completer.join();
    for (auto& kv : S.reqs)
        if (kv.second->out) S.blocks->put(kv.second->out);
    (void)hipDeviceSynchronize();
    S.blocks.reset();
    for (auto& kv : S.models) {
        if (kv.second->hi) (void)hipFree(kv.second->hi);
        if (kv.second->lo) (void)hipFree(kv.second->lo);
    }
    for (hipStream_t st : S.streams) (void)hipStreamDestroy(st);
    for (void* p : S.pin)
        if (p) (void)hipHostFree(p);
    if (S.pin_status) (void)hipHostFree(S.pin_status);
    erd_engines_destroy(S.pool);
    delete srv;
    return ERD_OK;
}

// Client engines as a view (for checking results against single fills).
int erd_server_client_engines(erd_server* srv, uint64_t client, erd_engines** view) {
    if (!srv || !view) return ERD_ERR_INVALID;
    Server& S = srv->s;
    int sl;
    {
        std::lock_guard<std::mutex> lk(S.mu);
        auto it = S.slot_of.find(client);
        if (it == S.slot_of.end()) return ERD_ERR_INVALID;
        sl = it->second;
    }
    return erd_engines_view(view, S.pool, (int64_t)sl * S.cfg.engines_per_client, S.cfg.engines_per_client);
}

}  // extern "C"
