// Serving runtime: continuous batching of sample requests from many clients.
//
// Every client owns E Vmt19937_64 engines that stay resident in device
// memory for the life of the server (seeded from the client id). Requests are
// queued as they arrive; a launcher thread repeatedly takes the oldest
// admissible requests (their client has no batch in flight, their output fits
// the HBM budget) and turns them into one segment program launch (+ the
// batched L*Z transform for MvNormal requests) on a free stream, while
// earlier batches are still running on other streams. A completion thread
// waits on each batch's event, stamps its requests done and frees the slot.
//
// A request's result depends only on its client's engine state and on the
// request itself: requests of one client run in submission order on that
// client's engines and never share them with another client, so a result is
// the same whether the request runs alone or in a mix.
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

constexpr int64_t kSplit = 1 << 16;  // outputs per engine before a request spreads further

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

struct Req {
    uint64_t ticket = 0;
    erd_serve_request r{};
    std::vector<double> param;  // MvNormal cov / multinomial weights
    std::vector<float> mean;
    int e_req = 1;              // engines the request is spread over
    int64_t per = 0;            // outputs (or columns) per engine
    int64_t out_elems = 0;
    size_t out_bytes = 0;
    void* out = nullptr;
    double t_submit = 0, t_done = 0;
    int state = 0;              // 0 queued, 1 running, 2 done, 3 failed
    int status = ERD_OK;
    int slot = -1;
    bool seed_first = false;    // first request of its client: seeds the engines
};

// Device factorisation cached per (client, covariance).
struct Factor {
    int d = 0;
    uint16_t* hi = nullptr;
    uint16_t* lo = nullptr;
};

struct Batch {
    hipStream_t stream = nullptr;
    int stream_idx = 0;
    hipEvent_t ev = nullptr;
    std::vector<Req*> reqs;
    std::unique_ptr<Scratch> scratch;
    int status = ERD_OK;
};

}  // namespace

struct Server {
    erd_server_config cfg{};
    erd_engines* pool = nullptr;
    std::mutex mu;
    std::condition_variable cv;
    std::deque<Req*> pending;
    std::unordered_map<uint64_t, std::unique_ptr<Req>> reqs;
    std::unordered_map<uint64_t, int> slot_of;  // client -> engine slot
    std::vector<uint64_t> slot_client;
    std::vector<char> slot_busy, slot_seeded;
    std::vector<hipStream_t> streams;
    std::vector<char> stream_free;
    std::deque<Batch*> inflight;
    std::map<std::pair<uint64_t, uint64_t>, Factor> factors;
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
        slot_busy.push_back(0);
        slot_seeded.push_back(0);
        slot_of[client] = s;
        return s;
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

// Builds and launches one batch (called without the server lock).
int Server::prepare_and_launch(Batch* b) {
    hipStream_t s = b->stream;
    b->scratch.reset(new Scratch(s));
    Scratch& sc = *b->scratch;
    const int E = cfg.engines_per_client;
    // Seed the engines of clients seen for the first time:
    // engine g of client c = Vmt19937_64(splitmix64(c ^ salt) + 4g).
    // (The launcher marks at most one request per new client as seed_first.)
    for (Req* q : b->reqs) {
        if (!q->seed_first) continue;
        std::vector<uint64_t> seeds(E);
        const uint64_t base = splitmix64(q->r.client ^ cfg.seed_salt);
        for (int g = 0; g < E; ++g) seeds[g] = base + 4ull * (uint64_t)g;
        uint64_t* d_seeds;
        if (sc.upload(&d_seeds, seeds) != hipSuccess) return ERD_ERR_HIP;
        if (launch_seed(pool->states + (int64_t)q->slot * E, d_seeds, E, s) != hipSuccess) return ERD_ERR_HIP;
    }
    std::vector<Segment> segs;
    std::vector<int32_t> seg_engine;
    std::vector<MvInstance> mv;
    for (Req* q : b->reqs) {
        const erd_serve_request& r = q->r;
        if (hipMallocAsync(&q->out, q->out_bytes, s) != hipSuccess) return ERD_ERR_OOM;
        const int32_t e0 = q->slot * E;
        if (r.kind == ERD_SERVE_MVNORMAL) {
            const int d = r.dim;
            const uint64_t key = hash_params(q->param);
            auto fk = std::make_pair(r.client, key);
            auto it = factors.find(fk);
            if (it == factors.end()) {
                double* a;
                if (sc.alloc(&a, (size_t)d * d) != hipSuccess) return ERD_ERR_OOM;
                if (hipMemcpyAsync(a, q->param.data(), sizeof(double) * d * d, hipMemcpyHostToDevice, s) !=
                        hipSuccess ||
                    hipStreamSynchronize(s) != hipSuccess)
                    return ERD_ERR_HIP;
                uint16_t *hi, *lo;
                float* l32;
                std::vector<int64_t> so;
                int rc = factor_and_split({d}, a, {0, (int64_t)d * d}, &hi, &lo, &l32, so, nullptr, s, sc);
                if (rc) return rc;
                // keep the bf16 planes for later requests of this client
                Factor f;
                f.d = d;
                const int dp = (d + 15) & ~15;
                if (hipMalloc(&f.hi, sizeof(uint16_t) * dp * dp) != hipSuccess ||
                    hipMalloc(&f.lo, sizeof(uint16_t) * dp * dp) != hipSuccess)
                    return ERD_ERR_OOM;
                hipMemcpyAsync(f.hi, hi, sizeof(uint16_t) * dp * dp, hipMemcpyDeviceToDevice, s);
                hipMemcpyAsync(f.lo, lo, sizeof(uint16_t) * dp * dp, hipMemcpyDeviceToDevice, s);
                it = factors.emplace(fk, f).first;
            }
            float* z;
            if (sc.alloc(&z, (size_t)q->out_elems) != hipSuccess) return ERD_ERR_OOM;
            for (int g = 0; g < q->e_req; ++g) {
                Segment sg{};
                sg.kind = SEG_STD_NORMAL_F32;
                sg.count = (int64_t)d * q->per;
                sg.out = z + (int64_t)g * d * q->per;
                segs.push_back(sg);
                seg_engine.push_back(e0 + g);
            }
            float* mean = nullptr;
            if (!q->mean.empty()) {
                if (sc.upload(&mean, q->mean) != hipSuccess) return ERD_ERR_HIP;
            }
            mv.push_back(MvInstance{d, 0, q->per * q->e_req, it->second.hi, it->second.lo, z, (float*)q->out,
                                    mean});
        } else if (r.kind == ERD_SERVE_MULTINOMIAL) {
            const int64_t K = r.dim;
            const int32_t T = (int32_t)r.a;
            HostTable t;
            int rc = build_table(q->param.data(), K, t);
            if (rc) return rc;
            int32_t* o = (int32_t*)q->out;
            const bool small = T < std::max<int64_t>(30 * (K - 1), 100);
            Segment proto{};
            if (small) {
                if (hipMemsetAsync(o, 0, q->out_bytes, s) != hipSuccess) return ERD_ERR_HIP;
                proto.kind = SEG_MULTINOMIAL_SMALL;
                proto.i0 = T;
                proto.d[0] = (double)K;
                rc = upload_table(t, proto, sc);
                if (rc) return rc;
            }
            std::vector<double> probs = q->param;
            double sum = 0;
            for (double v : probs) sum += v;
            for (double& v : probs) v /= sum;
            for (int g = 0; g < q->e_req; ++g) {
                int32_t* og = o + (int64_t)g * K * q->per;
                if (small) {
                    Segment sg = proto;
                    sg.count = q->per;
                    sg.out = og;
                    segs.push_back(sg);
                    seg_engine.push_back(e0 + g);
                    continue;
                }
                int32_t* rr;
                if (sc.upload(&rr, std::vector<int32_t>((size_t)q->per, T)) != hipSuccess) return ERD_ERR_HIP;
                double rest_p = 1 - probs[0];
                for (int64_t row = 0; row + 1 < K; ++row) {
                    Segment sg{};
                    sg.kind = SEG_MULTINOMIAL_ROW;
                    sg.count = q->per;
                    sg.out = og + row;
                    sg.in0 = rr;
                    sg.d[0] = row == 0 ? (float)probs[0] : (float)(probs[row] / rest_p);
                    sg.d[1] = (double)K;
                    if (row > 0) rest_p -= probs[row];
                    segs.push_back(sg);
                    seg_engine.push_back(e0 + g);
                }
                Segment last{};
                last.kind = SEG_MULTINOMIAL_LAST;
                last.count = q->per;
                last.out = og + (K - 1);
                last.in0 = rr;
                last.d[1] = (double)K;
                segs.push_back(last);
                seg_engine.push_back(e0 + g);
            }
        } else {
            for (int g = 0; g < q->e_req; ++g) {
                erd_request er{};
                er.dist = r.kind;
                er.engine = e0 + g;
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
    // Group the segments by engine, keeping each engine's order.
    std::vector<size_t> order(segs.size());
    for (size_t i = 0; i < order.size(); ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(),
                     [&](size_t x, size_t y) { return seg_engine[x] < seg_engine[y]; });
    std::vector<Segment> segs2;
    std::vector<int32_t> eng2;
    for (size_t i : order) {
        segs2.push_back(segs[i]);
        eng2.push_back(seg_engine[i]);
    }
    int rc = run_segments(pool, segs2, eng2, s);
    if (rc) return rc;
    if (!mv.empty()) {
        MvInstance* d_inst;
        if (sc.upload(&d_inst, mv) != hipSuccess) return ERD_ERR_HIP;
        rc = run_lz(mv, d_inst, s, sc);
        if (rc) return rc;
    }
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
                    std::vector<char> blocked(slot_client.size() + 1, 0);
                    std::vector<Req*> take;
                    int64_t take_elems = 0;
                    for (auto it = pending.begin(); it != pending.end();) {
                        Req* q = *it;
                        if ((int64_t)take.size() >= cfg.max_batch_requests) break;
                        if (!take.empty() && cfg.max_batch_elements > 0 &&
                            take_elems + q->out_elems > cfg.max_batch_elements)
                            break;
                        const int sl = q->slot;
                        if (blocked[sl] || slot_busy[sl] ||
                            budget_used + (int64_t)q->out_bytes > cfg.hbm_budget_bytes) {
                            blocked[sl] = 1;
                            ++it;
                            continue;
                        }
                        budget_used += (int64_t)q->out_bytes;
                        take_elems += q->out_elems;
                        q->state = 1;
                        if (!slot_seeded[sl]) {
                            slot_seeded[sl] = 1;
                            q->seed_first = true;
                        }
                        take.push_back(q);
                        it = pending.erase(it);
                    }
                    if (!take.empty()) {
                        for (Req* q : take) slot_busy[q->slot] = 1;
                        b = new Batch();
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
        hipEventCreateWithFlags(&b->ev, hipEventDisableTiming);
        b->status = prepare_and_launch(b);
        {
            std::lock_guard<std::mutex> lk(mu);
            inflight.push_back(b);
            ++batches;
        }
        cv.notify_all();
    }
}

void Server::completer_loop() {
    for (;;) {
        Batch* b = nullptr;
        {
            std::unique_lock<std::mutex> lk(mu);
            cv.wait(lk, [&] { return stop || !inflight.empty(); });
            if (inflight.empty()) return;
            b = inflight.front();
        }
        if (b->status == ERD_OK) {
            if (hipEventSynchronize(b->ev) != hipSuccess) b->status = ERD_ERR_HIP;
        } else {
            hipStreamSynchronize(b->stream);
        }
        const double t = now_s();
        b->scratch.reset();  // stream-ordered frees
        hipEventDestroy(b->ev);
        {
            std::lock_guard<std::mutex> lk(mu);
            inflight.pop_front();
            for (Req* q : b->reqs) {
                q->t_done = t;
                q->state = b->status == ERD_OK ? 2 : 3;
                q->status = b->status;
                slot_busy[q->slot] = 0;
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
    }
    S.launcher = std::thread([&S] { S.launcher_loop(); });
    S.completer = std::thread([&S] { S.completer_loop(); });
    *out = srv;
    return ERD_OK;
}

int erd_server_submit(erd_server* srv, const erd_serve_request* r, uint64_t* ticket) {
    if (!srv || !r || !ticket || r->count <= 0) return ERD_ERR_INVALID;
    Server& S = srv->s;
    std::unique_ptr<Req> q(new Req());
    q->r = *r;
    const int E = S.cfg.engines_per_client;
    if (r->kind == ERD_SERVE_MVNORMAL) {
        if (r->dim <= 0 || !r->host_param) return ERD_ERR_INVALID;
        q->param.assign(r->host_param, r->host_param + (size_t)r->dim * r->dim);
        if (r->host_mean) q->mean.assign(r->host_mean, r->host_mean + r->dim);
        q->e_req = (int)std::min<int64_t>(E, std::max<int64_t>(1, (int64_t)r->dim * r->count / erd::kSplit));
        q->per = erd::round_up((r->count + q->e_req - 1) / q->e_req, 8);
        q->out_elems = (int64_t)r->dim * q->per * q->e_req;
        q->out_bytes = sizeof(float) * (size_t)q->out_elems;
    } else if (r->kind == ERD_SERVE_MULTINOMIAL) {
        if (r->dim <= 0 || !r->host_param || r->a < 0) return ERD_ERR_INVALID;
        q->param.assign(r->host_param, r->host_param + r->dim);
        double sum = 0;
        for (double v : q->param) {
            if (v < 0) return ERD_ERR_INVALID;
            sum += v;
        }
        if (!std::isfinite(sum)) return ERD_ERR_NONFINITE;
        if (!(sum > 0)) return ERD_ERR_INVALID;
        q->e_req = (int)std::min<int64_t>(E, std::max<int64_t>(1, (int64_t)r->dim * r->count / erd::kSplit));
        q->per = erd::round_up((r->count + q->e_req - 1) / q->e_req, 8);
        q->out_elems = (int64_t)r->dim * q->per * q->e_req;
        q->out_bytes = sizeof(int32_t) * (size_t)q->out_elems;
    } else {
        if (r->kind < 0 || r->kind > ERD_DIST_BINOMIAL_I32 || r->kind == ERD_DIST_NORMAL_VOP_F32)
            return ERD_ERR_INVALID;  // per-element parameter kinds are not served
        erd_request probe{};
        probe.dist = r->kind;
        probe.count = erd::packet_width(r->kind);
        probe.out = (void*)16;
        probe.a = r->a;
        probe.b = r->b;
        erd::Segment tmp;
        int rc = erd::make_segment(probe, tmp);
        if (rc) return rc;
        q->e_req = (int)std::min<int64_t>(E, std::max<int64_t>(1, r->count / erd::kSplit));
        q->per = erd::round_up((r->count + q->e_req - 1) / q->e_req, erd::packet_width(r->kind));
        q->out_elems = q->per * q->e_req;
        q->out_bytes = erd::elem_bytes(r->kind) * (size_t)q->out_elems;
    }
    if ((int64_t)q->out_bytes > S.cfg.hbm_budget_bytes) return ERD_ERR_OOM;
    {
        std::lock_guard<std::mutex> lk(S.mu);
        const int sl = S.slot_for(r->client);
        if (sl < 0) return ERD_ERR_UNSUPPORTED;  // engine slots exhausted
        q->slot = sl;
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
    // The request is complete, so a stream-ordered free needs no ordering.
    return p ? (hipFreeAsync(p, nullptr) == hipSuccess ? ERD_OK : ERD_ERR_HIP) : ERD_OK;
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
    return ERD_OK;
}

int erd_server_reset_stats(erd_server* srv) {
    if (!srv) return ERD_ERR_INVALID;
    Server& S = srv->s;
    std::lock_guard<std::mutex> lk(S.mu);
    S.latencies.clear();
    S.done_requests = S.done_elems = S.batches = 0;
    S.t_first_submit = S.t_last_done = 0;
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
        if (kv.second->out) hipFreeAsync(kv.second->out, nullptr);
    hipDeviceSynchronize();
    for (auto& kv : S.factors) {
        hipFree(kv.second.hi);
        hipFree(kv.second.lo);
    }
    for (hipStream_t st : S.streams) hipStreamDestroy(st);
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
