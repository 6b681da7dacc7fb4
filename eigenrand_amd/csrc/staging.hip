// Pinned host staging for host-to-device uploads.
//
// A copy from pageable host memory makes the runtime wait for the stream, so
// every parameter upload would serialise the host with the device. Uploads
// instead go through a per-device ring of pinned memory: the bytes are copied
// into the ring, the DMA is enqueued on the caller's stream, and an event
// recorded behind it marks when that part of the ring may be reused. The host
// only waits when the ring is full (for the oldest upload still in flight).
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <vector>
#include "erd_internal.hpp"

namespace erd {
namespace {

class PinnedRing {
public:
    hipError_t upload(void* dst, const void* src, size_t bytes, hipStream_t s) {
        if (bytes == 0) return hipSuccess;
        std::lock_guard<std::mutex> lk(mu_);
        const size_t need = (bytes + 255) & ~(size_t)255;
        size_t off = 0;
        hipError_t e = reserve(need, &off);
        if (e != hipSuccess) return e;
        std::memcpy(base_ + off, src, bytes);
        e = hipMemcpyAsync(dst, base_ + off, bytes, hipMemcpyHostToDevice, s);
        if (e != hipSuccess) return e;
        hipEvent_t ev;
        e = take_event(&ev);
        if (e == hipSuccess) e = hipEventRecord(ev, s);
        if (e != hipSuccess) {
            // the chunk cannot be tracked: wait for the copy instead
            (void)hipStreamSynchronize(s);
            return e;
        }
        chunks_.push_back(Chunk{off, need, ev});
        return hipSuccess;
    }

private:
    struct Chunk {
        size_t off, len;
        hipEvent_t ev;
    };

    hipError_t take_event(hipEvent_t* ev) {
        if (!events_.empty()) {
            *ev = events_.back();
            events_.pop_back();
            return hipSuccess;
        }
        return hipEventCreateWithFlags(ev, hipEventDisableTiming);
    }

    // Drop finished chunks from the front (oldest first); `wait` blocks on the oldest.
    hipError_t retire(bool wait) {
        while (!chunks_.empty()) {
            Chunk& c = chunks_.front();
            hipError_t q = wait ? hipEventSynchronize(c.ev) : hipEventQuery(c.ev);
            if (q == hipErrorNotReady) return hipSuccess;
            if (q != hipSuccess) return q;
            events_.push_back(c.ev);
            chunks_.pop_front();
            if (wait) return hipSuccess;
        }
        return hipSuccess;
    }

    hipError_t reserve(size_t need, size_t* off) {
        if (need > cap_) {
            while (!chunks_.empty()) {
                hipError_t e = retire(true);
                if (e != hipSuccess) return e;
            }
            if (base_) (void)hipHostFree(base_);
            base_ = nullptr;
            const size_t cap = std::max<size_t>({need, 2 * cap_, (size_t)32 << 20});
            hipError_t e = hipHostMalloc((void**)&base_, cap, hipHostMallocDefault);
            if (e != hipSuccess) {
                cap_ = 0;
                return e;
            }
            cap_ = cap;
        }
        for (;;) {
            hipError_t e = retire(false);
            if (e != hipSuccess) return e;
            if (chunks_.empty()) {
                *off = 0;
                return hipSuccess;
            }
            const size_t tail = chunks_.front().off;
            const size_t head = chunks_.back().off + chunks_.back().len;
            if (head > tail) {  // in use: [tail, head)
                if (cap_ - head >= need) {
                    *off = head;
                    return hipSuccess;
                }
                if (tail >= need) {
                    *off = 0;
                    return hipSuccess;
                }
            } else if (head < tail && tail - head >= need) {  // in use: [tail, cap) + [0, head)
                *off = head;
                return hipSuccess;
            }
            e = retire(true);  // full: wait for the oldest upload
            if (e != hipSuccess) return e;
        }
    }

    std::mutex mu_;
    char* base_ = nullptr;
    size_t cap_ = 0;
    std::deque<Chunk> chunks_;
    std::vector<hipEvent_t> events_;
};

std::mutex g_rings_mu;
std::map<int, PinnedRing*> g_rings;  // per device; kept until the process ends

}  // namespace

namespace {

size_t size_class(size_t b) {
    if (b <= 4096) return 4096;
    size_t base = 4096;
    while (base * 2 < b) base *= 2;  // base < b <= 2 * base
    const size_t step = base / 4;
    return base + (b - base + step - 1) / step * step;
}

class ScratchCache {
public:
    hipError_t get(void** p, size_t bytes, hipStream_t s) {
        const size_t cls = size_class(bytes);
        {
            std::lock_guard<std::mutex> lk(mu_);
            for (size_t i = 0; i < free_.size(); ++i) {
                Blk& b = free_[i];
                if (b.cls != cls) continue;
                if (b.s != s && hipEventQuery(b.ev) != hipSuccess) continue;
                *p = b.p;
                events_.push_back(b.ev);
                free_[i] = free_.back();
                free_.pop_back();
                return hipSuccess;
            }
        }
        hipError_t e = hipMallocAsync(p, cls, s);
        if (e == hipErrorOutOfMemory) {
            (void)hipGetLastError();
            release_idle();
            e = hipMallocAsync(p, cls, s);
        }
        return e;
    }

    void put(void* p, size_t bytes, hipStream_t s) {
        std::lock_guard<std::mutex> lk(mu_);
        hipEvent_t ev = nullptr;
        if (!events_.empty()) {
            ev = events_.back();
            events_.pop_back();
        } else if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) {
            ev = nullptr;
        }
        if (!ev || hipEventRecord(ev, s) != hipSuccess) {
            // cannot track its stream position: give it back to the driver
            if (ev) events_.push_back(ev);
            (void)hipFreeAsync(p, s);
            return;
        }
        free_.push_back(Blk{p, size_class(bytes), s, ev});
    }

private:
    struct Blk {
        void* p;
        size_t cls;
        hipStream_t s;
        hipEvent_t ev;
    };

    // Frees every cached block (after its last use has completed).
    void release_idle() {
        std::lock_guard<std::mutex> lk(mu_);
        for (Blk& b : free_) {
            (void)hipEventSynchronize(b.ev);
            (void)hipFree(b.p);
            events_.push_back(b.ev);
        }
        free_.clear();
    }

    std::mutex mu_;
    std::vector<Blk> free_;
    std::vector<hipEvent_t> events_;
};

std::mutex g_cache_mu;
std::map<int, ScratchCache*> g_caches;  // per device; kept until the process ends

ScratchCache* cache_of_current_device() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return nullptr;
    std::lock_guard<std::mutex> lk(g_cache_mu);
    ScratchCache*& c = g_caches[dev];
    if (!c) c = new ScratchCache();
    return c;
}

std::mutex g_side_mu;
std::map<int, hipStream_t> g_side;  // per device; kept until the process ends

}  // namespace

hipError_t side_stream(hipStream_t* out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    std::lock_guard<std::mutex> lk(g_side_mu);
    hipStream_t& st = g_side[dev];
    if (!st) {
        e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
        if (e != hipSuccess) {
            st = nullptr;
            return e;
        }
    }
    *out = st;
    return hipSuccess;
}

hipError_t stream_after(hipStream_t to, hipStream_t from) {
    hipEvent_t ev;
    hipError_t e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
    if (e != hipSuccess) return e;
    e = hipEventRecord(ev, from);
    if (e == hipSuccess) e = hipStreamWaitEvent(to, ev, 0);
    const hipError_t d = hipEventDestroy(ev);  // released once the record completes
    return e != hipSuccess ? e : d;
}

hipError_t scratch_alloc(void** p, size_t bytes, hipStream_t s) {
    ScratchCache* c = cache_of_current_device();
    if (!c) return hipErrorInvalidDevice;
    return c->get(p, bytes, s);
}

void scratch_free(void* p, size_t bytes, hipStream_t s) {
    if (!p) return;
    ScratchCache* c = cache_of_current_device();
    if (c) c->put(p, bytes, s);
    else (void)hipFreeAsync(p, s);
}

hipError_t upload_async(void* dst, const void* src, size_t bytes, hipStream_t s) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    PinnedRing* ring;
    {
        std::lock_guard<std::mutex> lk(g_rings_mu);
        PinnedRing*& r = g_rings[dev];
        if (!r) r = new PinnedRing();
        ring = r;
    }
    return ring->upload(dst, src, bytes, s);
}

}  // namespace erd
