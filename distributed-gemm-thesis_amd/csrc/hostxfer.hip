// Host <-> device copies of column blocks, for callers whose matrices live
// in host memory (the full-offload path: global matrices on the host are
// decomposed onto the device tiles and the result is gathered back).
//
// Pinned host memory is moved by the DMA engines directly.  Pageable memory
// is not visible to them, so it goes through a ring of kSlots pinned slots
// of kSlotBytes: host threads copy one slot's share of the columns between
// the caller's buffer and a slot while the DMA engine moves the previous
// slot over PCIe.  The two overlap, and the host side copies at several
// threads' memory bandwidth instead of one.  On an MI355X host this moves
// pageable buffers at ~55 GB/s each way; the runtime's own pageable copies
// reach ~30 GB/s host-to-device and ~18 GB/s device-to-host, and pinned DMA
// 57.6 GB/s.
#include <algorithm>
#include <condition_variable>
#include <functional>
#include <mutex>
#include <thread>

#include <unistd.h>

#include "dg_internal.hpp"

namespace dg {

namespace {

constexpr size_t kSlotBytes = 32u << 20;  // one DMA transfer
constexpr int kSlots = 3;                  // two in flight while the third is filled or drained
constexpr int kMaxThreads = 8;             // host copy threads, the calling thread included
constexpr int64_t kMinParallel = 1 << 20;  // smaller pieces stay on the calling thread

// Worker threads kept for the life of the process.  run(parts, f) calls
// f(0), ..., f(parts - 1), part 0 on the calling thread, and returns when
// every part has returned.  One run at a time (callers hold the staging lock).
// A child created by fork() has none of the workers, so there every part
// runs on the calling thread.
class CopyPool {
   public:
    explicit CopyPool(int threads) : n_(threads), owner_(getpid()) {
        for (int i = 1; i < n_; ++i) th_.emplace_back([this, i] { loop(i); });
    }
    ~CopyPool() {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
        }
        go_.notify_all();
        for (auto& t : th_) t.join();
    }
    int size() const { return n_; }
    void run(int parts, const std::function<void(int)>& f) {
        parts = std::min(parts, n_);
        if (getpid() != owner_) {
            for (int i = 0; i < parts; ++i) f(i);
            return;
        }
        if (parts <= 1) {
            f(0);
            return;
        }
        {
            std::lock_guard<std::mutex> lk(mu_);
            job_ = &f;
            parts_ = parts;
            remaining_ = parts - 1;
            ++gen_;
        }
        go_.notify_all();
        f(0);
        std::unique_lock<std::mutex> lk(mu_);
        done_.wait(lk, [this] { return remaining_ == 0; });
        job_ = nullptr;
    }

   private:
    // A worker whose id is below the run's part count takes part `id`; run()
    // waits for it, so no worker can miss a run it has a part in.
    void loop(int id) {
        uint64_t seen = 0;
        std::unique_lock<std::mutex> lk(mu_);
        for (;;) {
            go_.wait(lk, [&] { return stop_ || gen_ != seen; });
            if (stop_) return;
            seen = gen_;
            if (id >= parts_) continue;
            const std::function<void(int)>* job = job_;
            lk.unlock();
            (*job)(id);
            lk.lock();
            if (--remaining_ == 0) done_.notify_one();
        }
    }
    int n_;
    pid_t owner_;
    std::vector<std::thread> th_;
    std::mutex mu_;
    std::condition_variable go_, done_;
    const std::function<void(int)>* job_ = nullptr;
    int parts_ = 0, remaining_ = 0;
    uint64_t gen_ = 0;
    bool stop_ = false;
};

CopyPool& pool() {
    static CopyPool p(std::max(1, std::min<int>(kMaxThreads, (int)std::thread::hardware_concurrency())));
    return p;
}

// The pinned ring: allocated at the first staged copy and kept until the
// process ends (a pinned allocation costs milliseconds; every later copy
// reuses it).  The lock serialises the staged copies of all threads.
struct Staging {
    std::mutex mu;
    char* slot[kSlots] = {};
};

Staging& staging() {
    static Staging s;
    return s;
}

int ensure_slots(Staging& s) {
    for (int i = 0; i < kSlots; ++i) {
        if (s.slot[i]) continue;
        const hipError_t e = hipHostMalloc((void**)&s.slot[i], kSlotBytes, hipHostMallocDefault);
        if (e != hipSuccess) {
            s.slot[i] = nullptr;
            return fail(DG_ERR_ALLOC, "pinned staging buffer of %zu bytes: %s", kSlotBytes, hipGetErrorString(e));
        }
    }
    return DG_OK;
}

enum class Mem { PAGEABLE, PINNED, DEVICE };

Mem memory_kind(const void* p) {
    hipPointerAttribute_t a;
    if (hipPointerGetAttributes(&a, p) != hipSuccess) {
        (void)hipGetLastError();  // memory the runtime does not know: pageable host memory
        return Mem::PAGEABLE;
    }
    switch (a.type) {
        case hipMemoryTypeUnregistered: return Mem::PAGEABLE;
        case hipMemoryTypeDevice:
        case hipMemoryTypeArray: return Mem::DEVICE;
        default: return Mem::PINNED;  // pinned, registered, managed or unified host memory
    }
}

// Events of one staged copy, one per slot.  On every return path the
// destructor waits for the DMA transfers recorded on them, so the slots are
// idle before the staging lock is released.
struct SlotEvents {
    hipEvent_t ev[kSlots] = {};
    int made = 0;
    int create() {
        for (; made < kSlots; ++made) DG_HIP(hipEventCreateWithFlags(&ev[made], hipEventDisableTiming));
        return DG_OK;
    }
    ~SlotEvents() {
        for (int i = 0; i < made; ++i) {
            (void)hipEventSynchronize(ev[i]);
            (void)hipEventDestroy(ev[i]);
        }
    }
};

// A staged transfer of `height` columns of `width` bytes moves in pieces of
// at most kSlotBytes: whole columns when a column fits a slot, otherwise
// one kSlotBytes-long part of one column.
struct Piece {
    int64_t c0, nc;   // first column, columns
    int64_t b0, len;  // byte range within each of those columns
};

struct Pieces {
    int64_t width, height, per;  // per: whole columns per piece (0: columns are split)
    int64_t count() const {
        if (width == 0 || height == 0) return 0;
        if (per > 0) return (height + per - 1) / per;
        const int64_t parts = (width + (int64_t)kSlotBytes - 1) / (int64_t)kSlotBytes;
        return height * parts;
    }
    Piece at(int64_t i) const {
        if (per > 0) return Piece{i * per, std::min(per, height - i * per), 0, width};
        const int64_t parts = (width + (int64_t)kSlotBytes - 1) / (int64_t)kSlotBytes;
        const int64_t c = i / parts, b0 = (i % parts) * (int64_t)kSlotBytes;
        return Piece{c, 1, b0, std::min((int64_t)kSlotBytes, width - b0)};
    }
};

Pieces make_pieces(int64_t width, int64_t height) {
    return Pieces{width, height, width <= (int64_t)kSlotBytes ? (int64_t)kSlotBytes / std::max<int64_t>(width, 1) : 0};
}

// Copies a piece between its slot (the piece's bytes back to back) and the
// host buffer (column j of the piece at host + (c0 + j) * pitch + b0), split
// evenly over the pool's threads.
void host_copy(char* slot, char* host, int64_t pitch, const Piece& p, bool to_slot) {
    const int64_t total = p.nc * p.len;
    char* first = host + p.c0 * pitch + p.b0;
    const int parts = total < kMinParallel ? 1 : pool().size();
    pool().run(parts, [&](int t) {
        int64_t x = total * t / parts;
        const int64_t end = total * (t + 1) / parts;
        while (x < end) {
            const int64_t j = x / p.len, o = x % p.len;
            const int64_t n = std::min(end - x, p.len - o);
            char* h = first + j * pitch + o;
            if (to_slot) memcpy(slot + x, h, (size_t)n);
            else memcpy(h, slot + x, (size_t)n);
            x += n;
        }
    });
}

// One DMA transfer of a piece between a slot and the device buffer.
hipError_t dma_piece(char* dev, int64_t dpitch, char* slot, const Piece& p, bool h2d, hipStream_t st) {
    char* d = dev + p.c0 * dpitch + p.b0;
    if (p.nc == 1 || dpitch == p.len)
        return h2d ? hipMemcpyAsync(d, slot, (size_t)(p.nc * p.len), hipMemcpyHostToDevice, st)
                   : hipMemcpyAsync(slot, d, (size_t)(p.nc * p.len), hipMemcpyDeviceToHost, st);
    return h2d ? hipMemcpy2DAsync(d, (size_t)dpitch, slot, (size_t)p.len, (size_t)p.len, (size_t)p.nc,
                                  hipMemcpyHostToDevice, st)
               : hipMemcpy2DAsync(slot, (size_t)p.len, d, (size_t)dpitch, (size_t)p.len, (size_t)p.nc,
                                  hipMemcpyDeviceToHost, st);
}

int staged_h2d(char* dev, int64_t dpitch, char* host, int64_t hpitch, int64_t width, int64_t height,
               hipStream_t st) {
    Staging& s = staging();
    std::lock_guard<std::mutex> lk(s.mu);
    int rc = ensure_slots(s);
    if (rc != DG_OK) return rc;
    SlotEvents ev;
    if ((rc = ev.create()) != DG_OK) return rc;
    const Pieces pieces = make_pieces(width, height);
    const int64_t n = pieces.count();
    for (int64_t i = 0; i < n; ++i) {
        const int k = (int)(i % kSlots);
        const Piece p = pieces.at(i);
        if (i >= kSlots) DG_HIP(hipEventSynchronize(ev.ev[k]));  // the slot's previous DMA has read it
        host_copy(s.slot[k], host, hpitch, p, true);
        DG_HIP(dma_piece(dev, dpitch, s.slot[k], p, true, st));
        DG_HIP(hipEventRecord(ev.ev[k], st));
    }
    if (n > 0) DG_HIP(hipEventSynchronize(ev.ev[(n - 1) % kSlots]));  // stream order: every piece has landed
    return DG_OK;
}

int staged_d2h(char* host, int64_t hpitch, char* dev, int64_t dpitch, int64_t width, int64_t height,
               hipStream_t st) {
    Staging& s = staging();
    std::lock_guard<std::mutex> lk(s.mu);
    int rc = ensure_slots(s);
    if (rc != DG_OK) return rc;
    SlotEvents ev;
    if ((rc = ev.create()) != DG_OK) return rc;
    const Pieces pieces = make_pieces(width, height);
    const int64_t n = pieces.count();
    // piece i's DMA runs while the host drains piece i - 1; piece i's slot
    // last held piece i - kSlots, drained two iterations earlier
    for (int64_t i = 0; i <= n; ++i) {
        if (i < n) {
            const int k = (int)(i % kSlots);
            DG_HIP(dma_piece(dev, dpitch, s.slot[k], pieces.at(i), false, st));
            DG_HIP(hipEventRecord(ev.ev[k], st));
        }
        if (i >= 1) {
            const int k = (int)((i - 1) % kSlots);
            DG_HIP(hipEventSynchronize(ev.ev[k]));
            host_copy(s.slot[k], host, hpitch, pieces.at(i - 1), false);
        }
    }
    return DG_OK;
}

int copy2d(const char* who, bool h2d, char* dev, int64_t dpitch, char* host, int64_t hpitch, int64_t width,
           int64_t height, void* stream) {
    if (width < 0 || height < 0) return fail(DG_ERR_INVALID, "%s: negative size %ld x %ld", who, (long)width,
                                             (long)height);
    if (width == 0 || height == 0) return DG_OK;
    if (!dev || !host) return fail(DG_ERR_INVALID, "%s: null pointer", who);
    if (dpitch < width || hpitch < width)
        return fail(DG_ERR_INVALID, "%s: pitches %ld (device) and %ld (host) must be >= width %ld", who,
                    (long)dpitch, (long)hpitch, (long)width);
    if (height > 1 && (dpitch > INT64_MAX / (height - 1) || hpitch > INT64_MAX / (height - 1)))
        return fail(DG_ERR_INVALID, "%s: buffer extent overflows 64 bits", who);
    int rc = require_device();
    if (rc != DG_OK) return rc;
    const hipStream_t st = as_stream(stream);
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    DG_HIP(hipStreamIsCapturing(st, &cap));
    if (cap != hipStreamCaptureStatusNone)
        return fail(DG_ERR_INVALID, "%s: host copies complete before returning and cannot be captured", who);
    if (memory_kind(dev) != Mem::DEVICE) return fail(DG_ERR_INVALID, "%s: the device buffer is not device memory", who);
    const Mem hk = memory_kind(host);
    if (hk == Mem::DEVICE) return fail(DG_ERR_INVALID, "%s: the host buffer is device memory", who);
    if (hk == Mem::PAGEABLE)
        return h2d ? staged_h2d(dev, dpitch, host, hpitch, width, height, st)
                   : staged_d2h(host, hpitch, dev, dpitch, width, height, st);
    // pinned or managed host memory: the DMA engines read or write it directly
    if (dpitch == width && hpitch == width)
        DG_HIP(h2d ? hipMemcpyAsync(dev, host, (size_t)(width * height), hipMemcpyHostToDevice, st)
                   : hipMemcpyAsync(host, dev, (size_t)(width * height), hipMemcpyDeviceToHost, st));
    else
        DG_HIP(h2d ? hipMemcpy2DAsync(dev, (size_t)dpitch, host, (size_t)hpitch, (size_t)width, (size_t)height,
                                      hipMemcpyHostToDevice, st)
                   : hipMemcpy2DAsync(host, (size_t)hpitch, dev, (size_t)dpitch, (size_t)width, (size_t)height,
                                      hipMemcpyDeviceToHost, st));
    DG_HIP(hipStreamSynchronize(st));
    return DG_OK;
}

}  // namespace

}  // namespace dg

extern "C" {

int dg_copy2d_h2d(void* dev, int64_t dev_pitch, const void* host, int64_t host_pitch, int64_t width, int64_t height,
                  void* stream) {
    return dg::copy2d("dg_copy2d_h2d", true, static_cast<char*>(dev), dev_pitch,
                      const_cast<char*>(static_cast<const char*>(host)), host_pitch, width, height, stream);
}

int dg_copy2d_d2h(void* host, int64_t host_pitch, const void* dev, int64_t dev_pitch, int64_t width, int64_t height,
                  void* stream) {
    return dg::copy2d("dg_copy2d_d2h", false, const_cast<char*>(static_cast<const char*>(dev)), dev_pitch,
                      static_cast<char*>(host), host_pitch, width, height, stream);
}

}  // extern "C"
