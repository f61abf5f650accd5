// Error state, device checks and the device table cache of libdgamd.
#include <cstdarg>
#include <memory>
#include <mutex>
#include <unordered_map>

#include "dg_internal.hpp"

namespace dg {

namespace {
thread_local char g_last_error[512] = "";
}

int fail(int code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
    va_end(ap);
    return code;
}

void clear_error() { g_last_error[0] = '\0'; }

int require_device() {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return fail(DG_ERR_NO_DEVICE, "no HIP device: %s", hipGetErrorString(e));
    static thread_local int checked_dev = -1;
    if (checked_dev == dev) return DG_OK;
    hipDeviceProp_t prop;
    e = hipGetDeviceProperties(&prop, dev);
    if (e != hipSuccess) return fail(DG_ERR_NO_DEVICE, "hipGetDeviceProperties: %s", hipGetErrorString(e));
    if (strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(DG_ERR_NO_DEVICE, "device %d is %s; this library is built for gfx950 only", dev,
                    prop.gcnArchName);
    checked_dev = dev;
    return DG_OK;
}

namespace {

// FNV-1a over the table bytes.
uint64_t hash_bytes(const void* p, size_t n) {
    const unsigned char* b = static_cast<const unsigned char*>(p);
    uint64_t h = 1469598103934665603ull;
    for (size_t i = 0; i < n; ++i) { h ^= b[i]; h *= 1099511628211ull; }
    return h;
}

// One resident table.  Its device buffer is never freed: an entry that is
// retired hands the buffer to its size class's free list, and only once the
// entry's last use has finished on the device.
//  * pinned: looked up while a stream was being captured (a graph holds the
//    address for as long as the caller keeps it), or used from two streams
//    (one event cannot cover both); never retired.
//  * done: recorded on the stream of the last use after that call's kernels
//    were enqueued (table_uses_enqueued); an entry is retired only when it is
//    complete, so no retirement waits on the device.
struct TableEntry {
    std::vector<unsigned char> bytes;  // host copy, compared on every hit
    void* dev = nullptr;
    int size_class = 0;                // buffer of 256 << size_class bytes
    uint64_t last_use = 0;
    bool pinned = false;
    int pending = 0;                   // lookups whose kernels are not enqueued yet
    hipStream_t last_stream = nullptr;
    bool used = false;                 // last_stream is meaningful
    hipEvent_t done = nullptr;
};

struct TableCache {
    std::mutex mu;
    std::unordered_map<uint64_t, std::vector<std::unique_ptr<TableEntry>>> map[64];  // per device: hash -> entries
    std::vector<void*> free_bufs[64][40];                                            // per device and size class
    uint64_t clock = 0;
    size_t entries[64] = {};
};

TableCache& cache() {
    static TableCache* c = new TableCache();  // never destroyed: device frees at exit are unsafe
    return *c;
}

// Soft cap on the resident entries per device: beyond it, a new table first
// retires the least recently used retirable entry.
constexpr size_t kMaxEntries = 256;

thread_local std::vector<TableEntry*> t_pending;  // this thread's lookups since its last enqueue mark

bool capturing(hipStream_t stream) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    return hipStreamIsCapturing(stream, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone;
}

int size_class(size_t bytes) {
    int c = 0;
    while (((size_t)256 << c) < bytes) ++c;
    return c;
}

// Retires the least recently used entry of `dev` that is not pinned, has no
// pending lookup and whose last use has completed (an event query: no wait).
void retire_one(TableCache& tc, int dev) {
    TableEntry* best = nullptr;
    uint64_t best_key = 0;
    for (auto& kv : tc.map[dev])
        for (auto& e : kv.second)
            if (!e->pinned && e->pending == 0 && (!best || e->last_use < best->last_use)) {
                if (e->used && (!e->done || hipEventQuery(e->done) != hipSuccess)) continue;
                best = e.get();
                best_key = kv.first;
            }
    if (!best) return;
    tc.free_bufs[dev][best->size_class].push_back(best->dev);
    if (best->done) (void)hipEventDestroy(best->done);
    auto& bucket = tc.map[dev][best_key];
    for (size_t i = 0; i < bucket.size(); ++i)
        if (bucket[i].get() == best) { bucket.erase(bucket.begin() + i); break; }
    if (bucket.empty()) tc.map[dev].erase(best_key);
    --tc.entries[dev];
}

}  // namespace

const void* cached_table(const void* host, size_t bytes, hipStream_t stream) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) {
        fail(DG_ERR_NO_DEVICE, "cached_table: no current device");
        return nullptr;
    }
    const bool cap = stream != nullptr && capturing(stream);
    const uint64_t h = hash_bytes(host, bytes);
    TableCache& tc = cache();
    std::lock_guard<std::mutex> lock(tc.mu);
    auto& bucket = tc.map[dev][h];
    for (auto& e : bucket) {
        if (e->bytes.size() == bytes && memcmp(e->bytes.data(), host, bytes) == 0) {
            e->last_use = ++tc.clock;
            if (cap) {
                e->pinned = true;
            } else {
                ++e->pending;
                t_pending.push_back(e.get());
            }
            return e->dev;
        }
    }
    if (cap) {
        fail(DG_ERR_INVALID,
             "descriptor table not resident while the stream is being captured; "
             "run the same call once before capture");
        return nullptr;
    }
    // Retiring may erase any hash bucket, this one included: look it up again after.
    if (tc.entries[dev] >= kMaxEntries) retire_one(tc, dev);
    auto& slot = tc.map[dev][h];
    auto e = std::make_unique<TableEntry>();
    e->bytes.assign(static_cast<const unsigned char*>(host), static_cast<const unsigned char*>(host) + bytes);
    e->size_class = size_class(bytes);
    if (e->size_class >= 40) {
        fail(DG_ERR_ALLOC, "cached_table: table of %zu bytes", bytes);
        return nullptr;
    }
    auto& fl = tc.free_bufs[dev][e->size_class];
    if (!fl.empty()) {
        e->dev = fl.back();
        fl.pop_back();
    } else {
        hipError_t err = hipMalloc(&e->dev, (size_t)256 << e->size_class);
        if (err != hipSuccess) {
            fail(DG_ERR_ALLOC, "cached_table: hipMalloc(%zu): %s", (size_t)256 << e->size_class, hipGetErrorString(err));
            return nullptr;
        }
    }
    // Synchronous copy from the entry's own host bytes: complete on return,
    // and the buffer (new, or retired after its last use completed) is idle.
    hipError_t err = hipMemcpy(e->dev, e->bytes.data(), bytes, hipMemcpyHostToDevice);
    if (err != hipSuccess) {
        fl.push_back(e->dev);
        fail(DG_ERR_HIP, "cached_table: hipMemcpy: %s", hipGetErrorString(err));
        return nullptr;
    }
    e->last_use = ++tc.clock;
    e->pending = 1;
    t_pending.push_back(e.get());
    void* d = e->dev;
    slot.push_back(std::move(e));
    ++tc.entries[dev];
    return d;
}

void table_uses_enqueued(hipStream_t stream) {
    if (t_pending.empty()) return;
    TableCache& tc = cache();
    std::lock_guard<std::mutex> lock(tc.mu);
    const bool cap = stream != nullptr && capturing(stream);
    for (TableEntry* e : t_pending) {
        --e->pending;
        if (e->pinned) continue;
        if (cap || (e->used && e->last_stream != stream)) {
            e->pinned = true;
            continue;
        }
        if (!e->done && hipEventCreateWithFlags(&e->done, hipEventDisableTiming) != hipSuccess) {
            e->done = nullptr;
            e->pinned = true;  // cannot track its completion: keep it
            continue;
        }
        if (hipEventRecord(e->done, stream) != hipSuccess) {
            e->pinned = true;
            continue;
        }
        e->last_stream = stream;
        e->used = true;
    }
    t_pending.clear();
}

}  // namespace dg

extern "C" {

const char* dg_last_error(void) { return dg::g_last_error; }

int dg_version(void) { return 100; }

int dg_device_ok(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return 0;
    return dg::require_device() == DG_OK ? 1 : 0;
}

}  // extern "C"
