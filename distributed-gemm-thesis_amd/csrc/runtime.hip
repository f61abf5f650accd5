// Error state, device checks and the device table cache of libdgamd.
#include <cstdarg>
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

struct TableEntry {
    std::vector<unsigned char> bytes;  // host copy, compared on every hit
    void* dev = nullptr;
    uint64_t last_use = 0;
};

struct TableCache {
    std::mutex mu;
    // key: (device, hash)
    std::unordered_map<uint64_t, std::vector<TableEntry>> map[64];
    uint64_t clock = 0;
    size_t entries = 0;
};

TableCache& cache() {
    static TableCache* c = new TableCache();  // never destroyed: device frees at exit are unsafe
    return *c;
}

constexpr size_t kMaxEntries = 256;

}  // namespace

const void* cached_table(const void* host, size_t bytes, hipStream_t stream) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) {
        fail(DG_ERR_NO_DEVICE, "cached_table: no current device");
        return nullptr;
    }
    const uint64_t h = hash_bytes(host, bytes);
    TableCache& tc = cache();
    std::lock_guard<std::mutex> lock(tc.mu);
    auto& bucket = tc.map[dev][h];
    for (auto& e : bucket) {
        if (e.bytes.size() == bytes && memcmp(e.bytes.data(), host, bytes) == 0) {
            e.last_use = ++tc.clock;
            return e.dev;
        }
    }
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (stream != nullptr && hipStreamIsCapturing(stream, &cs) == hipSuccess &&
        cs != hipStreamCaptureStatusNone) {
        fail(DG_ERR_INVALID,
             "descriptor table not resident while the stream is being captured; "
             "run the same call once before capture");
        return nullptr;
    }
    if (tc.entries >= kMaxEntries) {
        // Evict the least recently used entry of this device.  Wait for all
        // work first: a kernel in flight may still read it.
        uint64_t oldest = UINT64_MAX, okey = 0;
        size_t oidx = 0;
        for (auto& kv : tc.map[dev])
            for (size_t i = 0; i < kv.second.size(); ++i)
                if (kv.second[i].last_use < oldest) { oldest = kv.second[i].last_use; okey = kv.first; oidx = i; }
        if (oldest != UINT64_MAX) {
            if (hipDeviceSynchronize() != hipSuccess) { fail(DG_ERR_HIP, "cached_table: sync failed"); return nullptr; }
            auto& ob = tc.map[dev][okey];
            (void)hipFree(ob[oidx].dev);
            ob.erase(ob.begin() + oidx);
            if (ob.empty()) tc.map[dev].erase(okey);
            --tc.entries;
        }
    }
    TableEntry e;
    e.bytes.assign(static_cast<const unsigned char*>(host), static_cast<const unsigned char*>(host) + bytes);
    hipError_t err = hipMalloc(&e.dev, bytes);
    if (err != hipSuccess) { fail(DG_ERR_ALLOC, "cached_table: hipMalloc(%zu): %s", bytes, hipGetErrorString(err)); return nullptr; }
    // Synchronous copy from the entry's own host bytes: complete on return.
    err = hipMemcpy(e.dev, e.bytes.data(), bytes, hipMemcpyHostToDevice);
    if (err != hipSuccess) {
        (void)hipFree(e.dev);
        fail(DG_ERR_HIP, "cached_table: hipMemcpy: %s", hipGetErrorString(err));
        return nullptr;
    }
    e.last_use = ++tc.clock;
    void* d = e.dev;
    tc.map[dev][h].push_back(std::move(e));
    ++tc.entries;
    return d;
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
