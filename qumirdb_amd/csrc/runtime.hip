// Runtime: device init, stream, memory pool, pinned staging, errors, timing,
// staging of host batches, downloads, synthetic-data generators.
#include "runtime.h"

#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <map>
#include <mutex>
#include <shared_mutex>
#include <string>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "../../include/qdb_gen.h"
#include "device.h"

namespace qdbgpu {

namespace {

thread_local std::string g_error;
// The stream index this thread issues work on (0: the library stream).
thread_local int t_stream = 0;

struct Runtime {
    std::mutex mu;
    bool ready = false;
    int device = 0;
    hipStream_t stream = nullptr;
    int cus = 256;
    size_t max_lds = 64 * 1024;
    // pool: block size -> free blocks (with the order they were freed in);
    // block -> size
    struct FreeBlock {
        void* p;
        uint64_t freed;
        // The stream that freed it and, once more than one stream is in
        // use, an event recorded there at the free: another stream takes
        // the block only after the event has completed.
        hipStream_t s;
        hipEvent_t ev;
    };
    // Streams: [0] the library stream; [1..] created at first selection
    // (qdb_gpu_set_thread_stream), all at the default priority.
    hipStream_t streams[kMaxStreams] = {};
    bool multi = false;
    std::vector<hipEvent_t> pool_events;   // events of freed blocks (no timing)
    std::multimap<size_t, FreeBlock> free_dev;
    std::unordered_map<void*, size_t> dev_sizes;
    // Blocks in free_dev: a second dfree of one of them is refused (two
    // later dallocs would otherwise receive the same memory).
    std::unordered_set<void*> cached_set;
    uint64_t frees = 0;
    // counters: driver allocations and frees, free-memory queries
    int64_t n_malloc = 0, n_free = 0, n_meminfo = 0;
    size_t cached = 0;          // bytes in free_dev
    size_t reserve = 0;         // device memory the pool leaves free for the HIP runtime and other allocators
    std::map<size_t, std::vector<void*>> free_host;
    std::unordered_map<void*, size_t> host_sizes;
    // timing
    bool timing = false;
    struct Pair { hipEvent_t a, b; };
    std::map<std::string, std::vector<Pair>> timed;
    std::vector<hipEvent_t> spare_events;
};

Runtime& R()
{
    static Runtime* r = new Runtime();
    return *r;
}

size_t size_class(size_t bytes)
{
    if (bytes < 256) return 256;
    if (bytes <= (1u << 20)) {
        size_t c = 256;
        while (c < bytes) c <<= 1;
        return c;
    }
    if (bytes <= ((size_t)64 << 20)) {
        const size_t gran = 2u << 20;
        return (bytes + gran - 1) / gran * gran;
    }
    // Above 64 MiB: four classes per power of two (1, 1.25, 1.5, 1.75 x
    // 2^k), so that buffers of varying size (per-query tables and staging)
    // reuse cached blocks, at most 25% larger than asked.
    size_t p2 = (size_t)64 << 20;
    while (p2 * 2 <= bytes) p2 *= 2;
    for (int q = 1; q <= 4; ++q) {
        const size_t c = p2 + p2 / 4 * (size_t)q;
        if (bytes <= c) return bytes <= p2 ? p2 : c;
    }
    return p2 * 2;
}

}  // namespace

bool fail(const char* fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_error = buf;
    return false;
}

void clear_error()
{
    g_error.clear();
}

const std::string& last_error_text()
{
    return g_error;
}

bool hip_ok(hipError_t e, const char* what)
{
    if (e == hipSuccess) return true;
    return fail("%s: %s", what, hipGetErrorString(e));
}

bool ensure_init()
{
    Runtime& r = R();
    std::lock_guard<std::mutex> lock(r.mu);
    if (r.ready) return true;
    int count = 0;
    if (!hip_ok(hipGetDeviceCount(&count), "hipGetDeviceCount")) return false;
    if (count <= 0) return fail("no HIP device");
    if (!hip_ok(hipSetDevice(r.device), "hipSetDevice")) return false;
    hipDeviceProp_t prop;
    if (!hip_ok(hipGetDeviceProperties(&prop, r.device), "hipGetDeviceProperties")) return false;
    r.cus = prop.multiProcessorCount;
    // The pool leaves 1/8 of the device memory (36 GB of 288 GB) free.
    r.reserve = prop.totalGlobalMem / 8;
    // Largest dynamic LDS allocation of one workgroup (160 KiB on gfx950).
    int lds = 0;
    if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, r.device) == hipSuccess && lds > 0) {
        r.max_lds = (size_t)lds;
    }
    // Default priority, non-blocking with respect to the legacy stream.
    if (!hip_ok(hipStreamCreateWithFlags(&r.stream, hipStreamNonBlocking), "hipStreamCreate")) return false;
    r.streams[0] = r.stream;
    r.ready = true;
    return true;
}

hipStream_t stream()
{
    Runtime& r = R();
    return t_stream == 0 ? r.stream : r.streams[t_stream];
}

bool set_thread_stream(int index)
{
    if (index < 0 || index >= kMaxStreams) return fail("stream index %d outside [0, %d)", index, kMaxStreams);
    if (!ensure_init()) return false;
    Runtime& r = R();
    // A new thread's current device is device 0: select the library's.
    if (!hip_ok(hipSetDevice(r.device), "hipSetDevice")) return false;
    std::lock_guard<std::mutex> lock(r.mu);
    if (!r.streams[index]) {
        // Default priority, non-blocking with respect to the legacy stream.
        if (!hip_ok(hipStreamCreateWithFlags(&r.streams[index], hipStreamNonBlocking), "hipStreamCreate")) return false;
    }
    if (index != 0) r.multi = true;
    t_stream = index;
    return true;
}

int thread_stream_index()
{
    return t_stream;
}

size_t max_lds_per_block()
{
    return R().max_lds;
}

}  // namespace qdbgpu

// [compute units, largest LDS allocation per workgroup in bytes]; returns
// how many values were written, or -1 when the device is unavailable.
extern "C" int64_t qdb_gpu_device_limits(int64_t* out, int64_t n)
{
    if (!qdbgpu::ensure_init()) return -1;
    const int64_t v[2] = {(int64_t)qdbgpu::num_cus(), (int64_t)qdbgpu::max_lds_per_block()};
    int64_t k = 0;
    for (; k < n && k < 2; ++k) out[k] = v[k];
    return k;
}

namespace qdbgpu {

int num_cus()
{
    return R().cus;
}

// Largest cached block a request of class `c` may take: blocks above
// 64 MiB serve requests down to 2/3 of their size (per-query tables and
// staging vary in size from call to call); smaller blocks serve their own
// class only.
std::shared_mutex& capture_mutex()
{
    static std::shared_mutex m;
    return m;
}

static size_t reuse_limit(size_t c)
{
    return c > ((size_t)64 << 20) ? c + c / 2 : c;
}

static void evict_oldest_locked(Runtime& r)
{
    auto oldest = r.free_dev.begin();
    for (auto f = r.free_dev.begin(); f != r.free_dev.end(); ++f) {
        if (f->second.freed < oldest->second.freed) oldest = f;
    }
    r.cached -= oldest->first;
    r.dev_sizes.erase(oldest->second.p);
    r.cached_set.erase(oldest->second.p);
    if (oldest->second.ev) r.pool_events.push_back(oldest->second.ev);
    {
        std::shared_lock<std::shared_mutex> g(capture_mutex());
        (void)hipFree(oldest->second.p);
    }
    r.n_free += 1;
    r.free_dev.erase(oldest);
}

// Returns the blocks cached longest to the HIP runtime until the device has
// `need` bytes free beyond the reserve (or nothing is cached).
static void make_room_locked(Runtime& r, size_t need)
{
    while (!r.free_dev.empty()) {
        size_t free_b = 0, total_b = 0;
        r.n_meminfo += 1;
        if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) {
            (void)hipGetLastError();
            return;
        }
        if (free_b >= need + r.reserve) return;
        evict_oldest_locked(r);
    }
}

static void release_cached_locked(Runtime& r)
{
    for (auto& kv : r.free_dev) {
        r.dev_sizes.erase(kv.second.p);
        if (kv.second.ev) r.pool_events.push_back(kv.second.ev);
        {
            std::shared_lock<std::shared_mutex> g(capture_mutex());
            (void)hipFree(kv.second.p);
        }
        r.n_free += 1;
    }
    r.free_dev.clear();
    r.cached_set.clear();
    r.cached = 0;
}

void* dalloc(size_t bytes)
{
    if (!ensure_init()) return nullptr;
    Runtime& r = R();
    const size_t c = size_class(bytes);
    {
        std::lock_guard<std::mutex> lock(r.mu);
        // Best fit: the smallest cached block of at least class c that
        // this stream may use now (freed on it, or its free has completed).
        const hipStream_t cur = stream();
        for (auto it = r.free_dev.lower_bound(c); it != r.free_dev.end() && it->first <= reuse_limit(c); ++it) {
            Runtime::FreeBlock& b = it->second;
            if (b.ev && b.s != cur) {
                const hipError_t q = hipEventQuery(b.ev);
                if (q != hipSuccess) {
                    (void)hipGetLastError();
                    continue;
                }
            }
            void* p = b.p;
            if (b.ev) r.pool_events.push_back(b.ev);
            r.cached -= it->first;
            r.free_dev.erase(it);
            r.cached_set.erase(p);
            return p;
        }
    }
    void* p = nullptr;
    {
        // A new block leaves the reserve free, if cached blocks can make
        // room for it.
        std::lock_guard<std::mutex> lock(r.mu);
        if (r.cached) make_room_locked(r, c);
    }
    hipError_t e;
    {
        std::shared_lock<std::shared_mutex> g(capture_mutex());
        e = hipMalloc(&p, c);
    }
    if (e != hipSuccess) {
        // Return cached blocks to the HIP runtime and retry once.
        std::lock_guard<std::mutex> lock(r.mu);
        release_cached_locked(r);
        (void)hipGetLastError();
        {
            std::shared_lock<std::shared_mutex> g(capture_mutex());
            e = hipMalloc(&p, c);
        }
        if (e != hipSuccess) {
            fail("hipMalloc(%zu): %s", c, hipGetErrorString(e));
            return nullptr;
        }
    }
    std::lock_guard<std::mutex> lock(r.mu);
    r.dev_sizes[p] = c;
    r.n_malloc += 1;
    return p;
}

void dfree(void* p)
{
    if (!p) return;
    Runtime& r = R();
    std::lock_guard<std::mutex> lock(r.mu);
    auto it = r.dev_sizes.find(p);
    if (it == r.dev_sizes.end()) return;
    if (!r.cached_set.insert(p).second) {
        // Already free: keep the pool consistent and say so.
        fail("device pool: block %p freed twice", p);
        return;
    }
    const size_t sz = it->second;
    hipEvent_t ev = nullptr;
    if (r.multi) {
        // Work queued on this stream may still use the block.
        if (!r.pool_events.empty()) {
            ev = r.pool_events.back();
            r.pool_events.pop_back();
        } else if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) {
            (void)hipGetLastError();
            ev = nullptr;
        }
        if (ev && hipEventRecord(ev, stream()) != hipSuccess) {
            (void)hipGetLastError();
            (void)hipStreamSynchronize(stream());
            r.pool_events.push_back(ev);
            ev = nullptr;
        }
    }
    r.free_dev.emplace(sz, Runtime::FreeBlock{p, ++r.frees, stream(), ev});
    r.cached += sz;
    // Freed blocks stay cached while the device keeps the reserve free (the
    // HIP runtime and other allocators need memory of their own); past
    // that, the blocks cached longest go back to the HIP runtime first, so the
    // pool follows the sizes the current work asks for. hipFree waits for
    // the device, so a block still in use by queued work is safe to
    // release. Below 1/8 of the device cached, free memory is not checked.
    if (r.cached > r.reserve) make_room_locked(r, 0);
}

void trim_device_cache()
{
    Runtime& r = R();
    std::lock_guard<std::mutex> lock(r.mu);
    release_cached_locked(r);
}

size_t dalloc_size(const void* p)
{
    Runtime& r = R();
    std::lock_guard<std::mutex> lock(r.mu);
    auto it = r.dev_sizes.find(const_cast<void*>(p));
    return it == r.dev_sizes.end() ? 0 : it->second;
}

void* halloc_pinned(size_t bytes)
{
    if (!ensure_init()) return nullptr;
    Runtime& r = R();
    size_t c = size_class(bytes);
    {
        std::lock_guard<std::mutex> lock(r.mu);
        auto it = r.free_host.find(c);
        if (it != r.free_host.end() && !it->second.empty()) {
            void* p = it->second.back();
            it->second.pop_back();
            return p;
        }
    }
    void* p = nullptr;
    if (!hip_ok(hipHostMalloc(&p, c, hipHostMallocDefault), "hipHostMalloc")) return nullptr;
    std::lock_guard<std::mutex> lock(r.mu);
    r.host_sizes[p] = c;
    return p;
}

void hfree_pinned(void* p)
{
    if (!p) return;
    Runtime& r = R();
    std::lock_guard<std::mutex> lock(r.mu);
    auto it = r.host_sizes.find(p);
    if (it == r.host_sizes.end()) return;
    r.free_host[it->second].push_back(p);
}

bool sync(const char* what)
{
    return hip_ok(hipStreamSynchronize(stream()), what);
}

__global__ void k_zero_fill(uint64_t* p, size_t words, uint8_t* tail, size_t tail_bytes)
{
    size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t w = i; w < words; w += stride) p[w] = 0;
    if (i < tail_bytes) tail[i] = 0;
}

bool zero_fill(void* p, size_t bytes)
{
    if (bytes == 0) return true;
    size_t words = bytes / 8;
    uint8_t* tail = (uint8_t*)p + words * 8;
    int grid = grid_for((int64_t)(words > 64 ? words : 64), 256, 8);
    hipLaunchKernelGGL(k_zero_fill, dim3(grid), dim3(256), 0, stream(), (uint64_t*)p, words, tail, bytes - words * 8);
    return hip_ok(hipGetLastError(), "zero_fill launch");
}

int grid_for(int64_t items, int block, int per_cu)
{
    int64_t need = (items + block - 1) / block;
    int64_t cap = (int64_t)num_cus() * per_cu;
    if (need < 1) need = 1;
    return (int)(need < cap ? need : cap);
}

void* upload(const void* host, size_t bytes)
{
    void* d = dalloc(bytes);
    if (!d) return nullptr;
    if (!hip_ok(hipMemcpyAsync(d, host, bytes, hipMemcpyHostToDevice, stream()), "upload")) return nullptr;
    if (!sync("upload")) return nullptr;
    return d;
}

TimedRegion::TimedRegion(const char* family) : family_(family)
{
    Runtime& r = R();
    if (!r.timing) return;
    if (hipEventCreate(&start_) != hipSuccess) {
        start_ = nullptr;
        return;
    }
    (void)hipEventRecord(start_, stream());
}

TimedRegion::~TimedRegion()
{
    if (!start_) return;
    Runtime& r = R();
    hipEvent_t end = nullptr;
    if (hipEventCreate(&end) != hipSuccess) return;
    (void)hipEventRecord(end, stream());
    std::lock_guard<std::mutex> lock(r.mu);
    r.timed[family_].push_back({start_, end});
}

}  // namespace qdbgpu

using namespace qdbgpu;

// ---------------------------------------------------------------------------
// C-ABI: runtime
// ---------------------------------------------------------------------------

extern "C" int qdb_gpu_init(int device)
{
    Runtime& r = R();
    {
        std::lock_guard<std::mutex> lock(r.mu);
        if (!r.ready) r.device = device;
    }
    return ensure_init() ? 0 : -1;
}

extern "C" const char* qdb_gpu_last_error(void)
{
    return g_error.c_str();
}

extern "C" const char* qdb_gpu_version(void)
{
    return "qdb_gpu 0.1 gfx950";
}

extern "C" void* qdb_gpu_alloc(size_t bytes)
{
    return dalloc(bytes);
}

// Device pool counters: [driver allocations, driver frees, free-memory
// queries, cached bytes, bytes in blocks handed out]; returns how many
// values were written.
extern "C" int64_t qdb_gpu_pool_stats(int64_t* out, int64_t n)
{
    qdbgpu::Runtime& r = qdbgpu::R();
    std::lock_guard<std::mutex> lock(r.mu);
    size_t all = 0;
    for (const auto& kv : r.dev_sizes) all += kv.second;
    const int64_t v[5] = {r.n_malloc, r.n_free, r.n_meminfo, (int64_t)r.cached, (int64_t)(all - r.cached)};
    int64_t k = 0;
    for (; k < n && k < 5; ++k) out[k] = v[k];
    return k;
}

extern "C" void qdb_gpu_trim_cache(void)
{
    if (!qdbgpu::ensure_init()) return;
    (void)hipStreamSynchronize(qdbgpu::stream());
    qdbgpu::trim_device_cache();
}

extern "C" void qdb_gpu_free(void* p)
{
    dfree(p);
}

extern "C" int qdb_gpu_memcpy_h2d(void* dst, const void* src, size_t bytes)
{
    if (!ensure_init()) return -1;
    if (bytes == 0) return 0;
    // Through a pinned staging buffer so that the copy runs at DMA speed.
    const size_t chunk = 64u << 20;
    void* pin = halloc_pinned(bytes < chunk ? bytes : chunk);
    if (!pin) return -1;
    for (size_t off = 0; off < bytes; off += chunk) {
        size_t n = bytes - off < chunk ? bytes - off : chunk;
        memcpy(pin, (const char*)src + off, n);
        if (!hip_ok(hipMemcpyAsync((char*)dst + off, pin, n, hipMemcpyHostToDevice, stream()), "h2d") || !sync("h2d")) {
            hfree_pinned(pin);
            return -1;
        }
    }
    hfree_pinned(pin);
    return 0;
}

extern "C" int qdb_gpu_memcpy_d2h(void* dst, const void* src, size_t bytes)
{
    if (!ensure_init()) return -1;
    if (bytes == 0) return 0;
    if (!hip_ok(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, stream()), "d2h")) return -1;
    return sync("d2h") ? 0 : -1;
}

extern "C" int qdb_gpu_memset(void* dst, int value, size_t bytes)
{
    if (!ensure_init()) return -1;
    if (value == 0) return zero_fill(dst, bytes) && sync("memset") ? 0 : -1;
    if (!hip_ok(hipMemsetAsync(dst, value, bytes, stream()), "memset")) return -1;
    return sync("memset") ? 0 : -1;
}

extern "C" int qdb_gpu_synchronize(void)
{
    if (!ensure_init()) return -1;
    return sync("synchronize") ? 0 : -1;
}

extern "C" void qdb_gpu_timing_enable(int on)
{
    R().timing = on != 0;
}

extern "C" void qdb_gpu_timing_reset(void)
{
    Runtime& r = R();
    (void)hipStreamSynchronize(stream());
    std::lock_guard<std::mutex> lock(r.mu);
    for (auto& kv : r.timed) {
        for (auto& p : kv.second) {
            (void)hipEventDestroy(p.a);
            (void)hipEventDestroy(p.b);
        }
    }
    r.timed.clear();
}

extern "C" double qdb_gpu_timing_avg_ms(const char* family, int64_t* launches)
{
    Runtime& r = R();
    (void)hipStreamSynchronize(stream());
    std::lock_guard<std::mutex> lock(r.mu);
    auto it = r.timed.find(family);
    if (launches) *launches = 0;
    if (it == r.timed.end() || it->second.empty()) return 0.0;
    double total = 0.0;
    for (auto& p : it->second) {
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, p.a, p.b);
        total += ms;
    }
    if (launches) *launches = (int64_t)it->second.size();
    return total / (double)it->second.size();
}

// ---------------------------------------------------------------------------
// Staging: host TRowSet -> device TRowSet through pinned buffers.
// ---------------------------------------------------------------------------

namespace {

// Up to kSegs device byte ranges packed into one device buffer by one
// launch (blockIdx.y = segment).
constexpr int kSegs = 32;
struct CopySegs {
    const char* src[kSegs];
    uint64_t off[kSegs];
    uint64_t bytes[kSegs];
};

__global__ void k_gather_segments(CopySegs S, char* dst)
{
    const int g = blockIdx.y;
    const uint64_t n = S.bytes[g];
    const char* src = S.src[g];
    char* d = dst + S.off[g];
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += stride) d[i] = src[i];
}

// Copies device byte ranges to new host (malloc) buffers, one per range
// (nullptr for an empty range); false on a failed copy. Small totals are
// packed on the device and come back in one copy through pinned memory
// (one round trip instead of one per range); larger ones are copied range
// by range.
bool download_ranges(const std::vector<const void*>& src, const std::vector<size_t>& bytes, std::vector<void*>* host)
{
    const size_t k = src.size();
    host->assign(k, nullptr);
    std::vector<uint64_t> off(k, 0);
    size_t total = 0;
    for (size_t i = 0; i < k; ++i) {
        if (!src[i] || !bytes[i]) continue;
        off[i] = total;
        total += (bytes[i] + 15) / 16 * 16;
        (*host)[i] = heap_alloc(bytes[i]);
        if (!(*host)[i]) return fail("download: out of host memory");
    }
    if (total == 0) return true;
    if (total > ((size_t)16 << 20)) {
        for (size_t i = 0; i < k; ++i) {
            if (!(*host)[i]) continue;
            if (!hip_ok(hipMemcpyAsync((*host)[i], src[i], bytes[i], hipMemcpyDeviceToHost, stream()), "download")) return false;
        }
        return sync("download");
    }
    char* packed = (char*)dalloc(total);
    char* pin = (char*)halloc_pinned(total);
    bool ok = packed && pin;
    for (size_t i0 = 0; ok && i0 < k; i0 += kSegs) {
        CopySegs S;
        memset(&S, 0, sizeof(S));
        int n = 0;
        uint64_t most = 0;
        for (size_t i = i0; i < k && i < i0 + kSegs; ++i) {
            if (!(*host)[i]) continue;
            S.src[n] = (const char*)src[i];
            S.off[n] = off[i];
            S.bytes[n] = bytes[i];
            most = bytes[i] > most ? bytes[i] : most;
            ++n;
        }
        if (n == 0) continue;
        const int gx = (int)((most + 255) / 256 < 64 ? (most + 255) / 256 : 64);
        hipLaunchKernelGGL(k_gather_segments, dim3(gx > 0 ? gx : 1, n), dim3(256), 0, stream(), S, packed);
        ok = hip_ok(hipGetLastError(), "download gather");
    }
    ok = ok && hip_ok(hipMemcpyAsync(pin, packed, total, hipMemcpyDeviceToHost, stream()), "download") && sync("download");
    if (ok) {
        for (size_t i = 0; i < k; ++i) {
            if ((*host)[i]) memcpy((*host)[i], pin + off[i], bytes[i]);
        }
    }
    dfree(packed);
    hfree_pinned(pin);
    return ok;
}

// The builder of an output rowset keeps its TColumns and buffers until the
// rowset is published; a rowset still being built carries the builder in
// Private and destroy_builder as Destroy, so that a failure on the way frees
// everything.
void destroy_builder(TRowSet* rs)
{
    OwnedRowSet* o = (OwnedRowSet*)rs->Private;
    if (!o) return;
    for (void* p : o->dev) dfree(p);
    for (void* p : o->host) heap_free(p);
    delete o;
    rs->Private = nullptr;
    rs->Columns = nullptr;
}

// A published rowset's device buffers are listed in one host block of its
// owners list: [kDeviceLease, count, pointers...].
constexpr int64_t kDeviceLease = 0x5145444c45415345ll;

}  // namespace

namespace qdbgpu {

// ---------------------------------------------------------------------------
// Host heap and kernel-owned output rowsets.
// ---------------------------------------------------------------------------

namespace {

void* default_alloc(int64_t size)
{
    return malloc(size > 0 ? (size_t)size : 1);
}

void default_free(void* p)
{
    free(p);
}

struct HostHeap {
    QdbHeapAllocFn alloc = default_alloc;
    QdbHeapFreeFn free = default_free;
    QdbHeapDetachFn detach = nullptr;
};

HostHeap& heap()
{
    static HostHeap h;
    return h;
}

}  // namespace

void* heap_alloc(size_t bytes)
{
    return heap().alloc((int64_t)bytes);
}

void heap_free(void* p)
{
    if (p) heap().free(p);
}

void heap_detach(void* p)
{
    if (p && heap().detach) heap().detach(p);
}

OwnedRowSet* init_owned_rowset(TRowSet* out, int64_t rows, int ncols)
{
    OwnedRowSet* o = new OwnedRowSet();
    o->cols.resize((size_t)(ncols > 0 ? ncols : 1));
    memset(o->cols.data(), 0, o->cols.size() * sizeof(TColumn));
    memset(out, 0, sizeof(*out));
    out->Columns = o->cols.data();
    out->ColumnCount = ncols;
    out->RowCount = rows;
    out->Destroy = destroy_builder;
    out->Private = o;
    out->RefCount = 1;
    return o;
}

// The kernel-owned form of qdb/exec/kernel_rowset.h:14-38: Columns and every
// host buffer come from the host heap, Private is the owners list
// (owners[0] = count, then the blocks, the list itself allocated from the
// heap too) and Destroy is NULL: the caller adopts the rowset and installs its
// destroy callback, as the reference's processors do after
// agg_finish_rowset. Device buffers (an output left in HBM) are listed in one
// lease block of the owners list, which qdb_gpu_destroy_kernel_rowset
// returns to the device pool.
bool publish_rowset(TRowSet* rs)
{
    OwnedRowSet* o = (OwnedRowSet*)rs->Private;
    if (!o) return fail("publish: not a rowset under construction");
    const size_t ncols = (size_t)(rs->ColumnCount > 0 ? rs->ColumnCount : 0);
    TColumn* cols = (TColumn*)heap_alloc(std::max<size_t>(ncols, 1) * sizeof(TColumn));
    int64_t* lease = nullptr;
    if (!o->dev.empty()) lease = (int64_t*)heap_alloc((o->dev.size() + 2) * sizeof(int64_t));
    const size_t nowners = 1 + (lease ? 1 : 0) + o->host.size();
    int64_t* owners = (int64_t*)heap_alloc((nowners + 1) * sizeof(int64_t));
    if (!cols || (!o->dev.empty() && !lease) || !owners) {
        heap_free(cols);
        heap_free(lease);
        heap_free(owners);
        destroy_builder(rs);
        return fail("publish: out of host memory");
    }
    if (ncols) memcpy(cols, o->cols.data(), ncols * sizeof(TColumn));
    size_t k = 0;
    owners[++k] = (int64_t)(intptr_t)cols;
    if (lease) {
        lease[0] = kDeviceLease;
        lease[1] = (int64_t)o->dev.size();
        for (size_t i = 0; i < o->dev.size(); ++i) lease[2 + i] = (int64_t)(intptr_t)o->dev[i];
        owners[++k] = (int64_t)(intptr_t)lease;
    }
    for (void* p : o->host) owners[++k] = (int64_t)(intptr_t)p;
    owners[0] = (int64_t)k;
    delete o;
    rs->Columns = cols;
    rs->Private = owners;
    rs->Destroy = nullptr;
    return true;
}

void destroy_kernel_rowset(TRowSet* rs)
{
    if (!rs) return;
    if (rs->Destroy == destroy_builder) {
        destroy_builder(rs);
        return;
    }
    int64_t* owners = (int64_t*)rs->Private;
    if (!owners) return;
    for (int64_t i = 1; i <= owners[0]; ++i) {
        int64_t* block = (int64_t*)(intptr_t)owners[i];
        if (block && block[0] == kDeviceLease) {
            for (int64_t d = 0; d < block[1]; ++d) dfree((void*)(intptr_t)block[2 + d]);
        }
        heap_free(block);
    }
    heap_free(owners);
    rs->Private = nullptr;
    rs->Columns = nullptr;
    rs->Destroy = nullptr;
}

bool move_rowset_to_host(TRowSet* rs, const QdbValueType* types)
{
    OwnedRowSet* o = (OwnedRowSet*)rs->Private;
    const int64_t n = rs->RowCount;
    // Every data and mask buffer, then one download for all of them.
    std::vector<const void*> src;
    std::vector<size_t> bytes;
    for (int64_t c = 0; c < rs->ColumnCount; ++c) {
        const TColumn& col = o->cols[(size_t)c];
        const QdbValueType& t = types[c];
        src.push_back(col.Data);
        bytes.push_back(t.Kind == QDB_KIND_BOOL ? (size_t)((n + 7) / 8) : (size_t)n * (size_t)t.Width);
        src.push_back(col.Mask);
        bytes.push_back(col.Mask ? (size_t)((n + 7) / 8) : 0);
    }
    std::vector<void*> host;
    const bool ok = download_ranges(src, bytes, &host);
    for (size_t i = 0; i < host.size(); ++i) {
        if (!host[i] && (i % 2 == 0 || src[i])) host[i] = heap_alloc(1);  // a buffer for every column and mask present
    }
    if (!ok) {
        for (void* p : host) heap_free(p);
        return false;
    }
    for (int64_t c = 0; c < rs->ColumnCount; ++c) {
        TColumn& col = o->cols[(size_t)c];
        col.Data = (char*)host[2 * (size_t)c];
        col.Mask = col.Mask ? (const uint8_t*)host[2 * (size_t)c + 1] : nullptr;
    }
    for (void* p : o->dev) dfree(p);
    o->dev.clear();
    for (void* p : host) {
        if (p) o->host.push_back(p);
    }
    return true;
}

}  // namespace qdbgpu

extern "C" void qdb_gpu_set_host_heap(QdbHeapAllocFn alloc, QdbHeapFreeFn free_fn, QdbHeapDetachFn detach)
{
    HostHeap& h = heap();
    h.alloc = alloc ? alloc : default_alloc;
    h.free = free_fn ? free_fn : default_free;
    h.detach = detach;
}

extern "C" void qdb_gpu_destroy_kernel_rowset(TRowSet* rs)
{
    destroy_kernel_rowset(rs);
}

extern "C" void qdb_gpu_adopt_kernel_rowset(TRowSet* rs)
{
    if (!rs) return;
    int64_t* owners = (int64_t*)rs->Private;
    if (owners) {
        for (int64_t i = 1; i <= owners[0]; ++i) heap_detach((void*)(intptr_t)owners[i]);
        heap_detach(owners);
    }
    rs->Destroy = qdb_gpu_destroy_kernel_rowset;
}

extern "C" int qdb_gpu_stage_rowset(const TRowSet* host, const QdbColumnSchema* types, TRowSet* out)
{
    clear_error();
    memset(out, 0, sizeof(*out));
    if (!ensure_init()) return -1;
    if (host->ColumnCount < 0 || host->ColumnCount > types->NumColumns) return fail("stage: %lld columns, schema has %d", (long long)host->ColumnCount, types->NumColumns), -1;
    const int64_t n = host->RowCount;
    // Every buffer's bytes: data and mask per column, then selection and hash.
    std::vector<const void*> src;
    std::vector<size_t> bytes;
    for (int64_t c = 0; c < host->ColumnCount; ++c) {
        const TColumn& hc = host->Columns[c];
        const QdbValueType& t = types->Types[c];
        if (t.Kind == QDB_KIND_STRING) return fail("stage: string columns are not supported"), -1;
        src.push_back(hc.Data);
        bytes.push_back(t.Kind == QDB_KIND_BOOL ? (size_t)((n + hc.DataBitOffset + 7) / 8) : (size_t)n * (size_t)t.Width);
        src.push_back(hc.Mask);
        bytes.push_back(hc.Mask ? (size_t)((n + hc.MaskBitOffset + 7) / 8) : 0);
    }
    src.push_back(host->Selection);
    bytes.push_back(host->Selection ? (size_t)n : 0);
    src.push_back(host->Hash);
    bytes.push_back(host->Hash ? (size_t)n * 8 : 0);
    // One pinned staging buffer, one device block per buffer, every copy
    // queued before one synchronization.
    size_t total = 0;
    std::vector<size_t> off(src.size(), 0);
    for (size_t i = 0; i < src.size(); ++i) {
        off[i] = total;
        if (src[i] && bytes[i]) total += (bytes[i] + 255) / 256 * 256;
    }
    OwnedRowSet* o = init_owned_rowset(out, n, (int)host->ColumnCount);
    std::vector<char*> dev(src.size(), nullptr);
    char* pin = total ? (char*)halloc_pinned(total) : nullptr;
    bool ok = total == 0 || pin;
    for (size_t i = 0; ok && i < src.size(); ++i) {
        if (!src[i] || !bytes[i]) continue;
        dev[i] = (char*)dalloc(bytes[i]);
        ok = dev[i] != nullptr;
        if (!ok) break;
        o->dev.push_back(dev[i]);
        memcpy(pin + off[i], src[i], bytes[i]);
        ok = hip_ok(hipMemcpyAsync(dev[i], pin + off[i], bytes[i], hipMemcpyHostToDevice, stream()), "stage copy");
    }
    ok = ok && sync("stage");
    hfree_pinned(pin);
    if (!ok) {
        destroy_builder(out);
        memset(out, 0, sizeof(*out));
        return -1;
    }
    for (int64_t c = 0; c < host->ColumnCount; ++c) {
        const TColumn& hc = host->Columns[c];
        TColumn& dc = o->cols[(size_t)c];
        dc.Data = dev[2 * (size_t)c];
        dc.DataBitOffset = hc.DataBitOffset;
        dc.Mask = (const uint8_t*)dev[2 * (size_t)c + 1];
        dc.MaskBitOffset = hc.Mask ? hc.MaskBitOffset : 0;
    }
    const size_t k = 2 * (size_t)host->ColumnCount;
    out->Selection = (uint8_t*)dev[k];
    out->Hash = (uint64_t*)dev[k + 1];
    if (!publish_rowset(out)) return -1;
    out->Destroy = qdb_gpu_destroy_kernel_rowset;
    return 0;
}

extern "C" int qdb_gpu_download_rowset(const TRowSet* dev, const QdbValueType* types, TRowSet* out)
{
    clear_error();
    memset(out, 0, sizeof(*out));
    if (!ensure_init()) return -1;
    const int64_t n = dev->RowCount;
    // Every data and mask buffer, then one download for all of them.
    std::vector<const void*> src;
    std::vector<size_t> bytes;
    for (int64_t c = 0; c < dev->ColumnCount; ++c) {
        const TColumn& dc = dev->Columns[c];
        const QdbValueType& t = types[c];
        src.push_back(dc.Data);
        bytes.push_back(t.Kind == QDB_KIND_BOOL ? (size_t)((n + dc.DataBitOffset + 7) / 8) : (size_t)n * (size_t)t.Width);
        src.push_back(dc.Mask);
        bytes.push_back(dc.Mask ? (size_t)((n + dc.MaskBitOffset + 7) / 8) : 0);
    }
    OwnedRowSet* o = init_owned_rowset(out, n, (int)dev->ColumnCount);
    std::vector<void*> host;
    const bool fetched = download_ranges(src, bytes, &host);
    for (void* p : host) {
        if (p) o->host.push_back(p);
    }
    if (!fetched) {
        destroy_builder(out);
        memset(out, 0, sizeof(*out));
        return -1;
    }
    for (int64_t c = 0; c < dev->ColumnCount; ++c) {
        const TColumn& dc = dev->Columns[c];
        TColumn& hc = o->cols[(size_t)c];
        hc.Data = (char*)host[2 * (size_t)c];
        hc.DataBitOffset = dc.DataBitOffset;
        if (dc.Mask) {
            hc.Mask = (const uint8_t*)host[2 * (size_t)c + 1];
            hc.MaskBitOffset = dc.MaskBitOffset;
        }
    }
    if (!publish_rowset(out)) return -1;
    out->Destroy = qdb_gpu_destroy_kernel_rowset;
    return 0;
}

// ---------------------------------------------------------------------------
// Synthetic inputs in HBM (qdb_gen.h).
// ---------------------------------------------------------------------------

__global__ void k_gen_lineitem(uint64_t seed, int64_t row0, int64_t rows, int8_t* flag, int8_t* status,
    int32_t* shipdate, int64_t* quantity, int64_t* price, int64_t* discount, int64_t* tax)
{
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < rows; i += stride) {
        qdb_gen_lineitem_row(seed, row0 + i, &flag[i], &status[i], &shipdate[i], &quantity[i], &price[i], &discount[i], &tax[i]);
    }
}

__global__ void k_gen_join_build(uint64_t seed, int64_t row0, int64_t rows, int64_t build_rows, int64_t* key, int32_t* segment)
{
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < rows; i += stride) {
        qdb_gen_join_build_row(seed, row0 + i, build_rows, &key[i], &segment[i]);
    }
}

__global__ void k_gen_join_probe(uint64_t seed, int64_t row0, int64_t rows, int64_t build_rows, int64_t* fk, int64_t* value)
{
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < rows; i += stride) {
        qdb_gen_join_probe_row(seed, row0 + i, build_rows, &fk[i], &value[i]);
    }
}

__global__ void k_gen_groups(uint64_t seed, int64_t row0, int64_t rows, const double* cdf, int64_t distinct,
    int64_t* key, int64_t* vi, double* vf)
{
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < rows; i += stride) {
        qdb_gen_groups_row(seed, row0 + i, cdf, distinct, &key[i], &vi[i], &vf[i]);
    }
}

extern "C" int qdb_gpu_gen_lineitem(uint64_t seed, int64_t row0, int64_t rows, int8_t* flag, int8_t* status,
    int32_t* shipdate, int64_t* quantity, int64_t* price, int64_t* discount, int64_t* tax)
{
    if (!ensure_init()) return -1;
    if (rows <= 0) return 0;
    hipLaunchKernelGGL(k_gen_lineitem, dim3(grid_for(rows, 256, 16)), dim3(256), 0, stream(), seed, row0, rows, flag,
        status, shipdate, quantity, price, discount, tax);
    return hip_ok(hipGetLastError(), "gen_lineitem") && sync("gen_lineitem") ? 0 : -1;
}

extern "C" int qdb_gpu_gen_join_build(uint64_t seed, int64_t row0, int64_t rows, int64_t* key, int32_t* segment)
{
    if (!ensure_init()) return -1;
    if (rows <= 0) return 0;
    // The build table is the permutation of 1..rows of the whole table.
    hipLaunchKernelGGL(k_gen_join_build, dim3(grid_for(rows, 256, 16)), dim3(256), 0, stream(), seed, row0, rows,
        row0 + rows, key, segment);
    return hip_ok(hipGetLastError(), "gen_join_build") && sync("gen_join_build") ? 0 : -1;
}

extern "C" int qdb_gpu_gen_join_probe(uint64_t seed, int64_t row0, int64_t rows, int64_t build_rows, int64_t* fk, int64_t* value)
{
    if (!ensure_init()) return -1;
    if (rows <= 0) return 0;
    hipLaunchKernelGGL(k_gen_join_probe, dim3(grid_for(rows, 256, 16)), dim3(256), 0, stream(), seed, row0, rows,
        build_rows, fk, value);
    return hip_ok(hipGetLastError(), "gen_join_probe") && sync("gen_join_probe") ? 0 : -1;
}

extern "C" int qdb_gpu_gen_zipf_groups(uint64_t seed, int64_t row0, int64_t rows, const double* zipf_cdf,
    int64_t distinct, int64_t* key, int64_t* vi, double* vf)
{
    if (!ensure_init()) return -1;
    if (rows <= 0) return 0;
    hipLaunchKernelGGL(k_gen_groups, dim3(grid_for(rows, 256, 16)), dim3(256), 0, stream(), seed, row0, rows, zipf_cdf,
        distinct, key, vi, vf);
    return hip_ok(hipGetLastError(), "gen_groups") && sync("gen_groups") ? 0 : -1;
}

// Sizes of the ABI structs, in the order the header declares them, so that
// bindings can check their layouts.
extern "C" int qdb_gpu_set_thread_stream(int index)
{
    qdbgpu::clear_error();
    return qdbgpu::set_thread_stream(index) ? 0 : -1;
}

extern "C" int64_t qdb_gpu_struct_sizes(int64_t* out, int64_t n)
{
    const int64_t sizes[] = {
        (int64_t)sizeof(TColumn), (int64_t)sizeof(TRowSet), (int64_t)sizeof(QdbValueType), (int64_t)sizeof(QdbExprInsn),
        (int64_t)sizeof(QdbExprProgram), (int64_t)sizeof(QdbColumnSchema), (int64_t)sizeof(QdbFilterSpec),
        (int64_t)sizeof(QdbProjectSpec), (int64_t)sizeof(QdbAggregateFn), (int64_t)sizeof(QdbAggregateSpec),
        (int64_t)sizeof(QdbAggregateKernels), (int64_t)sizeof(QdbRowJoinSpec), (int64_t)sizeof(QdbRowJoinKernels),
        (int64_t)sizeof(QdbPairBuffer), (int64_t)sizeof(QdbFusedResult), (int64_t)sizeof(QdbTuneEntry),
    };
    const int64_t count = (int64_t)(sizeof(sizes) / sizeof(sizes[0]));
    for (int64_t i = 0; i < n && i < count; ++i) out[i] = sizes[i];
    return count;
}
