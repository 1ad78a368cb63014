// Runtime join filter on gfx950: the probe-side filter a filtered build
// publishes for its join (qdb/exec/runtime_filter.cpp,
// qdb/kernel/join/runtime_filter.oz; docs/arch/scheduler_runtime.md:383-443).
//
// Contents follow the reference's kernel, so a probe keeps exactly the rows
// the reference keeps for the same hashes:
//   - the filter holds the selected build rows' join-key hashes (rh_hash);
//   - up to MaxKeys (16K) distinct hashes it is an EXACT set: an open-
//     addressing table of pow2 >= 2 * distinct (at least 8) u64 slots with
//     linear probing from hash & mask, 0 = empty, hash 0 stored as 1
//     (rf_exact_*); a probe keeps a row whose hash is in the set;
//   - beyond that it is a blocked BLOOM filter of bit_floor(BloomBytes)
//     bytes (2 MiB by default): a hash picks the 64-byte block
//     rf_mix(hash) & block_mask and, in each of the block's eight words, the
//     bit (rf_mix(hash ^ 0x9E3779B97F4A7C15) >> 8i) & 63 (rf_bloom_*); a
//     probe keeps a row whose eight bits are set;
//   - with less than 64 bytes of Bloom storage and too many keys the filter
//     cannot answer (kind 0): a probe keeps every selected row.
// rf_mix is the splitmix64 finalizer.
//
// Layout in HBM: the build's selected hashes (u64, appended per build
// batch), then either the exact table (u64 slots) or the Bloom words. Probes
// are HBM-streaming over the probe hashes with one 64-byte gather per row
// into an L2-resident filter (2 MiB < one XCD's 4 MB L2).
#include <hip/hip_runtime.h>

#include <string.h>

#include "device.h"
#include "runtime.h"

namespace qdbgpu {

namespace {

constexpr int kBlock = 256;
constexpr int64_t kBloomBytesPerBlock = 64;

__device__ __forceinline__ uint64_t rf_mix(uint64_t v)
{
    v ^= v >> 30;
    v *= 0xBF58476D1CE4E5B9ull;
    v ^= v >> 27;
    v *= 0x94D049BB133111EBull;
    return v ^ (v >> 31);
}

__device__ __forceinline__ uint64_t rf_stored(uint64_t h)
{
    return h == 0 ? 1ull : h;
}

// Selected hashes of a build batch appended at hashes_out[*count...].
__global__ void __launch_bounds__(kBlock) k_rf_collect(const uint64_t* hashes, const uint8_t* sel, int64_t n,
    uint64_t* out, unsigned long long* count)
{
    const int lane = lane_id();
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t base = (int64_t)blockIdx.x * blockDim.x + threadIdx.x - lane; base < n; base += stride) {
        const int64_t i = base + lane;
        const bool keep = i < n && (!sel || sel[i]);
        const uint64_t m = __ballot(keep);
        unsigned long long at = 0;
        if (lane == 0 && m) at = atomicAdd(count, (unsigned long long)__popcll(m));
        at = shfl_u64(at, 0);
        if (keep) out[at + (unsigned long long)__popcll(m & ((1ull << lane) - 1ull))] = hashes[i];
    }
}

constexpr unsigned long long kSetEmpty = ~0ull;

__global__ void __launch_bounds__(kBlock) k_rf_fill(unsigned long long* p, int64_t n, unsigned long long v)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
}

// The distinct raw hashes (rf_set_insert's deduplicating set): each hash is
// inserted into an open-addressing set of at least twice as many slots as
// hashes (empty = all ones; the all-ones hash has a flag word of its own),
// and a hash that was not there yet is appended to `distinct`
// (counters[1] counts them).
__global__ void __launch_bounds__(kBlock) k_rf_dedup(const uint64_t* keys, int64_t n, unsigned long long* set, uint64_t mask,
    uint64_t* distinct, unsigned long long* counters)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const uint64_t h = keys[i];
        bool fresh = false;
        if (h == kSetEmpty) {
            fresh = atomicCAS(&counters[3], 0ull, 1ull) == 0ull;
            if (fresh) distinct[atomicAdd(&counters[1], 1ull)] = h;
            continue;
        }
        uint64_t slot = rf_mix(h) & mask;
        for (;;) {
            const unsigned long long old = atomicCAS(&set[slot], kSetEmpty, (unsigned long long)h);
            if (old == kSetEmpty) {
                // Appended inside the loop (no bool lives out of it).
                distinct[atomicAdd(&counters[1], 1ull)] = h;
                break;
            }
            if (old == (unsigned long long)h) break;
            slot = (slot + 1) & mask;
        }
    }
}

// The exact table: every distinct hash inserted with linear probing.
__global__ void __launch_bounds__(kBlock) k_rf_exact_build(const uint64_t* distinct, int64_t n, unsigned long long* table,
    uint64_t mask)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const uint64_t h = distinct[i];
        const unsigned long long v = rf_stored(h);
        uint64_t slot = h & mask;
        // The table is at least twice the distinct count: a free slot exists.
        for (;;) {
            const unsigned long long old = atomicCAS(&table[slot], 0ull, v);
            if (old == 0ull || old == v) break;
            slot = (slot + 1) & mask;
        }
    }
}

__device__ __forceinline__ void rf_bloom_slots(uint64_t h, uint64_t block_mask, uint64_t* base, uint64_t* bits)
{
    *base = (rf_mix(h) & block_mask) * 8;
    *bits = rf_mix(h ^ 0x9E3779B97F4A7C15ull);
}

__global__ void __launch_bounds__(kBlock) k_rf_bloom_add(const uint64_t* hashes, int64_t n, unsigned long long* words,
    uint64_t block_mask)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        uint64_t base, bits;
        rf_bloom_slots(hashes[i], block_mask, &base, &bits);
#pragma unroll
        for (int w = 0; w < 8; ++w) atomicOr(&words[base + w], 1ull << ((bits >> (8 * w)) & 63));
    }
}

__global__ void __launch_bounds__(kBlock) k_rf_popcount(const unsigned long long* words, int64_t n, unsigned long long* total)
{
    unsigned long long c = 0;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) c += __popcll(words[i]);
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) c += shfl_xor_u64(c, o);
    if (lane_id() == 0 && c) atomicAdd(total, c);
}

// output[i] = selected and in the filter; per-block kept counts into part.
template <int kKind>
__global__ void __launch_bounds__(kBlock) k_rf_probe(const uint64_t* hashes, const uint8_t* sel, int64_t n,
    const unsigned long long* words, uint64_t mask, uint8_t* out, unsigned long long* part)
{
    __shared__ unsigned long long wsum[kBlock / kWave];
    unsigned long long kept = 0;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        bool keep = !sel || sel[i];
        if (keep && kKind == 1) {
            // The answer is stored inside the probe loop: a bool that lives
            // out of this divergent loop was compiled (ROCm 7.2, gfx950) to
            // the LAST iteration's compare mask, which is 0 for every lane
            // that left the loop earlier.
            const uint64_t h = hashes[i];
            const unsigned long long v = rf_stored(h);
            uint64_t slot = h & mask;
            for (;;) {
                const unsigned long long found = words[slot];
                if (found == v) {
                    out[i] = 1;
                    kept += 1;
                    break;
                }
                if (found == 0ull) {
                    out[i] = 0;
                    break;
                }
                slot = (slot + 1) & mask;
            }
            continue;
        } else if (keep && kKind == 2) {
            uint64_t base, bits;
            rf_bloom_slots(hashes[i], mask, &base, &bits);
            // The block's eight words in two 32-byte loads.
            const ulonglong4* blk = (const ulonglong4*)(words + base);
            const ulonglong4 a = blk[0], b = blk[1];
            const unsigned long long wv[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
#pragma unroll
            for (int w = 0; w < 8; ++w) keep = keep && ((wv[w] >> ((bits >> (8 * w)) & 63)) & 1ull);
        }
        out[i] = keep ? 1 : 0;
        kept += keep;
    }
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) kept += shfl_xor_u64(kept, o);
    if (lane_id() == 0) wsum[threadIdx.x / kWave] = kept;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
        for (int w = 0; w < kBlock / kWave; ++w) t += wsum[w];
        part[blockIdx.x] = t;
    }
}

__global__ void k_rf_sum(const unsigned long long* part, int n, unsigned long long* out)
{
    unsigned long long t = 0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) t += part[i];
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) t += shfl_xor_u64(t, o);
    __shared__ unsigned long long w[kBlock / kWave];
    if (lane_id() == 0) w[threadIdx.x / kWave] = t;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long s = 0;
        for (int i = 0; i < (int)(blockDim.x / kWave); ++i) s += w[i];
        *out = s;
    }
}

}  // namespace

struct RuntimeFilterState {
    int64_t max_keys = 0;
    int64_t bloom_bytes = 0;          // bit_floor of the request, 0 below 64
    // build: the selected hashes so far (device), their count (device and host)
    uint64_t* keys = nullptr;
    int64_t keys_cap = 0;
    unsigned long long* counters = nullptr;   // [0] hashes, [1] runs, [2] popcount, [3] kept
    // published
    int kind = -1;                    // -1 building, 0 cannot answer, 1 exact, 2 Bloom
    unsigned long long* words = nullptr;
    int64_t word_count = 0;
    int64_t exact_keys = 0;
    int64_t bloom_set_bits = 0;
    int64_t build_rows = 0;
    ~RuntimeFilterState()
    {
        dfree(keys);
        dfree(counters);
        dfree(words);
    }
};

}  // namespace qdbgpu

using namespace qdbgpu;

extern "C" void* qdb_gpu_rf_new(int64_t max_keys, int64_t bloom_bytes)
{
    clear_error();
    if (!ensure_init()) return nullptr;
    if (max_keys < 0 || bloom_bytes < 0) return fail("runtime filter: negative size"), nullptr;
    RuntimeFilterState* f = new RuntimeFilterState();
    f->max_keys = max_keys;
    int64_t b = 0;
    if (bloom_bytes >= kBloomBytesPerBlock) {
        b = 1;
        while (b * 2 <= bloom_bytes) b *= 2;
    }
    f->bloom_bytes = b;
    f->counters = (unsigned long long*)dalloc(64);
    if (!f->counters || !zero_fill(f->counters, 64)) {
        delete f;
        return nullptr;
    }
    return f;
}

extern "C" void qdb_gpu_rf_free(void* filter)
{
    delete (RuntimeFilterState*)filter;
}

extern "C" int qdb_gpu_rf_build(void* filter, const uint64_t* hashes, const uint8_t* selection, int64_t count)
{
    clear_error();
    RuntimeFilterState* f = (RuntimeFilterState*)filter;
    if (!f) return fail("null filter"), -1;
    if (f->kind >= 0) return fail("runtime filter: build after publish"), -1;
    if (count <= 0) return 0;
    if (!hashes) return fail("null hashes"), -1;
    // Room for every row of this batch (the reference's Keys invariant).
    const int64_t need = f->build_rows + count;
    if (need > f->keys_cap) {
        int64_t cap = f->keys_cap ? f->keys_cap : 4096;
        while (cap < need) cap *= 2;
        uint64_t* k = (uint64_t*)dalloc((size_t)cap * 8);
        if (!k) return -1;
        if (f->keys && f->build_rows > 0
            && !hip_ok(hipMemcpyAsync(k, f->keys, (size_t)f->build_rows * 8, hipMemcpyDeviceToDevice, stream()), "rf keys copy")) {
            dfree(k);
            return -1;
        }
        dfree(f->keys);
        f->keys = k;
        f->keys_cap = cap;
    }
    hipLaunchKernelGGL(k_rf_collect, dim3(grid_for(count, kBlock, 8)), dim3(kBlock), 0, stream(), hashes, selection, count,
        f->keys, f->counters);
    if (!hip_ok(hipGetLastError(), "rf collect")) return -1;
    f->build_rows += count;
    return 0;
}

extern "C" int qdb_gpu_rf_publish(void* filter)
{
    clear_error();
    RuntimeFilterState* f = (RuntimeFilterState*)filter;
    if (!f) return fail("null filter"), -1;
    if (f->kind >= 0) return f->kind;
    unsigned long long n = 0;
    if (!hip_ok(hipMemcpyAsync(&n, f->counters, 8, hipMemcpyDeviceToHost, stream()), "rf count") || !sync("rf publish")) return -1;
    // Distinct hashes through a deduplicating set of >= 2 (n + 1) slots.
    int64_t cap = 16;
    while (cap < 2 * ((int64_t)n + 1)) cap *= 2;
    unsigned long long* set = (unsigned long long*)dalloc((size_t)cap * 8);
    uint64_t* distinct_keys = (uint64_t*)dalloc((size_t)(n ? n : 1) * 8);
    bool ok = set && distinct_keys;
    if (ok) {
        hipLaunchKernelGGL(k_rf_fill, dim3(grid_for(cap, kBlock, 8)), dim3(kBlock), 0, stream(), set, cap, kSetEmpty);
        if (n) {
            hipLaunchKernelGGL(k_rf_dedup, dim3(grid_for((int64_t)n, kBlock, 8)), dim3(kBlock), 0, stream(), f->keys, (int64_t)n, set,
                (uint64_t)(cap - 1), distinct_keys, f->counters);
        }
        ok = hip_ok(hipGetLastError(), "rf dedup");
    }
    unsigned long long distinct = 0;
    ok = ok && hip_ok(hipMemcpyAsync(&distinct, f->counters + 1, 8, hipMemcpyDeviceToHost, stream()), "rf distinct")
        && sync("rf publish");
    if (ok && (int64_t)distinct <= f->max_keys) {
        // Exact: pow2 >= 2 * distinct slots, at least 8 (rf_publish).
        int64_t size = 8;
        while (size < 2 * (int64_t)distinct) size *= 2;
        f->words = (unsigned long long*)dalloc((size_t)size * 8);
        ok = f->words && zero_fill(f->words, (size_t)size * 8);
        if (ok && distinct) {
            hipLaunchKernelGGL(k_rf_exact_build, dim3(grid_for((int64_t)distinct, kBlock, 8)), dim3(kBlock), 0, stream(),
                distinct_keys, (int64_t)distinct, f->words, (uint64_t)(size - 1));
            ok = hip_ok(hipGetLastError(), "rf exact build") && sync("rf exact build");
        }
        f->word_count = size;
        f->exact_keys = (int64_t)distinct;
        f->kind = 1;
    } else if (ok && f->bloom_bytes == 0) {
        f->kind = 0;  // no Bloom storage: the filter cannot answer
    } else if (ok) {
        const int64_t words = f->bloom_bytes / 8;
        f->words = (unsigned long long*)dalloc((size_t)words * 8);
        ok = f->words && zero_fill(f->words, (size_t)words * 8);
        if (ok) {
            hipLaunchKernelGGL(k_rf_bloom_add, dim3(grid_for((int64_t)n, kBlock, 8)), dim3(kBlock), 0, stream(), f->keys, (int64_t)n,
                f->words, (uint64_t)(words / 8 - 1));
            hipLaunchKernelGGL(k_rf_popcount, dim3(grid_for(words, kBlock, 4)), dim3(kBlock), 0, stream(), f->words, words,
                f->counters + 2);
            unsigned long long bits = 0;
            ok = hip_ok(hipGetLastError(), "rf bloom")
                && hip_ok(hipMemcpyAsync(&bits, f->counters + 2, 8, hipMemcpyDeviceToHost, stream()), "rf bits") && sync("rf bloom");
            f->bloom_set_bits = (int64_t)bits;
        }
        f->word_count = words;
        f->kind = 2;
    }
    dfree(set);
    dfree(distinct_keys);
    if (!ok) {
        f->kind = -1;
        return -1;
    }
    // The collected hashes are no longer needed.
    dfree(f->keys);
    f->keys = nullptr;
    f->keys_cap = 0;
    return f->kind;
}

extern "C" int64_t qdb_gpu_rf_probe(const void* filter, const uint64_t* hashes, const uint8_t* selection, uint8_t* output,
    int64_t count)
{
    clear_error();
    const RuntimeFilterState* f = (const RuntimeFilterState*)filter;
    if (!f) return fail("null filter"), -1;
    if (f->kind < 0) return fail("runtime filter: probe before publish"), -1;
    if (count <= 0) return 0;
    if (!output || (f->kind > 0 && !hashes)) return fail("null argument"), -1;
    const int grid = grid_for(count, kBlock, 8);
    unsigned long long* part = (unsigned long long*)dalloc((size_t)grid * 8 + 8);
    if (!part) return -1;
    TimedRegion timer("rf_probe");
    const uint64_t mask = f->kind == 1 ? (uint64_t)(f->word_count - 1) : (uint64_t)(f->word_count / 8 - 1);
    switch (f->kind) {
    case 0: hipLaunchKernelGGL(k_rf_probe<0>, dim3(grid), dim3(kBlock), 0, stream(), hashes, selection, count, f->words, mask, output, part); break;
    case 1: hipLaunchKernelGGL(k_rf_probe<1>, dim3(grid), dim3(kBlock), 0, stream(), hashes, selection, count, f->words, mask, output, part); break;
    default: hipLaunchKernelGGL(k_rf_probe<2>, dim3(grid), dim3(kBlock), 0, stream(), hashes, selection, count, f->words, mask, output, part); break;
    }
    hipLaunchKernelGGL(k_rf_sum, dim3(1), dim3(kBlock), 0, stream(), part, grid, part + grid);
    unsigned long long kept = 0;
    const bool ok = hip_ok(hipGetLastError(), "rf probe")
        && hip_ok(hipMemcpyAsync(&kept, part + grid, 8, hipMemcpyDeviceToHost, stream()), "rf kept") && sync("rf probe");
    dfree(part);
    return ok ? (int64_t)kept : -1;
}

extern "C" int64_t qdb_gpu_rf_stats(const void* filter, int64_t* out, int64_t n)
{
    const RuntimeFilterState* f = (const RuntimeFilterState*)filter;
    if (!f) return -1;
    const int64_t v[5] = {f->kind, f->exact_keys, f->kind == 1 ? f->word_count : 0, f->kind == 2 ? f->word_count : 0,
                          f->bloom_set_bits};
    int64_t k = 0;
    for (; k < n && k < 5; ++k) out[k] = v[k];
    return k;
}
