// Equi hash join on gfx950: jr_dispatch / jr_materialize.
//
// Reference: the row join table (qdb/kernel/join/join_rows.oz,
// qdb/kernel/row_table.oz), generated append / probe / dispatch
// (qdb/kernel/join_gen.cpp:2017-2375), the processor protocol
// (qdb/exec/row_join_exec.cpp:110-276) and op-codes (qdb/kernel/compiler.h:
// 363-383).
//
// Device design.
//   Append   selected build rows with a valid key are compacted, in row
//            order, into a columnar arena in HBM (packed key words, the
//            build row ordinal, payload columns); a block count, a scan and
//            a scatter keep the order deterministic.
//   Build    an open-addressing table of 16-byte slots {key, head, count}
//            (capacity a power of two >= 2 x rows, like rt_init) links equal
//            keys into chains in build order (as jr_build_rows does). A
//            blocked Bloom filter over the same hash is built beside it.
//   Probe    one streaming pass over the probe batch. Each workgroup first
//            copies either the whole table (small builds) or the Bloom
//            filter into LDS; only rows that pass the filter touch the HBM
//            table. Matches are appended to the pair buffer with one atomic
//            per workgroup iteration, and the optional fused global
//            aggregate (COUNT, exact SUM, MIN/MAX) over the matched pairs is
//            reduced in registers and combined once per workgroup.
//   Materialize gathers probe and build columns for a range of pairs.
#include <hip/hip_runtime.h>

#include <string.h>

#include <mutex>
#include <unordered_map>
#include <vector>

#include "device.h"
#include "runtime.h"
#include "slots.h"

namespace qdbgpu {

namespace {

constexpr int kBlock = 256;
constexpr int kProbeBlock = 1024;
constexpr uint32_t kLockHead = 0xffffffffu;
constexpr int kMaxFused = QDB_JOIN_MAX_FUSED;

struct JKeyField {
    int32_t slot;
    int32_t kind;
    int32_t width;
    int32_t word;
    int32_t shift;
};

struct JKeyDesc {
    int nkeys;
    JKeyField f[QDB_JOIN_MAX_KEYS];
};

template <int KW>
struct Slot {
    uint64_t key[KW];
    uint32_t head;   // arena row + 1 of the first row of the chain; 0 = empty
    uint32_t count;  // rows in the chain
};

// Packed key of a row; false when a key field is NULL.
template <int KW>
__device__ __forceinline__ bool join_key(const DevCols& cols, const JKeyDesc& K, int64_t row, uint64_t (&key)[KW])
{
#pragma unroll
    for (int w = 0; w < KW; ++w) key[w] = 0;
    bool valid = true;
    for (int f = 0; f < K.nkeys; ++f) {
        const JKeyField kf = K.f[f];
        const DevCol& c = cols.c[kf.slot];
        if (!col_valid(c, row)) valid = false;
        uint64_t lo = 0, hi = 0;
        if (kf.kind == QDB_KIND_FLOAT) {
            lo = canonical_f64_bits(col_f64(c, row));
        } else if (kf.kind == QDB_KIND_BOOL) {
            lo = (uint64_t)bit_at((const uint8_t*)c.data, row, c.bit_off);
        } else if (kf.width == 16) {
            const uint64_t* p = (const uint64_t*)(c.data + 16 * row);
            lo = p[0];
            hi = p[1];
        } else {
            switch (kf.width) {
            case 1: lo = ((const uint8_t*)c.data)[row]; break;
            case 2: lo = ((const uint16_t*)c.data)[row]; break;
            case 4: lo = ((const uint32_t*)c.data)[row]; break;
            default: lo = ((const uint64_t*)c.data)[row]; break;
            }
        }
#pragma unroll
        for (int w = 0; w < KW; ++w) {
            if (w == kf.word) key[w] |= lo << kf.shift;
            if (kf.width == 16 && w == kf.word + 1) key[w] |= hi;
        }
    }
    return valid;
}

// ---------------------------------------------------------------------------
// Append: count, scan, scatter.
// ---------------------------------------------------------------------------

template <int KW>
__global__ void __launch_bounds__(kBlock) k_append_count(DevCols cols, JKeyDesc K, int64_t n, const uint8_t* sel,
    unsigned int* block_counts)
{
    __shared__ unsigned int total;
    if (threadIdx.x == 0) total = 0;
    __syncthreads();
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    bool keep = false;
    if (i < n && (!sel || sel[i])) {
        uint64_t key[KW];
        keep = join_key<KW>(cols, K, i, key);
    }
    const uint64_t b = __ballot(keep);
    if (lane_id() == 0 && b) atomicAdd(&total, (unsigned int)__popcll(b));
    __syncthreads();
    if (threadIdx.x == 0) block_counts[blockIdx.x] = total;
}

// Exclusive scan of `n` counts in place (one workgroup); writes the total.
__global__ void __launch_bounds__(1024) k_scan_counts(unsigned int* counts, unsigned int n, unsigned long long* total)
{
    __shared__ unsigned long long partial[1024];
    const unsigned int per = (n + blockDim.x - 1) / blockDim.x;
    const unsigned int lo = threadIdx.x * per;
    const unsigned int hi = lo + per < n ? lo + per : n;
    unsigned long long s = 0;
    for (unsigned int i = lo; i < hi; ++i) s += counts[i];
    partial[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long run = 0;
        for (unsigned int t = 0; t < blockDim.x; ++t) {
            unsigned long long v = partial[t];
            partial[t] = run;
            run += v;
        }
        *total = run;
    }
    __syncthreads();
    unsigned long long run = partial[threadIdx.x];
    for (unsigned int i = lo; i < hi; ++i) {
        unsigned int v = counts[i];
        counts[i] = (unsigned int)run;
        run += v;
    }
}

struct ArenaCols {
    char* data[64];
    uint8_t* valid[64];
    int32_t width[64];
    int32_t kind[64];
    int32_t src_slot[64];
};

// The arena payload is written for every build column the plan keeps.
template <int KW>
__global__ void __launch_bounds__(kBlock) k_append_write(DevCols cols, JKeyDesc K, int64_t n, const uint8_t* sel,
    const unsigned int* block_offsets, int64_t arena_base, int64_t ordinal_base, uint64_t* arena_keys,
    int64_t* arena_ordinal, int32_t* ord_to_arena, DevCols payload_cols, int npayload, ArenaCols AC)
{
    __shared__ unsigned int wave_base[kBlock / kWave];
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int lane = lane_id();
    const int wave = threadIdx.x / kWave;
    bool keep = false;
    uint64_t key[KW];
    if (i < n && (!sel || sel[i])) keep = join_key<KW>(cols, K, i, key);
    const uint64_t b = __ballot(keep);
    if (lane == 0) wave_base[wave] = (unsigned int)__popcll(b);
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned int run = 0;
        for (int w = 0; w < kBlock / kWave; ++w) {
            unsigned int v = wave_base[w];
            wave_base[w] = run;
            run += v;
        }
    }
    __syncthreads();
    if (i < n) ord_to_arena[ordinal_base + i] = -1;
    if (!keep) return;
    const unsigned int rank = wave_base[wave] + (unsigned int)__popcll(b & ((1ull << lane) - 1ull));
    const int64_t dst = arena_base + block_offsets[blockIdx.x] + rank;
#pragma unroll
    for (int w = 0; w < KW; ++w) arena_keys[dst * KW + w] = key[w];
    arena_ordinal[dst] = ordinal_base + i;
    ord_to_arena[ordinal_base + i] = (int32_t)dst;
    for (int p = 0; p < npayload; ++p) {
        const DevCol& c = payload_cols.c[AC.src_slot[p]];
        const bool v = col_valid(c, i);
        if (AC.valid[p]) AC.valid[p][dst] = v ? 1 : 0;
        char* d = AC.data[p];
        if (AC.kind[p] == QDB_KIND_BOOL) {
            ((uint8_t*)d)[dst] = v ? bit_at((const uint8_t*)c.data, i, c.bit_off) : 0;
            continue;
        }
        switch (AC.width[p]) {
        case 1: ((uint8_t*)d)[dst] = v ? ((const uint8_t*)c.data)[i] : 0; break;
        case 2: ((uint16_t*)d)[dst] = v ? ((const uint16_t*)c.data)[i] : 0; break;
        case 4: ((uint32_t*)d)[dst] = v ? ((const uint32_t*)c.data)[i] : 0; break;
        case 8: ((uint64_t*)d)[dst] = v ? ((const uint64_t*)c.data)[i] : 0; break;
        default: {
            const uint64_t* s = (const uint64_t*)(c.data + 16 * i);
            uint64_t* t = (uint64_t*)(d + 16 * dst);
            t[0] = v ? s[0] : 0;
            t[1] = v ? s[1] : 0;
        }
        }
    }
}

// ---------------------------------------------------------------------------
// Build: insert, order chains, Bloom filter.
// ---------------------------------------------------------------------------

__device__ __forceinline__ uint32_t bloom_word(uint64_t h, uint32_t words_mask)
{
    return (uint32_t)(h >> 20) & words_mask;
}

__device__ __forceinline__ uint32_t bloom_bits(uint64_t h)
{
    return (1u << ((h >> 52) & 31u)) | (1u << ((h >> 58) & 31u));
}

// One insert attempt of arena row `r`; false when the slot it needs is
// being filled by another lane. The caller retries in a loop that the whole
// wave runs together: a lane that spun here could wait forever on a lane of
// its own wave whose publishing store the compiler placed after the loop.
template <int KW>
__device__ __forceinline__ bool build_insert_once(Slot<KW>* table, uint64_t cmask, const uint64_t (&key)[KW], uint64_t h,
    int64_t r, uint32_t* next)
{
    uint64_t s = h & cmask;
    for (;;) {
        Slot<KW>* sl = &table[s];
        uint32_t head = __hip_atomic_load(&sl->head, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
        if (head == 0u) {
            head = atomicCAS(&sl->head, 0u, kLockHead);
            if (head == 0u) {
#pragma unroll
                for (int w = 0; w < KW; ++w) __hip_atomic_store(&sl->key[w], key[w], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                next[r] = 0u;
                __hip_atomic_store(&sl->head, (uint32_t)(r + 1), __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
                return true;
            }
        }
        if (head == kLockHead) return false;
        bool eq = true;
#pragma unroll
        for (int w = 0; w < KW; ++w) eq = eq && __hip_atomic_load(&sl->key[w], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == key[w];
        if (eq) {
            // Push on the chain; the order is fixed up afterwards. A
            // published head is never the lock value again.
            next[r] = atomicExch(&sl->head, (uint32_t)(r + 1));
            return true;
        }
        s = (s + 1) & cmask;
    }
}

template <int KW>
__global__ void __launch_bounds__(kBlock) k_build_insert(Slot<KW>* table, uint64_t cmask, const uint64_t* arena_keys,
    int64_t rows, uint32_t* next, uint32_t* bloom, uint32_t bloom_mask)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    // Every thread of the block runs the same number of iterations.
    for (int64_t b = blockIdx.x * (int64_t)blockDim.x; b < rows; b += stride) {
        const int64_t r = b + threadIdx.x;
        bool pending = r < rows;
        uint64_t key[KW];
#pragma unroll
        for (int w = 0; w < KW; ++w) key[w] = pending ? arena_keys[r * KW + w] : 0ull;
        const uint64_t h = key_place_hash<KW>(key);
        if (pending) atomicOr(&bloom[bloom_word(h, bloom_mask)], bloom_bits(h));
        while (__any(pending)) {
            if (pending && build_insert_once<KW>(table, cmask, key, h, r, next)) pending = false;
        }
    }
}

// Orders every chain by arena row (= build order) and counts it.
template <int KW>
__global__ void __launch_bounds__(kBlock) k_build_chains(Slot<KW>* table, uint64_t capacity, uint32_t* next,
    unsigned long long* keys_out)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    unsigned long long keys = 0;
    for (uint64_t s = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; s < capacity; s += stride) {
        uint32_t head = table[s].head;
        if (head == 0u) continue;
        ++keys;
        // Insertion sort of the singly linked list by row index.
        uint32_t sorted = 0u;
        uint32_t count = 0;
        uint32_t cur = head;
        while (cur) {
            uint32_t nxt = next[cur - 1];
            if (sorted == 0u || cur < sorted) {
                next[cur - 1] = sorted;
                sorted = cur;
            } else {
                uint32_t p = sorted;
                while (next[p - 1] && next[p - 1] < cur) p = next[p - 1];
                next[cur - 1] = next[p - 1];
                next[p - 1] = cur;
            }
            ++count;
            cur = nxt;
        }
        table[s].head = sorted;
        table[s].count = count;
    }
    if (keys) atomicAdd(keys_out, keys);
}

// ---------------------------------------------------------------------------
// Probe.
// ---------------------------------------------------------------------------

enum FusedKind : uint8_t { FK_COUNT_PAIRS, FK_COUNT_VALID, FK_SUM, FK_MIN, FK_MAX };

struct FusedDesc {
    int32_t kind;
    int32_t from_build;   // argument from the build arena
    int32_t slot;         // probe column slot
    int32_t payload;      // build payload index
};

struct ProbeArgs {
    DevCols cols;         // probe columns used (keys, fused args, selection none)
    JKeyDesc K;
    int64_t n;
    const uint8_t* sel;
    const void* table;
    uint64_t cmask;
    const uint32_t* next;
    const int64_t* ordinal;
    const uint32_t* bloom;
    uint32_t bloom_words;       // power of two, or 0 (no filter)
    uint32_t lds_table_slots;   // > 0: the whole table is copied to LDS
    int emit_unmatched;
    int probe_is_left;
    int64_t* pairs;             // 2 i64 per pair
    unsigned long long* pair_count;
    unsigned long long pair_cap;
    int nfused;
    FusedDesc fused[kMaxFused];
    const char* payload_data[kMaxFused];
    const uint8_t* payload_valid[kMaxFused];
    int32_t payload_width[kMaxFused];
    unsigned long long* fused_acc;  // per fused agg: [count, p0, p1, p2, min, max] words
};

constexpr int kAccWords = 6;

// Probe of one valid key: the chain head and its length, or 0.
template <int KW, bool kLdsTable>
__device__ __forceinline__ uint32_t probe_key(const ProbeArgs& P, const Slot<KW>* lds_table, const uint64_t (&key)[KW],
    uint64_t h, uint32_t* count)
{
    *count = 0;
    if (kLdsTable) {
        const uint32_t mask = P.lds_table_slots - 1u;
        uint32_t s = (uint32_t)h & mask;
        for (;;) {
            const Slot<KW>& sl = lds_table[s];
            if (sl.head == 0u) return 0u;
            bool eq = true;
#pragma unroll
            for (int w = 0; w < KW; ++w) eq = eq && sl.key[w] == key[w];
            if (eq) {
                *count = sl.count;
                return sl.head;
            }
            s = (s + 1u) & mask;
        }
    }
    const Slot<KW>* table = (const Slot<KW>*)P.table;
    uint64_t s = h & P.cmask;
    for (;;) {
        const Slot<KW> sl = table[s];
        if (sl.head == 0u) return 0u;
        bool eq = true;
#pragma unroll
        for (int w = 0; w < KW; ++w) eq = eq && sl.key[w] == key[w];
        if (eq) {
            *count = sl.count;
            return sl.head;
        }
        s = (s + 1) & P.cmask;
    }
}

struct FusedAcc {
    unsigned long long count;
    u128 sum;
    int64_t mn, mx;
};

template <int KW, bool kLdsTable>
__global__ void __launch_bounds__(kProbeBlock) k_probe(ProbeArgs P)
{
    extern __shared__ uint32_t lds[];
    __shared__ unsigned int wave_tot[kProbeBlock / kWave];
    __shared__ unsigned long long block_base;
    const Slot<KW>* lds_table = (const Slot<KW>*)lds;
    const uint32_t* lds_bloom = lds;
    if (kLdsTable) {
        const uint32_t* src = (const uint32_t*)P.table;
        const uint32_t words = P.lds_table_slots * (uint32_t)(sizeof(Slot<KW>) / 4);
        for (uint32_t i = threadIdx.x; i < words; i += blockDim.x) lds[i] = src[i];
    } else if (P.bloom_words) {
        for (uint32_t i = threadIdx.x; i < P.bloom_words; i += blockDim.x) lds[i] = P.bloom[i];
    }
    __syncthreads();
    const int lane = lane_id();
    const int wave = threadIdx.x / kWave;
    FusedAcc acc[kMaxFused];
#pragma unroll
    for (int f = 0; f < kMaxFused; ++f) {
        acc[f].count = 0;
        acc[f].sum = 0;
        acc[f].mn = 0x7fffffffffffffffll;
        acc[f].mx = (int64_t)0x8000000000000000ll;
    }
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    // Every thread of the block runs the same number of iterations (barriers).
    for (int64_t base = blockIdx.x * (int64_t)blockDim.x; base < P.n; base += stride) {
        const int64_t i = base + threadIdx.x;
        uint32_t head = 0, count = 0;
        bool unmatched = false;
        if (i < P.n && (!P.sel || P.sel[i])) {
            uint64_t key[KW];
            const bool valid = join_key<KW>(P.cols, P.K, i, key);
            if (valid) {
                const uint64_t h = key_place_hash<KW>(key);
                bool maybe = true;
                if (!kLdsTable && P.bloom_words) {
                    const uint32_t bits = bloom_bits(h);
                    maybe = (lds_bloom[bloom_word(h, P.bloom_words - 1u)] & bits) == bits;
                }
                if (maybe) head = probe_key<KW, kLdsTable>(P, lds_table, key, h, &count);
            }
            unmatched = head == 0u && P.emit_unmatched;
        }
        const unsigned int emit = count + (unmatched ? 1u : 0u);
        // Block-wide exclusive scan of the pair counts: one atomic per block.
        unsigned int incl = emit;
#pragma unroll
        for (int off = 1; off < kWave; off <<= 1) {
            unsigned int y = (unsigned int)__shfl_up((int)incl, off, kWave);
            if (lane >= off) incl += y;
        }
        if (lane == kWave - 1) wave_tot[wave] = incl;
        __syncthreads();
        if (threadIdx.x == 0) {
            unsigned long long run = 0;
            for (int w = 0; w < kProbeBlock / kWave; ++w) {
                unsigned int v = wave_tot[w];
                wave_tot[w] = (unsigned int)run;
                run += v;
            }
            block_base = run ? atomicAdd(P.pair_count, run) : 0ull;
        }
        __syncthreads();
        unsigned long long pos = block_base + wave_tot[wave] + (incl - emit);
        const int64_t stream_id = (int64_t)((uint64_t)-1 << 32) + i;
        if (emit) {
            if (unmatched) {
                if (pos < P.pair_cap) {
                    P.pairs[2 * pos] = P.probe_is_left ? stream_id : -1;
                    P.pairs[2 * pos + 1] = P.probe_is_left ? -1 : stream_id;
                }
            }
            for (uint32_t r = head; r; r = P.next[r - 1]) {
                const int64_t ord = P.ordinal[r - 1];
                if (pos < P.pair_cap) {
                    P.pairs[2 * pos] = P.probe_is_left ? stream_id : ord;
                    P.pairs[2 * pos + 1] = P.probe_is_left ? ord : stream_id;
                }
                ++pos;
                // Fused global aggregate over this pair.
#pragma unroll
                for (int f = 0; f < kMaxFused; ++f) {
                    if (f >= P.nfused) continue;
                    const FusedDesc d = P.fused[f];
                    if (d.kind == FK_COUNT_PAIRS) {
                        acc[f].count += 1;
                        continue;
                    }
                    bool v;
                    i128 x;
                    if (d.from_build) {
                        const uint8_t* vv = P.payload_valid[f];
                        v = !vv || vv[r - 1];
                        const char* pd = P.payload_data[f];
                        switch (P.payload_width[f]) {
                        case 1: x = ((const int8_t*)pd)[r - 1]; break;
                        case 2: x = ((const int16_t*)pd)[r - 1]; break;
                        case 4: x = ((const int32_t*)pd)[r - 1]; break;
                        case 8: x = ((const int64_t*)pd)[r - 1]; break;
                        default: {
                            const uint64_t* p = (const uint64_t*)(pd + 16 * (size_t)(r - 1));
                            x = (i128)(((u128)p[1] << 64) | (u128)p[0]);
                        }
                        }
                    } else {
                        const DevCol& c = P.cols.c[d.slot];
                        v = col_valid(c, i);
                        x = v ? col_int(c, i) : 0;
                    }
                    if (!v) continue;
                    acc[f].count += 1;
                    if (d.kind == FK_SUM) acc[f].sum += (u128)x;
                    else if (d.kind == FK_MIN) acc[f].mn = (int64_t)x < acc[f].mn ? (int64_t)x : acc[f].mn;
                    else if (d.kind == FK_MAX) acc[f].mx = (int64_t)x > acc[f].mx ? (int64_t)x : acc[f].mx;
                }
            }
        }
        __syncthreads();
    }
    // Combine the fused accumulators: wave reduction, then one atomic per
    // wave and word (add words hold the sum as bits [0,32), [32,64), [64,128)).
    for (int f = 0; f < P.nfused; ++f) {
        unsigned long long w[kAccWords];
        w[0] = acc[f].count;
        w[1] = (unsigned long long)((uint64_t)acc[f].sum & 0xffffffffull);
        w[2] = (unsigned long long)((uint64_t)acc[f].sum >> 32);
        w[3] = (unsigned long long)(uint64_t)(acc[f].sum >> 64);
        w[4] = (unsigned long long)acc[f].mn;
        w[5] = (unsigned long long)acc[f].mx;
#pragma unroll
        for (int off = 1; off < kWave; off <<= 1) {
            w[0] += shfl_xor_u64(w[0], off);
            w[1] += shfl_xor_u64(w[1], off);
            w[2] += shfl_xor_u64(w[2], off);
            w[3] += shfl_xor_u64(w[3], off);
            int64_t a = (int64_t)shfl_xor_u64(w[4], off);
            int64_t b = (int64_t)shfl_xor_u64(w[5], off);
            w[4] = (unsigned long long)((int64_t)w[4] < a ? (int64_t)w[4] : a);
            w[5] = (unsigned long long)((int64_t)w[5] > b ? (int64_t)w[5] : b);
        }
        if (lane == 0) {
            unsigned long long* dst = P.fused_acc + (size_t)f * kAccWords;
            if (w[0]) {
                atomicAdd(&dst[0], w[0]);
                atomicAdd(&dst[1], w[1]);
                atomicAdd(&dst[2], w[2]);
                atomicAdd(&dst[3], w[3]);
                atomicMin((long long*)&dst[4], (long long)w[4]);
                atomicMax((long long*)&dst[5], (long long)w[5]);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Materialize.
// ---------------------------------------------------------------------------

struct GatherCol {
    int32_t from_build;
    int32_t kind;
    int32_t width;
    int32_t nullable_out;
    int32_t slot;      // probe column slot
    int32_t payload;   // build payload index
    const char* arena_data;
    const uint8_t* arena_valid;
    char* out_data;
    uint8_t* out_mask;
};

struct GatherArgs {
    DevCols probe;
    const int64_t* pairs;
    int64_t start;
    int64_t n;
    int probe_is_left;
    const int32_t* ord_to_arena;
    int ncols;
    GatherCol col[40];
};

constexpr int kGatherMax = 40;

__global__ void __launch_bounds__(kBlock) k_materialize(GatherArgs G)
{
    const int64_t nbytes = (G.n + 7) / 8;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t b = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; b < nbytes; b += stride) {
        for (int c = 0; c < G.ncols; ++c) {
            const GatherCol gc = G.col[c];
            uint8_t mbits = 0, dbits = 0;
            for (int64_t o = b * 8; o < b * 8 + 8 && o < G.n; ++o) {
                const int64_t* pr = G.pairs + 2 * (G.start + o);
                const int64_t stream_id = G.probe_is_left ? pr[0] : pr[1];
                const int64_t build = G.probe_is_left ? pr[1] : pr[0];
                bool valid;
                u128 v = 0;
                if (gc.from_build) {
                    const int64_t r = build < 0 ? -1 : (int64_t)G.ord_to_arena[build];
                    valid = r >= 0 && (!gc.arena_valid || gc.arena_valid[r]);
                    if (valid) {
                        switch (gc.kind == QDB_KIND_BOOL ? 1 : gc.width) {
                        case 1: v = ((const uint8_t*)gc.arena_data)[r]; break;
                        case 2: v = ((const uint16_t*)gc.arena_data)[r]; break;
                        case 4: v = ((const uint32_t*)gc.arena_data)[r]; break;
                        case 8: v = ((const uint64_t*)gc.arena_data)[r]; break;
                        default: {
                            const uint64_t* p = (const uint64_t*)(gc.arena_data + 16 * r);
                            v = ((u128)p[1] << 64) | (u128)p[0];
                        }
                        }
                    }
                } else {
                    const int64_t row = stream_id - (int64_t)((uint64_t)-1 << 32);
                    const DevCol& pc = G.probe.c[gc.slot];
                    valid = col_valid(pc, row);
                    if (valid) {
                        if (gc.kind == QDB_KIND_BOOL) v = (u128)bit_at((const uint8_t*)pc.data, row, pc.bit_off);
                        else if (gc.kind == QDB_KIND_FLOAT && gc.width == 8) v = (u128)(uint64_t)__double_as_longlong(((const double*)pc.data)[row]);
                        else if (gc.width == 16) {
                            const uint64_t* p = (const uint64_t*)(pc.data + 16 * row);
                            v = ((u128)p[1] << 64) | (u128)p[0];
                        } else {
                            switch (gc.width) {
                            case 1: v = ((const uint8_t*)pc.data)[row]; break;
                            case 2: v = ((const uint16_t*)pc.data)[row]; break;
                            case 4: v = ((const uint32_t*)pc.data)[row]; break;
                            default: v = ((const uint64_t*)pc.data)[row]; break;
                            }
                        }
                    }
                }
                const int bit = (int)(o & 7);
                if (valid) mbits |= (1u << bit);
                if (gc.kind == QDB_KIND_BOOL) {
                    if (valid && v) dbits |= (1u << bit);
                    continue;
                }
                char* d = gc.out_data;
                switch (gc.width) {
                case 1: ((uint8_t*)d)[o] = v; break;
                case 2: ((uint16_t*)d)[o] = (uint16_t)v; break;
                case 4: ((uint32_t*)d)[o] = (uint32_t)v; break;
                case 8: ((uint64_t*)d)[o] = (uint64_t)v; break;
                default: {
                    uint64_t* p = (uint64_t*)(d + 16 * o);
                    p[0] = (uint64_t)v;
                    p[1] = (uint64_t)(v >> 64);
                }
                }
            }
            if (gc.kind == QDB_KIND_BOOL) ((uint8_t*)gc.out_data)[b] = dbits;
            if (gc.out_mask) gc.out_mask[b] = mbits;
        }
    }
}

// ---------------------------------------------------------------------------
// Host.
// ---------------------------------------------------------------------------

struct JoinPlan {
    QdbRowJoinSpec spec;
    int KW = 1;
    JKeyDesc build_key;   // slots index build-side DevCols
    JKeyDesc probe_key;   // slots index probe-side DevCols
    int build_nslots = 0, build_slot_col[kMaxDevCols];
    int probe_nslots = 0, probe_slot_col[kMaxDevCols];
    int build_key_slots = 0;  // build slots used by keys (payload slots follow)
    int npayload = 0;         // every build column is kept in the arena
    FusedDesc fused[kMaxFused];
    int nout = 0;
    QdbValueType out_types[128];
    GatherCol gather_proto[128];
};

// The device state behind the RowArena blob.
struct JoinBuild {
    JoinPlan* plan = nullptr;
    int64_t rows = 0, cap = 0;
    uint64_t* keys = nullptr;
    int64_t* ordinal = nullptr;
    std::vector<char*> pdata;
    std::vector<uint8_t*> pvalid;
    int64_t appended = 0, ord_cap = 0;
    int32_t* ord_to_arena = nullptr;
    // table
    void* table = nullptr;
    uint64_t capacity = 0;
    uint32_t* next = nullptr;
    uint32_t* bloom = nullptr;
    uint32_t bloom_words = 0;
    int64_t keys_distinct = 0;
    bool built = false;
    // fused aggregate words, device
    unsigned long long* fused_acc = nullptr;
    unsigned long long* pair_count = nullptr;   // device counter
    // the probe batch the pair buffer refers to (for materialize)
    TRowSet last_probe;
};

struct RowArenaBlob {
    JoinBuild* build;
    int64_t BlockCount;
    int64_t BlockCapacity;
    int64_t BlockRows;
    int64_t RowBytes;
    int64_t Rows;
    int64_t CurRows;
};
static_assert(sizeof(RowArenaBlob) == QDB_ROW_ARENA_SIZE, "RowArena blob size");

struct RowTableBlob {
    void* entries;
    int64_t Capacity;
    uint64_t Mask;
    int64_t Shift;
    int64_t Rows;
    int64_t Keys;
};
static_assert(sizeof(RowTableBlob) == QDB_ROW_TABLE_SIZE, "RowTable blob size");

static_assert(sizeof(QdbPairBuffer) == QDB_PAIR_BUFFER_SIZE, "PairBuffer blob size");

SlotTable<JoinPlan>& join_slots()
{
    static SlotTable<JoinPlan>* t = new SlotTable<JoinPlan>();
    return *t;
}

// Pair buffer blob -> the join that filled it (materialize has no arena).
std::mutex g_pairs_mu;
std::unordered_map<const void*, JoinBuild*>& pair_owner()
{
    static auto* m = new std::unordered_map<const void*, JoinBuild*>();
    return *m;
}

int key_bytes(const QdbValueType& t)
{
    if (t.Kind == QDB_KIND_BOOL) return 1;
    if (t.Kind == QDB_KIND_FLOAT) return 8;
    return t.Width;
}

bool compile_join(const QdbRowJoinSpec& spec, JoinPlan* P)
{
    P->spec = spec;
    if (spec.NumKeys < 1 || spec.NumKeys > QDB_JOIN_MAX_KEYS) return fail("bad join key count %d", spec.NumKeys);
    if (spec.JoinType < 0 || spec.JoinType > QDB_JOIN_RIGHT_OUTER) return fail("unsupported join type %d", spec.JoinType);
    if (spec.Build.NumColumns > 64 || spec.Probe.NumColumns > 64) return fail("too many columns");
    if (spec.NumFused < 0 || spec.NumFused > kMaxFused) return fail("bad fused aggregate count");
    // Keys: build and probe pack identically.
    int byte = 0;
    for (int k = 0; k < spec.NumKeys; ++k) {
        const int bc = spec.BuildKeyColumns[k], pc = spec.ProbeKeyColumns[k];
        if (bc < 0 || bc >= spec.Build.NumColumns || pc < 0 || pc >= spec.Probe.NumColumns) return fail("join key column out of range");
        const QdbValueType& bt = spec.Build.Types[bc];
        const QdbValueType& pt = spec.Probe.Types[pc];
        if (bt.Kind == QDB_KIND_STRING || pt.Kind == QDB_KIND_STRING) return fail("string join keys are not supported on this device");
        const int w = key_bytes(bt);
        if (w != key_bytes(pt) || (bt.Kind == QDB_KIND_FLOAT) != (pt.Kind == QDB_KIND_FLOAT)) {
            return fail("join key %d: build and probe types differ", k);
        }
        if (bt.Kind == QDB_KIND_FLOAT && bt.Width != 8) return fail("float join keys must be f64");
        const int align = w > 8 ? 8 : w;
        byte = (byte + align - 1) / align * align;
        JKeyField f;
        f.kind = bt.Kind;
        f.width = w;
        f.word = (byte / 8);
        f.shift = ((byte % 8) * 8);
        f.slot = P->build_nslots;
        P->build_key.f[k] = f;
        P->build_slot_col[P->build_nslots++] = bc;
        f.kind = pt.Kind;
        f.slot = P->probe_nslots;
        P->probe_key.f[k] = f;
        P->probe_slot_col[P->probe_nslots++] = pc;
        byte += w;
    }
    P->build_key.nkeys = P->probe_key.nkeys = spec.NumKeys;
    P->KW = byte <= 8 ? 1 : byte <= 16 ? 2 : 0;
    if (P->KW == 0) return fail("join key of %d bytes exceeds 16", byte);
    P->build_key_slots = P->build_nslots;
    // Payload: every build column.
    P->npayload = spec.Build.NumColumns;
    for (int c = 0; c < spec.Build.NumColumns; ++c) {
        if (spec.Build.Types[c].Kind == QDB_KIND_STRING) return fail("string build columns are not supported on this device");
    }
    // Fused aggregates.
    for (int f = 0; f < spec.NumFused; ++f) {
        const QdbAggregateFn& fn = spec.Fused[f];
        FusedDesc& d = P->fused[f];
        memset(&d, 0, sizeof(d));
        if (fn.Func == QDB_AGG_COUNT_STAR) {
            d.kind = FK_COUNT_PAIRS;
            continue;
        }
        const bool from_build = fn.Column <= -2;
        const int col = from_build ? -fn.Column - 2 : fn.Column;
        const QdbColumnSchema& sch = from_build ? spec.Build : spec.Probe;
        if (col < 0 || col >= sch.NumColumns) return fail("fused aggregate %d: bad column", f);
        const QdbValueType& t = sch.Types[col];
        if (t.Kind == QDB_KIND_FLOAT || t.Kind == QDB_KIND_STRING || t.Kind == QDB_KIND_BOOL) {
            return fail("fused aggregate %d: only integer, date and decimal arguments", f);
        }
        if ((fn.Func == QDB_AGG_MIN || fn.Func == QDB_AGG_MAX) && t.Width == 16) {
            return fail("fused aggregate %d: MIN/MAX of 16-byte decimals", f);
        }
        d.kind = fn.Func == QDB_AGG_COUNT ? FK_COUNT_VALID : fn.Func == QDB_AGG_SUM ? FK_SUM : fn.Func == QDB_AGG_MIN ? FK_MIN : FK_MAX;
        if (fn.Func == QDB_AGG_SUM && t.Kind == QDB_KIND_DATE) return fail("function sum(DATE) does not exist");
        d.from_build = from_build;
        if (from_build) {
            d.payload = col;
        } else {
            int s = -1;
            for (int q = 0; q < P->probe_nslots; ++q) {
                if (P->probe_slot_col[q] == col) s = q;
            }
            if (s < 0) {
                if (P->probe_nslots >= kMaxDevCols) return fail("too many probe columns");
                s = P->probe_nslots;
                P->probe_slot_col[P->probe_nslots++] = col;
            }
            d.slot = s;
        }
    }
    // Output columns (left side first).
    const bool outer = spec.JoinType != QDB_JOIN_INNER;
    auto add_side = [&](bool build_side) -> bool {
        const QdbColumnSchema& sch = build_side ? spec.Build : spec.Probe;
        for (int c = 0; c < sch.NumColumns; ++c) {
            if (P->nout >= 128) return fail("too many output columns");
            QdbValueType t = sch.Types[c];
            if (build_side && outer) t.Nullable = true;
            P->out_types[P->nout] = t;
            GatherCol& g = P->gather_proto[P->nout];
            memset(&g, 0, sizeof(g));
            g.from_build = build_side;
            g.kind = t.Kind;
            g.width = (t.Kind == QDB_KIND_BOOL ? 1 : t.Width);
            g.nullable_out = t.Nullable;
            g.payload = c;
            g.slot = -1;
            ++P->nout;
        }
        return true;
    };
    if (spec.ProbeIsLeft) {
        if (!add_side(false) || !add_side(true)) return false;
    } else {
        if (!add_side(true) || !add_side(false)) return false;
    }
    return true;
}

bool bind_side(const int* slot_col, int nslots, const QdbColumnSchema& sch, const TRowSet* rs, DevCols* cols)
{
    memset(cols, 0, sizeof(*cols));
    for (int s = 0; s < nslots; ++s) {
        const int c = slot_col[s];
        if (c >= rs->ColumnCount) return fail("batch has %lld columns, the join reads column %d", (long long)rs->ColumnCount, c);
        const TColumn& col = rs->Columns[c];
        const QdbValueType& t = sch.Types[c];
        DevCol& d = cols->c[s];
        d.data = col.Data;
        d.mask = col.Mask;
        d.mask_off = col.MaskBitOffset;
        d.bit_off = col.DataBitOffset;
        d.kind = t.Kind;
        d.width = t.Width;
        d.nullable = t.Nullable;
    }
    return true;
}

void build_free(JoinBuild* J)
{
    if (!J) return;
    dfree(J->keys);
    dfree(J->ordinal);
    for (char* p : J->pdata) dfree(p);
    for (uint8_t* p : J->pvalid) dfree(p);
    dfree(J->ord_to_arena);
    dfree(J->table);
    dfree(J->next);
    dfree(J->bloom);
    dfree(J->fused_acc);
    dfree(J->pair_count);
    delete J;
}

template <typename T>
bool grow_buf(T** p, int64_t old_elems, int64_t new_elems, size_t elem_bytes)
{
    T* q = (T*)dalloc((size_t)new_elems * elem_bytes);
    if (!q) return false;
    if (*p && old_elems > 0) {
        if (!hip_ok(hipMemcpyAsync(q, *p, (size_t)old_elems * elem_bytes, hipMemcpyDeviceToDevice, stream()), "arena grow")) return false;
        if (!sync("arena grow")) return false;
    }
    dfree(*p);
    *p = q;
    return true;
}

bool arena_reserve(JoinBuild* J, int64_t more_rows, int64_t more_ord)
{
    JoinPlan* P = J->plan;
    if (J->rows + more_rows > J->cap) {
        int64_t cap = J->cap ? J->cap : 1024;
        while (cap < J->rows + more_rows) cap *= 2;
        if (!grow_buf(&J->keys, J->rows * P->KW, cap * P->KW, 8)) return false;
        if (!grow_buf(&J->ordinal, J->rows, cap, 8)) return false;
        for (int p = 0; p < P->npayload; ++p) {
            const QdbValueType& t = P->spec.Build.Types[p];
            const size_t w = t.Kind == QDB_KIND_BOOL ? 1 : (size_t)t.Width;
            if (!grow_buf(&J->pdata[(size_t)p], J->rows, cap, w)) return false;
            if (t.Nullable && !grow_buf(&J->pvalid[(size_t)p], J->rows, cap, 1)) return false;
        }
        J->cap = cap;
    }
    if (J->appended + more_ord > J->ord_cap) {
        int64_t cap = J->ord_cap ? J->ord_cap : 1024;
        while (cap < J->appended + more_ord) cap *= 2;
        if (!grow_buf(&J->ord_to_arena, J->appended, cap, 4)) return false;
        J->ord_cap = cap;
    }
    return true;
}

template <int KW>
bool do_append(JoinBuild* J, const TRowSet* batch)
{
    JoinPlan* P = J->plan;
    const int64_t n = batch->RowCount;
    if (n <= 0) return true;
    if (n >= (int64_t)1 << 31) return fail("append batches are limited to 2^31 rows");
    if (!arena_reserve(J, n, n)) return false;
    DevCols cols;
    if (!bind_side(P->build_slot_col, P->build_nslots, P->spec.Build, batch, &cols)) return false;
    // Payload: every build column of the batch.
    DevCols pcols;
    ArenaCols AC;
    memset(&AC, 0, sizeof(AC));
    memset(&pcols, 0, sizeof(pcols));
    // Payload columns go through a second DevCols in chunks of kMaxDevCols.
    const int nblocks = (int)((n + kBlock - 1) / kBlock);
    unsigned int* counts = (unsigned int*)dalloc((size_t)nblocks * 4 + 16);
    unsigned long long* total = (unsigned long long*)dalloc(16);
    if (!counts || !total) return false;
    hipLaunchKernelGGL((k_append_count<KW>), dim3(nblocks), dim3(kBlock), 0, stream(), cols, P->build_key, n, batch->Selection, counts);
    hipLaunchKernelGGL(k_scan_counts, dim3(1), dim3(1024), 0, stream(), counts, (unsigned int)nblocks, total);
    int64_t added = 0;
    bool ok = hip_ok(hipGetLastError(), "append count");
    for (int p0 = 0; ok && (p0 < P->npayload || p0 == 0); p0 += kMaxDevCols) {
        const int np = P->npayload - p0 < kMaxDevCols ? P->npayload - p0 : kMaxDevCols;
        int slot_col[kMaxDevCols];
        for (int q = 0; q < np; ++q) {
            slot_col[q] = p0 + q;
            const QdbValueType& t = P->spec.Build.Types[p0 + q];
            AC.data[q] = J->pdata[(size_t)(p0 + q)];
            AC.valid[q] = t.Nullable ? J->pvalid[(size_t)(p0 + q)] : nullptr;
            AC.width[q] = (t.Kind == QDB_KIND_BOOL ? 1 : t.Width);
            AC.kind[q] = t.Kind;
            AC.src_slot[q] = q;
        }
        if (!bind_side(slot_col, np, P->spec.Build, batch, &pcols)) return false;
        hipLaunchKernelGGL((k_append_write<KW>), dim3(nblocks), dim3(kBlock), 0, stream(), cols, P->build_key, n,
            batch->Selection, counts, J->rows, J->appended, J->keys, J->ordinal, J->ord_to_arena, pcols, np, AC);
        ok = hip_ok(hipGetLastError(), "append write");
        if (P->npayload == 0) break;
    }
    unsigned long long h_total = 0;
    if (ok) ok = hip_ok(hipMemcpyAsync(&h_total, total, 8, hipMemcpyDeviceToHost, stream()), "append total") && sync("append");
    dfree(counts);
    dfree(total);
    if (!ok) return false;
    added = (int64_t)h_total;
    J->rows += added;
    J->appended += n;
    return true;
}

template <int KW>
bool do_build(JoinBuild* J)
{
    uint64_t capacity = 1024;
    while (capacity < (uint64_t)J->rows * 2) capacity *= 2;
    J->capacity = capacity;
    J->table = dalloc(capacity * sizeof(Slot<KW>));
    J->next = (uint32_t*)dalloc((size_t)(J->rows > 0 ? J->rows : 1) * 4);
    // Bloom filter: about 16 bits per row, between 8 KiB and 128 KiB.
    uint32_t words = 2048;
    while (words < 32768u && (uint64_t)words * 32 < (uint64_t)J->rows * 16) words *= 2;
    J->bloom_words = words;
    J->bloom = (uint32_t*)dalloc((size_t)words * 4);
    unsigned long long* keys = (unsigned long long*)dalloc(16);
    if (!J->table || !J->next || !J->bloom || !keys) return false;
    if (!zero_fill(J->table, capacity * sizeof(Slot<KW>)) || !zero_fill(J->bloom, (size_t)words * 4) || !zero_fill(keys, 16)) {
        return false;
    }
    if (J->rows > 0) {
        TimedRegion timer("join_build");
        hipLaunchKernelGGL((k_build_insert<KW>), dim3(grid_for(J->rows, kBlock, 8)), dim3(kBlock), 0, stream(),
            (Slot<KW>*)J->table, capacity - 1, J->keys, J->rows, J->next, J->bloom, words - 1u);
        hipLaunchKernelGGL((k_build_chains<KW>), dim3(grid_for((int64_t)capacity, kBlock, 8)), dim3(kBlock), 0, stream(),
            (Slot<KW>*)J->table, capacity, J->next, keys);
    }
    unsigned long long h_keys = 0;
    bool ok = hip_ok(hipGetLastError(), "build launch")
        && hip_ok(hipMemcpyAsync(&h_keys, keys, 8, hipMemcpyDeviceToHost, stream()), "build keys") && sync("build");
    dfree(keys);
    if (!ok) return false;
    J->keys_distinct = (int64_t)h_keys;
    J->built = true;
    return true;
}

template <int KW>
bool do_probe(JoinBuild* J, const TRowSet* batch, QdbPairBuffer* pairs)
{
    JoinPlan* P = J->plan;
    const int64_t n = batch->RowCount;
    pairs->Count = 0;
    J->last_probe = *batch;
    if (n <= 0) return true;
    if (n >= (int64_t)1 << 32) return fail("probe batches are limited to 2^32 rows");
    ProbeArgs A;
    memset(&A, 0, sizeof(A));
    if (!bind_side(P->probe_slot_col, P->probe_nslots, P->spec.Probe, batch, &A.cols)) return false;
    A.K = P->probe_key;
    A.n = n;
    A.sel = batch->Selection;
    A.table = J->table;
    A.cmask = J->capacity - 1;
    A.next = J->next;
    A.ordinal = J->ordinal;
    A.bloom = J->bloom;
    A.bloom_words = J->bloom_words;
    A.emit_unmatched = P->spec.JoinType != QDB_JOIN_INNER;
    A.probe_is_left = P->spec.ProbeIsLeft;
    A.nfused = P->spec.NumFused;
    for (int f = 0; f < A.nfused; ++f) {
        A.fused[f] = P->fused[f];
        if (P->fused[f].from_build) {
            const int col = P->fused[f].payload;
            const QdbValueType& t = P->spec.Build.Types[col];
            A.payload_data[f] = J->pdata[(size_t)col];
            A.payload_valid[f] = t.Nullable ? J->pvalid[(size_t)col] : nullptr;
            A.payload_width[f] = t.Width;
        }
    }
    A.fused_acc = J->fused_acc;
    // Whole table in LDS when it fits in 64 KiB; else the Bloom filter.
    const size_t table_bytes = J->capacity * sizeof(Slot<KW>);
    const bool lds_table = table_bytes <= 64 * 1024;
    A.lds_table_slots = lds_table ? (uint32_t)J->capacity : 0u;
    const size_t lds = lds_table ? table_bytes : (size_t)J->bloom_words * 4;
    // Pair capacity: the sum of the chain lengths bounds the matches; with
    // distinct build keys every probe row matches at most once.
    unsigned long long cap = (unsigned long long)n;
    if (J->keys_distinct != J->rows) cap = (unsigned long long)n * (unsigned long long)(J->rows - J->keys_distinct + 1);
    if (A.emit_unmatched) cap += (unsigned long long)n;
    for (int attempt = 0; attempt < 2; ++attempt) {
        if ((unsigned long long)pairs->Capacity < cap) {
            dfree(pairs->Data);
            pairs->Data = (int64_t*)dalloc((size_t)cap * 16);
            if (!pairs->Data) return false;
            pairs->Capacity = (int64_t)cap;
        }
        A.pairs = pairs->Data;
        A.pair_cap = (unsigned long long)pairs->Capacity;
        A.pair_count = J->pair_count;
        if (!zero_fill(J->pair_count, 8)) return false;
        // Fused accumulators are per probe call; they are summed on the host.
        if (A.nfused) {
            std::vector<unsigned long long> init((size_t)A.nfused * kAccWords, 0ull);
            for (int f = 0; f < A.nfused; ++f) {
                init[(size_t)f * kAccWords + 4] = 0x7fffffffffffffffull;
                init[(size_t)f * kAccWords + 5] = 0x8000000000000000ull;
            }
            if (!hip_ok(hipMemcpyAsync(J->fused_acc, init.data(), init.size() * 8, hipMemcpyHostToDevice, stream()), "fused init")) return false;
        }
        const int grid = grid_for(n, kProbeBlock, 1);
        static bool attr_set = false;
        if (!attr_set) {
            (void)hipFuncSetAttribute((const void*)k_probe<KW, true>, hipFuncAttributeMaxDynamicSharedMemorySize, 128 * 1024);
            (void)hipFuncSetAttribute((const void*)k_probe<KW, false>, hipFuncAttributeMaxDynamicSharedMemorySize, 128 * 1024);
            attr_set = true;
        }
        {
            TimedRegion timer("join_probe");
            if (lds_table) hipLaunchKernelGGL((k_probe<KW, true>), dim3(grid), dim3(kProbeBlock), lds, stream(), A);
            else hipLaunchKernelGGL((k_probe<KW, false>), dim3(grid), dim3(kProbeBlock), lds, stream(), A);
        }
        unsigned long long count = 0;
        if (!hip_ok(hipGetLastError(), "probe launch")) return false;
        if (!hip_ok(hipMemcpyAsync(&count, J->pair_count, 8, hipMemcpyDeviceToHost, stream()), "pair count") || !sync("probe")) return false;
        if (count <= A.pair_cap) {
            pairs->Count = (int64_t)count;
            return true;
        }
        cap = count;  // cannot happen with the bound above; retry once
    }
    return fail("probe pair buffer overflow");
}

// Host-side fused results of the probes since Init.
struct FusedTotals {
    std::vector<QdbFusedResult> r;
};

std::mutex g_fused_mu;
std::unordered_map<JoinBuild*, FusedTotals>& fused_totals()
{
    static auto* m = new std::unordered_map<JoinBuild*, FusedTotals>();
    return *m;
}

bool collect_fused(JoinBuild* J)
{
    JoinPlan* P = J->plan;
    const int nf = P->spec.NumFused;
    if (!nf) return true;
    std::vector<unsigned long long> w((size_t)nf * kAccWords);
    if (!hip_ok(hipMemcpyAsync(w.data(), J->fused_acc, w.size() * 8, hipMemcpyDeviceToHost, stream()), "fused read") || !sync("fused")) return false;
    std::lock_guard<std::mutex> lock(g_fused_mu);
    FusedTotals& T = fused_totals()[J];
    if (T.r.empty()) {
        T.r.resize((size_t)nf);
        for (int f = 0; f < nf; ++f) {
            memset(&T.r[(size_t)f], 0, sizeof(QdbFusedResult));
        }
    }
    for (int f = 0; f < nf; ++f) {
        const unsigned long long* a = &w[(size_t)f * kAccWords];
        QdbFusedResult& r = T.r[(size_t)f];
        const FusedDesc d = P->fused[f];
        const bool first = r.Count == 0;
        r.Count += (int64_t)a[0];
        if (d.kind == FK_SUM) {
            u128 prev = ((u128)(uint64_t)r.Hi << 64) | (u128)(uint64_t)r.Lo;
            u128 add = (u128)a[1] + ((u128)a[2] << 32) + ((u128)a[3] << 64);
            u128 s = prev + add;
            r.Lo = (int64_t)(uint64_t)s;
            r.Hi = (int64_t)(uint64_t)(s >> 64);
        } else if (d.kind == FK_MIN && a[0]) {
            int64_t v = (int64_t)a[4];
            if (first || v < r.Lo) r.Lo = v;
            r.Hi = r.Lo < 0 ? -1 : 0;
        } else if (d.kind == FK_MAX && a[0]) {
            int64_t v = (int64_t)a[5];
            if (first || v > r.Lo) r.Lo = v;
            r.Hi = r.Lo < 0 ? -1 : 0;
        } else if (d.kind == FK_COUNT_PAIRS || d.kind == FK_COUNT_VALID) {
            r.Lo = r.Count;
            r.Hi = 0;
        }
        // COUNT is never NULL; SUM/MIN/MAX of no value is NULL.
        r.Valid = (d.kind == FK_COUNT_PAIRS || d.kind == FK_COUNT_VALID) ? 1 : (r.Count > 0);
    }
    return true;
}

bool jr_dispatch_impl(JoinPlan* P, void* arena, void* table, void* scratch, void* pairs, TRowSet* batch, const int64_t* key_columns,
    int64_t op)
{
    (void)scratch;
    (void)key_columns;  // the compiled spec fixes the key columns
    clear_error();
    if (!P) return fail("join slot is empty");
    RowArenaBlob* AB = (RowArenaBlob*)arena;
    RowTableBlob* TB = (RowTableBlob*)table;
    QdbPairBuffer* PB = (QdbPairBuffer*)pairs;
    switch (op) {
    case QDB_JR_OP_INIT: {
        if (AB->build) {
            std::lock_guard<std::mutex> lock(g_fused_mu);
            fused_totals().erase(AB->build);
        }
        build_free(AB->build);
        memset(AB, 0, sizeof(*AB));
        JoinBuild* J = new JoinBuild();
        J->plan = P;
        J->pdata.assign((size_t)P->npayload, nullptr);
        J->pvalid.assign((size_t)P->npayload, nullptr);
        J->fused_acc = (unsigned long long*)dalloc((size_t)kMaxFused * kAccWords * 8);
        J->pair_count = (unsigned long long*)dalloc(16);
        if (!J->fused_acc || !J->pair_count) {
            build_free(J);
            return false;
        }
        AB->build = J;
        AB->RowBytes = (int64_t)P->KW * 8 + 8;
        AB->BlockRows = 16384;
        return true;
    }
    case QDB_JR_OP_APPEND: {
        JoinBuild* J = AB->build;
        if (!J) return fail("join append before Init");
        if (J->built) return fail("join append after Build");
        if (!batch) return true;
        bool ok = P->KW == 1 ? do_append<1>(J, batch) : do_append<2>(J, batch);
        AB->Rows = J->rows;
        return ok;
    }
    case QDB_JR_OP_BUILD: {
        JoinBuild* J = AB->build;
        if (!J) return fail("join build before Init");
        bool ok = P->KW == 1 ? do_build<1>(J) : do_build<2>(J);
        if (!ok) return false;
        TB->entries = J->table;
        TB->Capacity = (int64_t)J->capacity;
        TB->Mask = J->capacity - 1;
        int shift = 64;
        for (uint64_t c = J->capacity; c > 1; c >>= 1) --shift;
        TB->Shift = shift;
        TB->Rows = J->rows;
        TB->Keys = J->keys_distinct;
        return true;
    }
    case QDB_JR_OP_PROBE: {
        JoinBuild* J = AB->build;
        if (!J || !J->built) return fail("join probe before Build");
        if (!batch) return fail("join probe without a batch");
        bool ok = P->KW == 1 ? do_probe<1>(J, batch, PB) : do_probe<2>(J, batch, PB);
        if (ok) ok = collect_fused(J);
        if (ok) {
            std::lock_guard<std::mutex> lock(g_pairs_mu);
            pair_owner()[pairs] = J;
        }
        return ok;
    }
    default:
        return false;  // SEMI/ANTI ops are not served by this kernel
    }
}

int64_t jr_materialize_impl(JoinPlan* P, void* pairs, TRowSet* stream_rs, int64_t start, int64_t limit, TRowSet* out)
{
    clear_error();
    QdbPairBuffer* PB = (QdbPairBuffer*)pairs;
    JoinBuild* J = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_pairs_mu);
        auto it = pair_owner().find(pairs);
        if (it != pair_owner().end()) J = it->second;
    }
    if (!P || !J || J->plan != P) {
        fail("materialize: the pair buffer was not filled by this join");
        return 0;
    }
    if (start < 0 || start >= PB->Count || limit <= 0) return 0;
    const int64_t n = PB->Count - start < limit ? PB->Count - start : limit;
    const QdbColumnSchema& ps = P->spec.Probe;
    OwnedRowSet* o = init_owned_rowset(out, n, P->nout);
    const size_t mask_bytes = (size_t)((n + 7) / 8) + 8;
    std::vector<GatherCol> all((size_t)P->nout);
    for (int c = 0; c < P->nout; ++c) {
        GatherCol g = P->gather_proto[c];
        const QdbValueType& t = P->out_types[c];
        const size_t bytes = t.Kind == QDB_KIND_BOOL ? mask_bytes : (size_t)n * (size_t)t.Width + 8;
        char* data = (char*)dalloc(bytes);
        if (!data) return 0;
        o->dev.push_back(data);
        o->cols[(size_t)c].Data = data;
        g.out_data = data;
        if (t.Nullable) {
            uint8_t* m = (uint8_t*)dalloc(mask_bytes);
            if (!m) return 0;
            o->dev.push_back(m);
            o->cols[(size_t)c].Mask = m;
            g.out_mask = m;
        }
        if (g.from_build) {
            g.arena_data = J->pdata[(size_t)g.payload];
            g.arena_valid = P->spec.Build.Types[g.payload].Nullable ? J->pvalid[(size_t)g.payload] : nullptr;
        }
        all[(size_t)c] = g;
    }
    // Launches of at most kGatherMax output columns and kMaxDevCols probe columns.
    size_t c = 0;
    while (c < all.size()) {
        GatherArgs g;
        memset(&g, 0, sizeof(g));
        g.pairs = PB->Data;
        g.start = start;
        g.n = n;
        g.probe_is_left = P->spec.ProbeIsLeft;
        g.ord_to_arena = J->ord_to_arena;
        int slot_col[kMaxDevCols];
        int ns = 0;
        while (c < all.size() && g.ncols < kGatherMax) {
            GatherCol gc = all[c];
            if (!gc.from_build) {
                if (ns == kMaxDevCols) break;
                gc.slot = ns;
                slot_col[ns++] = gc.payload;
            }
            g.col[g.ncols++] = gc;
            ++c;
        }
        if (!bind_side(slot_col, ns, ps, stream_rs, &g.probe)) return 0;
        TimedRegion timer("join_materialize");
        hipLaunchKernelGGL(k_materialize, dim3(grid_for((n + 7) / 8, kBlock, 8)), dim3(kBlock), 0, stream(), g);
        if (!hip_ok(hipGetLastError(), "materialize launch")) return 0;
    }
    if (!sync("materialize")) return 0;
    if (P->spec.OutputOnHost && !move_rowset_to_host(out, P->out_types)) return 0;
    return n;
}

template <int N>
struct JoinDispatchTramp {
    static bool call(void* arena, void* table, void* scratch, void* pairs, TRowSet* batch, const int64_t* key_columns, int64_t op)
    {
        return jr_dispatch_impl(join_slots().get(N), arena, table, scratch, pairs, batch, key_columns, op);
    }
};

template <int N>
struct JoinMaterializeTramp {
    static int64_t call(void* pairs, TRowSet* stream_rs, int64_t start, int64_t limit, TRowSet* out)
    {
        return jr_materialize_impl(join_slots().get(N), pairs, stream_rs, start, limit, out);
    }
};

const auto kJoinDispatchFns = make_tramp_table<QdbJoinDispatchFn, JoinDispatchTramp>(std::make_integer_sequence<int, kSlots>{});
const auto kJoinMaterializeFns = make_tramp_table<QdbJoinMaterializeFn, JoinMaterializeTramp>(std::make_integer_sequence<int, kSlots>{});

}  // namespace

void release_join_slot(int slot)
{
    join_slots().release(slot);
}

}  // namespace qdbgpu

using namespace qdbgpu;

extern "C" int qdb_gpu_compile_row_join(const QdbRowJoinSpec* spec, QdbRowJoinKernels* out)
{
    clear_error();
    if (!spec || !out) return fail("null argument"), -1;
    if (!ensure_init()) return -1;
    JoinPlan* P = new JoinPlan();
    if (!compile_join(*spec, P)) {
        delete P;
        return -1;
    }
    const int slot = join_slots().add(P);
    if (slot < 0) {
        delete P;
        return fail("all join slots are in use"), -1;
    }
    memset(out, 0, sizeof(*out));
    out->Dispatch = kJoinDispatchFns.fns[slot];
    out->Materialize = kJoinMaterializeFns.fns[slot];
    out->Slot = slot;
    out->NumOutputColumns = P->nout;
    for (int c = 0; c < P->nout; ++c) out->OutputTypes[c] = P->out_types[c];
    return 0;
}

extern "C" int64_t qdb_gpu_join_fused_results(void* arena, QdbFusedResult* out, int64_t max_out)
{
    RowArenaBlob* AB = (RowArenaBlob*)arena;
    if (!AB || !AB->build) return 0;
    JoinBuild* J = AB->build;
    const int nf = J->plan->spec.NumFused;
    std::lock_guard<std::mutex> lock(g_fused_mu);
    auto it = fused_totals().find(J);
    int64_t n = 0;
    for (int f = 0; f < nf && f < max_out; ++f) {
        if (it == fused_totals().end()) {
            memset(&out[f], 0, sizeof(QdbFusedResult));
            const uint8_t k = J->plan->fused[f].kind;
            out[f].Valid = (k == FK_COUNT_PAIRS || k == FK_COUNT_VALID) ? 1 : 0;
        } else {
            out[f] = it->second.r[(size_t)f];
        }
        ++n;
    }
    return n;
}

extern "C" void qdb_gpu_join_destroy(void* arena, void* table, void* scratch, void* pairs)
{
    (void)scratch;
    RowArenaBlob* AB = (RowArenaBlob*)arena;
    if (AB && AB->build) {
        {
            std::lock_guard<std::mutex> lock(g_fused_mu);
            fused_totals().erase(AB->build);
        }
        {
            std::lock_guard<std::mutex> lock(g_pairs_mu);
            for (auto it = pair_owner().begin(); it != pair_owner().end();) {
                if (it->second == AB->build) it = pair_owner().erase(it);
                else ++it;
            }
        }
        build_free(AB->build);
        memset(AB, 0, sizeof(*AB));
    }
    if (table) memset(table, 0, QDB_ROW_TABLE_SIZE);
    QdbPairBuffer* PB = (QdbPairBuffer*)pairs;
    if (PB) {
        dfree(PB->Data);
        memset(PB, 0, sizeof(*PB));
    }
}
