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

#include <stdlib.h>
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
// Words per fused aggregate: [count, sum partial words x3, min, max].
constexpr int kAccWords = 6;
// Host readback words: the pair count, one spare, the fused words.
constexpr int kReadbackWords = 2 + kMaxFused * kAccWords;

// The probe's counters before a launch: the pair count 0 and the first
// `words` fused words at their identities (sums 0, MIN INT64_MAX, MAX
// INT64_MIN); thread t sets word t.
__device__ __forceinline__ void probe_counters_init(unsigned long long* pair_count, unsigned long long* acc, int words, int t)
{
    if (t == 0) pair_count[0] = 0ull;
    if (t < words) {
        const int k = t % kAccWords;
        acc[t] = k == 4 ? 0x7fffffffffffffffull : k == 5 ? 0x8000000000000000ull : 0ull;
    }
}

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

// Kept rows per block; with `block_range` (one-word keys) also the block's
// largest ~key and largest key of its kept rows (0, 0 when none).
template <int KW>
__global__ void __launch_bounds__(kBlock) k_append_count(DevCols cols, JKeyDesc K, int64_t n, const uint8_t* sel,
    unsigned int* block_counts, unsigned long long* block_range)
{
    __shared__ unsigned int total;
    __shared__ unsigned long long range[2];
    if (threadIdx.x == 0) {
        total = 0;
        range[0] = 0;
        range[1] = 0;
    }
    __syncthreads();
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    bool keep = false;
    uint64_t key[KW];
    key[0] = 0;
    if (i < n && (!sel || sel[i])) keep = join_key<KW>(cols, K, i, key);
    const uint64_t b = __ballot(keep);
    if (lane_id() == 0 && b) atomicAdd(&total, (unsigned int)__popcll(b));
    if (KW == 1 && block_range && keep) {
        atomicMax(&range[0], (unsigned long long)~key[0]);
        atomicMax(&range[1], (unsigned long long)key[0]);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        block_counts[blockIdx.x] = total;
        if (KW == 1 && block_range) {
            block_range[2 * blockIdx.x] = range[0];
            block_range[2 * blockIdx.x + 1] = range[1];
        }
    }
}

// Exclusive scan of `n` block counts in place (one workgroup), offset by
// the arena's row count: each block's first arena row. The arena's words
// then take the new row count and, with `block_range`, the key range of
// the rows so far. `fresh`: the arena is empty and its words not yet set
// (the first append writes all of them).
__global__ void __launch_bounds__(1024) k_scan_counts(unsigned int* counts, unsigned int n, unsigned long long* arena,
    const unsigned long long* block_range, int fresh)
{
    __shared__ unsigned long long partial[1024];
    __shared__ unsigned long long base;
    __shared__ unsigned long long wrange[1024 / kWave][2];
    if (threadIdx.x == 0) base = fresh ? 0ull : arena[0];
    if (block_range) {
        unsigned long long nmin = 0, mx = 0;
        for (unsigned int b = threadIdx.x; b < n; b += blockDim.x) {
            nmin = block_range[2 * b] > nmin ? block_range[2 * b] : nmin;
            mx = block_range[2 * b + 1] > mx ? block_range[2 * b + 1] : mx;
        }
#pragma unroll
        for (int o = 1; o < kWave; o <<= 1) {
            const unsigned long long a = shfl_xor_u64(nmin, o), c = shfl_xor_u64(mx, o);
            nmin = a > nmin ? a : nmin;
            mx = c > mx ? c : mx;
        }
        if (lane_id() == 0) {
            wrange[threadIdx.x / kWave][0] = nmin;
            wrange[threadIdx.x / kWave][1] = mx;
        }
    }
    const unsigned int per = (n + blockDim.x - 1) / blockDim.x;
    const unsigned int lo = threadIdx.x * per;
    const unsigned int hi = lo + per < n ? lo + per : n;
    unsigned long long s = 0;
    for (unsigned int i = lo; i < hi; ++i) s += counts[i];
    // Exclusive scan of the threads' sums: within each wave by shuffles,
    // then over the waves' totals.
    const int lane = lane_id(), wave = threadIdx.x / kWave, nw = blockDim.x / kWave;
    unsigned long long incl = s;
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        const unsigned int ylo = (unsigned int)__shfl_up((int)(unsigned int)incl, o, kWave);
        const unsigned int yhi = (unsigned int)__shfl_up((int)(unsigned int)(incl >> 32), o, kWave);
        if (lane >= o) incl += ((unsigned long long)yhi << 32) | ylo;
    }
    if (lane == kWave - 1) partial[wave] = incl;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long run = base;
        for (int w = 0; w < nw; ++w) {
            const unsigned long long v = partial[w];
            partial[w] = run;
            run += v;
        }
        arena[0] = run;
        unsigned long long nmin = fresh ? 0ull : arena[2], mx = fresh ? 0ull : arena[3];
        if (block_range) {
            for (int w = 0; w < nw; ++w) {
                nmin = wrange[w][0] > nmin ? wrange[w][0] : nmin;
                mx = wrange[w][1] > mx ? wrange[w][1] : mx;
            }
        }
        if (fresh) arena[1] = 0ull;
        arena[2] = nmin;
        arena[3] = mx;
    }
    __syncthreads();
    unsigned long long run = partial[wave] + incl - s;
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
    const unsigned int* block_offsets, int64_t ordinal_base, uint64_t* arena_keys, int64_t* arena_ordinal,
    int32_t* ord_to_arena, DevCols payload_cols, int npayload, ArenaCols AC)
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
    const int64_t dst = (int64_t)block_offsets[blockIdx.x] + rank;
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
// Build: distinct keys, CSR order, the probe table, Bloom filter.
// ---------------------------------------------------------------------------
//
// The insert-time slot {key, head, count} becomes the probe-time slot after
// the build: count 0 = empty; count 1: head = the key's build ordinal;
// count > 1: head = the first of the key's `count` rows in CSR order (the
// rows of a key contiguous, in build order). A match costs one 16-byte slot
// read (KW = 1).

__device__ __forceinline__ uint32_t bloom_word(uint64_t h, uint32_t words_mask)
{
    return (uint32_t)(h >> 20) & words_mask;
}

__device__ __forceinline__ uint32_t bloom_bits(uint64_t h)
{
    return (1u << ((h >> 52) & 31u)) | (1u << ((h >> 58) & 31u));
}

// One insert attempt of arena row `r`: the slot of its key, or -1 when the
// slot it needs is being filled by another lane. The caller retries in a
// loop that the whole wave runs together: a lane that spun here could wait
// forever on a lane of its own wave whose publishing store the compiler
// placed after the loop.
template <int KW>
__device__ __forceinline__ int64_t build_insert_once(Slot<KW>* table, uint64_t cmask, const uint64_t (&key)[KW], uint64_t h,
    int64_t r)
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
                __hip_atomic_store(&sl->head, (uint32_t)(r + 1), __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
                return (int64_t)s;
            }
        }
        if (head == kLockHead) return -1;
        bool eq = true;
#pragma unroll
        for (int w = 0; w < KW; ++w) eq = eq && __hip_atomic_load(&sl->key[w], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == key[w];
        if (eq) return (int64_t)s;
        s = (s + 1) & cmask;
    }
}

// One-word keys when a value no build key takes is known (`empty`, the
// table's key words start there): a slot is claimed by one 64-bit CAS of its
// key word, which also tells a later row of the same key where it went, so
// no acquire or release is needed (on this part those write back or
// invalidate L2 on every insert). Emptiness is read from the row count by
// every later kernel.
__device__ __forceinline__ int64_t build_insert_cas(Slot<1>* table, uint64_t cmask, uint64_t key, uint64_t h, uint64_t empty)
{
    uint64_t s = h & cmask;
    for (;;) {
        const unsigned long long old = atomicCAS((unsigned long long*)&table[s].key[0], (unsigned long long)empty,
            (unsigned long long)key);
        if (old == empty || old == key) return (int64_t)s;
        s = (s + 1) & cmask;
    }
}

// Key words of a one-word-key table to `empty`, heads and counts to 0.
__global__ void __launch_bounds__(kBlock) k_fill_slots(Slot<1>* table, uint64_t capacity, uint64_t empty)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t s = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; s < capacity; s += stride) {
        table[s].key[0] = empty;
        table[s].head = 0u;
        table[s].count = 0u;
    }
}

// Inserts every arena row's key (distinct keys get a slot), counts the rows
// of each key and records each row's slot. `flags[0]` is set when a key has
// more than one row. use_empty (KW = 1): claim slots with build_insert_cas.
template <int KW>
__global__ void __launch_bounds__(kBlock) k_build_insert(Slot<KW>* table, uint64_t cmask, const uint64_t* arena_keys,
    int64_t rows, uint32_t* row_slot, unsigned int* flags, int use_empty, uint64_t empty)
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
        int64_t s = -1;
        if constexpr (KW == 1) {
            if (use_empty) {
                if (pending) s = build_insert_cas((Slot<1>*)table, cmask, key[0], h, empty);
                pending = false;
            }
        }
        while (__any(pending)) {
            if (pending) {
                s = build_insert_once<KW>(table, cmask, key, h, r);
                pending = s < 0;
            }
        }
        if (r < rows) {
            row_slot[r] = (uint32_t)s;
            if (atomicAdd(&table[s].count, 1u) != 0u) flags[0] = 1u;
        }
    }
}

// Per-slot row counts (for the scan) and the Bloom filter bits of every
// distinct key.
// Per-slot row counts (for the scan), the Bloom filter bits of every
// distinct key, and per workgroup the number of distinct keys and the most
// rows one key has (part[2 * block], part[2 * block + 1]).
template <int KW>
__global__ void __launch_bounds__(kBlock) k_slot_counts(const Slot<KW>* table, uint64_t capacity, uint32_t* counts,
    uint32_t* bloom, uint32_t bloom_mask, unsigned long long* distinct_part)
{
    __shared__ unsigned long long wsum[kBlock / kWave], wmax[kBlock / kWave];
    unsigned long long used = 0, most = 0;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t s = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; s < capacity; s += stride) {
        const Slot<KW> sl = table[s];
        counts[s] = sl.count;
        used += sl.count != 0u ? 1u : 0u;
        most = sl.count > most ? sl.count : most;
        if (sl.count == 0u || !bloom) continue;
        uint64_t key[KW];
#pragma unroll
        for (int w = 0; w < KW; ++w) key[w] = sl.key[w];
        const uint64_t h = key_place_hash<KW>(key);
        atomicOr(&bloom[bloom_word(h, bloom_mask)], bloom_bits(h));
    }
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        used += shfl_xor_u64(used, o);
        const unsigned long long m = shfl_xor_u64(most, o);
        most = m > most ? m : most;
    }
    if (lane_id() == 0) {
        wsum[threadIdx.x / kWave] = used;
        wmax[threadIdx.x / kWave] = most;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0, m = 0;
        for (int w = 0; w < kBlock / kWave; ++w) {
            t += wsum[w];
            m = wmax[w] > m ? wmax[w] : m;
        }
        distinct_part[2 * blockIdx.x] = t;
        distinct_part[2 * blockIdx.x + 1] = m;
    }
}

// Distinct keys: each row goes to the first position of its slot.
__global__ void __launch_bounds__(kBlock) k_place_unique(const uint32_t* row_slot, const uint32_t* start, int64_t rows,
    uint32_t* perm)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < rows; r += stride) perm[start[row_slot[r]]] = (uint32_t)r;
}

__global__ void __launch_bounds__(kBlock) k_iota(uint32_t* v, int64_t n)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += stride) v[i] = (uint32_t)i;
}

// The arena keys and ordinals in CSR order, and ord_to_arena.
template <int KW>
__global__ void __launch_bounds__(kBlock) k_build_csr(const uint32_t* perm, int64_t rows, const uint64_t* keys,
    const int64_t* ordinal, uint64_t* new_keys, int64_t* new_ordinal, int32_t* ord_to_arena)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < rows; p += stride) {
        const uint32_t r = perm[p];
#pragma unroll
        for (int w = 0; w < KW; ++w) new_keys[p * KW + w] = keys[(int64_t)r * KW + w];
        const int64_t ord = ordinal[r];
        new_ordinal[p] = ord;
        ord_to_arena[ord] = (int32_t)p;
    }
}

// Insert-time slots -> probe-time slots.
template <int KW>
__global__ void __launch_bounds__(kBlock) k_finish_slots(Slot<KW>* table, uint64_t capacity, const uint32_t* start,
    const int64_t* csr_ordinal)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t s = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; s < capacity; s += stride) {
        const uint32_t c = table[s].count;
        if (c == 0u) continue;
        table[s].head = c == 1u ? (uint32_t)csr_ordinal[start[s]] : start[s];
    }
}

// Smallest and largest packed key of a one-word-key arena: a partial per
// workgroup (part[2 * block], part[2 * block + 1]), then k_reduce_partials.
__global__ void __launch_bounds__(kBlock) k_key_minmax(const uint64_t* keys, const unsigned long long* rows_dev,
    unsigned long long* part)
{
    __shared__ unsigned long long s_lo[kBlock / kWave], s_hi[kBlock / kWave];
    const int64_t rows = (int64_t)*rows_dev;
    unsigned long long lo = ~0ull, hi = 0ull;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < rows; r += stride) {
        const unsigned long long k = keys[r];
        lo = k < lo ? k : lo;
        hi = k > hi ? k : hi;
    }
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        const unsigned long long a = shfl_xor_u64(lo, o), b = shfl_xor_u64(hi, o);
        lo = a < lo ? a : lo;
        hi = b > hi ? b : hi;
    }
    if (lane_id() == 0) {
        s_lo[threadIdx.x / kWave] = lo;
        s_hi[threadIdx.x / kWave] = hi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < kBlock / kWave; ++w) {
            lo = s_lo[w] < lo ? s_lo[w] : lo;
            hi = s_hi[w] > hi ? s_hi[w] : hi;
        }
        part[2 * blockIdx.x] = lo;
        part[2 * blockIdx.x + 1] = hi;
    }
}

// Reduces `n` partials of `words` u64 words each (partial p, word k at
// part[p * stride + k]) into out[k]: op[k] 0 = sum, 1 = signed min, 2 =
// signed max, 3 = unsigned min, 4 = unsigned max. One workgroup of kBlock
// threads; each wave folds whole words.
struct ReduceOps {
    int32_t op[kMaxFused * 6];
};

__global__ void __launch_bounds__(kBlock) k_reduce_partials(const unsigned long long* part, int n, int stride, int words,
    ReduceOps ops, unsigned long long* out)
{
    // Wave w folds words w, w + kBlock / kWave, ...: its lanes' loads of a
    // word are independent, and no barrier is needed.
    const int lane = lane_id();
    for (int k = threadIdx.x / kWave; k < words; k += kBlock / kWave) {
        const int op = ops.op[k];
        auto comb = [op](unsigned long long a, unsigned long long b) -> unsigned long long {
            switch (op) {
            case 0: return a + b;
            case 1: return (int64_t)b < (int64_t)a ? b : a;
            case 2: return (int64_t)b > (int64_t)a ? b : a;
            case 3: return b < a ? b : a;
            default: return b > a ? b : a;
            }
        };
        const unsigned long long id = op == 0 ? 0ull : op == 1 ? 0x7fffffffffffffffull : op == 2 ? 0x8000000000000000ull
                                                   : op == 3 ? ~0ull : 0ull;
        unsigned long long x = id;
        for (int p = lane; p < n; p += kWave) x = comb(x, part[(size_t)p * stride + k]);
#pragma unroll
        for (int o = 1; o < kWave; o <<= 1) x = comb(x, shfl_xor_u64(x, o));
        if (lane == 0) out[k] = x;
    }
}

constexpr uint32_t kAbsent = 0xffffffffu;   // dense table: no build row for this key

// The arena's device words: [0] the row count, [1] the dense build's flags
// (u32 repeated key, u32 range too wide), [2] the largest ~key and [3] the
// largest key of the rows appended (one-word keys; 0 and 0 while empty).
constexpr int kArenaWords = 4;

// The key range [min, max] of the dense table when it holds at most
// dense_limit(rows) keys.
__device__ __forceinline__ bool dense_range(const unsigned long long* arena, uint64_t* lo, uint64_t* range)
{
    const uint64_t rows = arena[0];
    *lo = ~arena[2];
    const uint64_t hi = arena[3];
    *range = hi - *lo + 1;
    return rows > 0 && hi >= *lo && *range != 0 && *range <= (rows * 8 > (1ull << 16) ? rows * 8 : (1ull << 16));
}

// Clears the dense table over the key range (no membership bits, kAbsent
// ordinals), or flags a range too wide.
__global__ void __launch_bounds__(kBlock) k_dense_prepare(unsigned long long* arena, uint32_t* dbits, uint32_t* dtab)
{
    uint64_t lo = 0, range = 0;
    const bool ok = dense_range(arena, &lo, &range);
    if (!ok) {
        if (blockIdx.x == 0 && threadIdx.x == 0) ((unsigned int*)(arena + 1))[1] = 1u;
        return;
    }
    const uint64_t words = (range + 31) / 32;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < range; i += stride) {
        dtab[i] = kAbsent;
        if (i < words) dbits[i] = 0u;
    }
}

// The dense table: a membership bit and the build ordinal of every row at
// key - min (kAbsent where no row has the key); the repeated-key flag set
// when a key repeats. Workgroup 0 also sets the first probe's counters.
__global__ void __launch_bounds__(kBlock) k_dense_fill(const uint64_t* keys, const int64_t* ordinal, unsigned long long* arena,
    uint32_t* dbits, uint32_t* dtab, unsigned long long* pair_count, unsigned long long* acc, int acc_words)
{
    uint64_t dmin = 0, range = 0;
    if (!dense_range(arena, &dmin, &range)) return;
    if (blockIdx.x == 0) probe_counters_init(pair_count, acc, acc_words, (int)threadIdx.x);
    const int64_t rows = (int64_t)arena[0];
    unsigned int* flags = (unsigned int*)(arena + 1);
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < rows; r += stride) {
        const uint64_t idx = keys[r] - dmin;
        const uint32_t bit = 1u << (idx & 31u);
        if (atomicOr(&dbits[idx >> 5], bit) & bit) flags[0] = 1u;
        dtab[idx] = (uint32_t)ordinal[r];
    }
}

// The ordinal table packed at w bits per key (w < 32; kAbsent becomes all
// ones): thread per output word j assembles the entries that overlap bits
// [32 j, 32 j + 32). The last word is slack for two-word reads.
__global__ void __launch_bounds__(kBlock) k_dense_pack(const uint32_t* dtab, uint64_t range, uint32_t w, uint32_t* pack,
    uint64_t nwords)
{
    const uint64_t mask = (1ull << w) - 1ull;
    for (uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; j < nwords; j += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t s = j * 32;
        uint64_t word = 0;
        const uint64_t e1 = (s + 31) / w;
        for (uint64_t e = s / w; e <= e1 && e < range; ++e) {
            const uint32_t o = dtab[e];
            const uint64_t v = o == kAbsent ? mask : (uint64_t)o;
            const uint64_t at = e * w;
            word |= at >= s ? (v << (at - s)) : (v >> (s - at));
        }
        pack[j] = (uint32_t)word;
    }
}

// One payload column (and its validity bytes) gathered into CSR order.
__global__ void __launch_bounds__(kBlock) k_permute_column(const uint32_t* perm, int64_t rows, const char* src, char* dst,
    int width, const uint8_t* vsrc, uint8_t* vdst)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < rows; p += stride) {
        const int64_t r = perm[p];
        switch (width) {
        case 1: ((uint8_t*)dst)[p] = ((const uint8_t*)src)[r]; break;
        case 2: ((uint16_t*)dst)[p] = ((const uint16_t*)src)[r]; break;
        case 4: ((uint32_t*)dst)[p] = ((const uint32_t*)src)[r]; break;
        case 8: ((uint64_t*)dst)[p] = ((const uint64_t*)src)[r]; break;
        default: {
            const uint64_t* s = (const uint64_t*)(src + 16 * r);
            uint64_t* d = (uint64_t*)(dst + 16 * p);
            d[0] = s[0];
            d[1] = s[1];
        }
        }
        if (vsrc) vdst[p] = vsrc[r];
    }
}

// ---------------------------------------------------------------------------
// Probe.
// ---------------------------------------------------------------------------

enum FusedKind : int32_t { FK_COUNT_PAIRS, FK_COUNT_VALID, FK_SUM, FK_MIN, FK_MAX };

struct FusedDesc {
    int32_t kind;
    int32_t from_build;   // argument from the build arena
    int32_t slot;         // probe column slot
    int32_t payload;      // build payload index
};

struct ProbeArgs {
    DevCols cols;         // probe columns used (keys, fused args)
    JKeyDesc K;
    int32_t key_i64;      // one non-nullable 8-byte key in cols.c[0]: direct loads
    int32_t pad0;
    int64_t n;
    const uint8_t* sel;
    const void* table;          // probe-time slots
    uint64_t cmask;
    const int64_t* ordinal;     // build ordinals in CSR order
    const int32_t* ord_to_arena;
    const uint32_t* bloom;
    uint32_t bloom_words;       // power of two, or 0 (no filter)
    uint32_t lds_words;         // dense: bitmap words staged in LDS
    uint64_t dmin;              // dense: smallest build key
    uint64_t drange;            // dense: max - min + 1
    const uint32_t* dbits;      // dense: key - min membership bitmap
    const uint32_t* dtab;       // dense: build ordinal by key - min
    int32_t emit_unmatched;
    int32_t probe_is_left;
    int32_t val_slot;           // the one probe column of the fused aggregates (kMode 2)
    int64_t* pairs;             // 2 i64 per pair
    unsigned long long* pair_count;
    unsigned long long pair_cap;
    int32_t nfused;
    int32_t pad2;
    FusedDesc fused[kMaxFused];
    const char* payload_data[kMaxFused];
    const uint8_t* payload_valid[kMaxFused];
    int32_t payload_width[kMaxFused];
    // Per workgroup and fused agg: [count, p0, p1, p2, min, max] words,
    // which k_reduce_partials folds into the join's totals; when null, each
    // workgroup adds its words to fused_acc with atomics instead.
    unsigned long long* fused_part;
    unsigned long long* fused_acc;
    // dense, lean probe: the ordinals packed in dpack_w bits each (all ones:
    // no build row), or null
    const uint32_t* dpack;
    uint32_t dpack_w;
    uint32_t pad3;
};

// Probe kernel modes: kGeneric reads keys through the column descriptors;
// kDirectKey reads one non-nullable 8-byte integer key with a streaming
// (non-temporal) load; kDirectVal also streams the one non-nullable 8-byte
// integer probe column that every probe-side fused aggregate reads, beside
// the key, so that the fused aggregate needs no load of its own.
constexpr int kGeneric = 0, kDirectKey = 1, kDirectVal = 2;

// Build tables the probe reads: the hash table in HBM behind a Bloom filter
// in LDS, the whole hash table in LDS, or the dense direct table (the
// membership bitmap in LDS as far as it fits, the ordinal table in HBM).
constexpr int kTabHash = 0, kTabLds = 1, kTabDense = 2;
// Device scratch of the allocation-free probe (qdb_gpu_join_probe_fused_
// noalloc): the fused partials of at most 512 workgroups.
constexpr size_t kNoAllocScratchBytes = 128 * 1024;
// Pairs the first probe attempt allocates room for at most (2 GiB of pairs).
constexpr unsigned long long kFirstPairBudget = 1ull << 27;

// The key of selected probe row i; false when a key field is NULL.
template <int KW, bool kDirect>
__device__ __forceinline__ bool probe_row_key(const ProbeArgs& P, int64_t i, uint64_t (&key)[KW])
{
    if (kDirect) {
        key[0] = __builtin_nontemporal_load((const uint64_t*)P.cols.c[0].data + i);
#pragma unroll
        for (int w = 1; w < KW; ++w) key[w] = 0;
        return true;
    }
    return join_key<KW>(P.cols, P.K, i, key);
}

// The rest of a lookup after the slot `sl` at position s was read: the
// key's row count (0 = no match) and the slot's head.
template <int KW>
__device__ __forceinline__ uint32_t lookup_from(const ProbeArgs& P, const Slot<KW>* T, const uint64_t (&key)[KW], uint64_t s,
    Slot<KW> sl, uint32_t* head)
{
    for (;;) {
        if (sl.count == 0u) return 0u;
        bool eq = true;
#pragma unroll
        for (int w = 0; w < KW; ++w) eq = eq && sl.key[w] == key[w];
        if (eq) {
            *head = sl.head;
            return sl.count;
        }
        s = (s + 1) & P.cmask;
        sl = T[s];
    }
}

// Per-lane state of one fused aggregate: count and either the i128 sum
// (lo, hi) or the running min / max in `lo`.
struct FusedLane {
    unsigned long long count;
    uint64_t lo, hi;
};

__device__ __forceinline__ void fused_add(FusedLane& a, int kind, uint32_t times, i128 x)
{
    a.count += times;
    if (kind == FK_SUM) {
        const u128 s = ((((u128)a.hi) << 64) | (u128)a.lo) + (u128)x * (u128)times;
        a.lo = (uint64_t)s;
        a.hi = (uint64_t)(s >> 64);
    } else if (kind == FK_MIN) {
        a.lo = (int64_t)x < (int64_t)a.lo ? (uint64_t)(int64_t)x : a.lo;
    } else if (kind == FK_MAX) {
        a.lo = (int64_t)x > (int64_t)a.lo ? (uint64_t)(int64_t)x : a.lo;
    }
}

constexpr int kSub = 4;                                  // rows per lane per load round
constexpr int kSubRounds = 3;                            // rounds per wave tile
constexpr int kTileRows = kSub * kSubRounds;             // rows per lane per wave tile
constexpr int kWaveTile = kWave * kTileRows;             // rows per output reservation

typedef long long ll2 __attribute__((ext_vector_type(2)));

// Persistent probe. Every wave works alone on tiles of kWaveTile rows (no
// workgroup barrier after the filter is loaded): per round of kSub rows per
// lane, the keys, then the first slot of every row that passes the Bloom
// filter, are loaded as independent rounds. The tile's pairs are placed with
// one atomic per wave tile; within the tile they are ordered row-major across
// the wave so that every store instruction writes a contiguous run. The
// fused aggregate of a probe column is applied once per matched row (times
// its match count) into per-lane state, combined once at the end.
// Writes a workgroup's fused state: wave reduction (add words hold the sum
// as bits [0,32), [32,64), [64,128)), the waves' partials through LDS, and
// one partial per workgroup with plain stores; k_reduce_partials combines the
// workgroups. (Atomics from every wave to the same few words serialize in
// the memory system and cost more than the probe itself.) Every thread of
// the workgroup calls it. kShared: the state is the pair count and the sum,
// min and max of the one probe column (shared by every fused aggregate).
template <int kAcc, bool kShared>
__device__ __forceinline__ void fused_flush(const ProbeArgs& P, const FusedLane (&acc)[kAcc], unsigned long long v_pairs,
    uint64_t v_lo, uint64_t v_hi, int64_t v_min, int64_t v_max, int lane)
{
    __shared__ unsigned long long part[kProbeBlock / kWave][kMaxFused][kAccWords];
    const int wave = threadIdx.x / kWave;
#pragma unroll
    for (int f = 0; f < kMaxFused; ++f) {
        if (f >= P.nfused) continue;
        const int kind = P.fused[f].kind;
        unsigned long long w[kAccWords];
        if (kShared) {
            w[0] = v_pairs;
            w[1] = kind == FK_SUM ? (v_lo & 0xffffffffull) : 0ull;
            w[2] = kind == FK_SUM ? (v_lo >> 32) : 0ull;
            w[3] = kind == FK_SUM ? v_hi : 0ull;
            w[4] = kind == FK_MIN ? (uint64_t)v_min : 0x7fffffffffffffffull;
            w[5] = kind == FK_MAX ? (uint64_t)v_max : 0x8000000000000000ull;
        } else {
            const FusedLane a = acc[f < kAcc ? f : 0];
            w[0] = a.count;
            w[1] = kind == FK_SUM ? (a.lo & 0xffffffffull) : 0ull;
            w[2] = kind == FK_SUM ? (a.lo >> 32) : 0ull;
            w[3] = kind == FK_SUM ? a.hi : 0ull;
            w[4] = kind == FK_MIN ? a.lo : 0x7fffffffffffffffull;
            w[5] = kind == FK_MAX ? a.lo : 0x8000000000000000ull;
        }
#pragma unroll
        for (int o = 1; o < kWave; o <<= 1) {
            w[0] += shfl_xor_u64(w[0], o);
            w[1] += shfl_xor_u64(w[1], o);
            w[2] += shfl_xor_u64(w[2], o);
            w[3] += shfl_xor_u64(w[3], o);
            const int64_t x4 = (int64_t)shfl_xor_u64(w[4], o);
            const int64_t x5 = (int64_t)shfl_xor_u64(w[5], o);
            w[4] = (unsigned long long)((int64_t)w[4] < x4 ? (int64_t)w[4] : x4);
            w[5] = (unsigned long long)((int64_t)w[5] > x5 ? (int64_t)w[5] : x5);
        }
        if (lane == 0) {
#pragma unroll
            for (int k = 0; k < kAccWords; ++k) part[wave][f][k] = w[k];
        }
    }
    __syncthreads();
    if (wave == 0 && lane < P.nfused * kAccWords) {
        const int f = lane / kAccWords, k = lane % kAccWords;
        unsigned long long x = part[0][f][k];
        for (int v = 1; v < kProbeBlock / kWave; ++v) {
            const unsigned long long y = part[v][f][k];
            if (k < 4) x += y;
            else if (k == 4) x = (int64_t)y < (int64_t)x ? y : x;
            else x = (int64_t)y > (int64_t)x ? y : x;
        }
        if (P.fused_part) {
            P.fused_part[((size_t)blockIdx.x * kMaxFused + f) * kAccWords + k] = x;
        } else {
            // The workgroup's words straight into the probe's totals (set to
            // their identities before the launch): the sum words are partial
            // words that add without overflow, MIN/MAX are signed.
            unsigned long long* t = P.fused_acc + (size_t)f * kAccWords + k;
            if (k < 4) {
                if (x) atomicAdd(t, x);
            } else if (k == 4) {
                if ((int64_t)x != 0x7fffffffffffffffll) atomicMin((long long*)t, (long long)x);
            } else {
                if (x != 0x8000000000000000ull) atomicMax((long long*)t, (long long)x);
            }
        }
    }
}


// Copies `words` 32-bit words from global memory (16-byte aligned) to LDS:
// each thread's 16-byte loads of a pass are in flight together, then
// stored (a load-store loop would wait one memory round trip per word).
__device__ __forceinline__ void fill_lds_words(uint32_t* dst, const uint32_t* src, uint64_t words)
{
    constexpr int kU = 10;  // 10 x 16 B x 1024 threads = 160 KiB in one pass
    const uint64_t v4 = words / 4;
    const uint4* s4 = (const uint4*)src;
    uint4* d4 = (uint4*)dst;
    for (uint64_t base = 0; base < v4; base += (uint64_t)kU * blockDim.x) {
        uint4 t[kU];
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const uint64_t i = base + (uint64_t)u * blockDim.x + threadIdx.x;
            t[u] = s4[i < v4 ? i : 0];
        }
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const uint64_t i = base + (uint64_t)u * blockDim.x + threadIdx.x;
            if (i < v4) d4[i] = t[u];
        }
    }
    for (uint64_t i = v4 * 4 + threadIdx.x; i < words; i += blockDim.x) dst[i] = src[i];
}


// Places a block tile's output rows: every wave passes its row count `run`;
// returns the wave's first output position. One atomic per block tile (the
// waves' counts are combined through LDS between two barriers); every
// thread of the block calls it once per tile.
__device__ __forceinline__ unsigned long long reserve_block_tile(unsigned long long* counter, uint32_t run, int wave, int lane)
{
    constexpr int nw = kProbeBlock / kWave;
    __shared__ unsigned int wave_run[nw];
    __shared__ unsigned long long block_base;
    if (lane == 0) wave_run[wave] = run;
    __syncthreads();
    if (wave == 0) {
        const unsigned int v = lane < nw ? wave_run[lane] : 0u;
        unsigned int incl = v;
#pragma unroll
        for (int o = 1; o < nw; o <<= 1) {
            const unsigned int y = (unsigned int)__shfl_up((int)incl, o, kWave);
            if (lane >= o) incl += y;
        }
        if (lane < nw) wave_run[lane] = incl - v;
        const unsigned int total = (unsigned int)__shfl((int)incl, nw - 1, kWave);
        if (lane == 0) block_base = total ? atomicAdd(counter, (unsigned long long)total) : 0ull;
    }
    __syncthreads();
    return block_base + wave_run[wave];
}

// Dense direct table probe (distinct one-word integer keys), one pass. A
// block takes tiles of kProbeBlock / kWave wave tiles; per wave tile, with
// kTileRows rows per lane loaded as independent rounds: the keys; the
// bitmap words (LDS for the keys it covers; beyond it the ordinal table is
// read directly, kAbsent = no build row); the ordinals of the matches and
// the fused aggregate's probe column. The block's output rows are placed
// with one atomic per tile; within a wave tile they are in probe row order
// (row i = tile start + r * 64 + lane, r-major). A key matches at most one
// build row, so a row's output is at most one pair (or its unmatched pair in
// an outer join).
//
// probe_dense_tiles runs the block tiles starting at b_begin, b_begin +
// b_step, ... below b_end (every thread of the block calls it with the same
// range) and adds the block's fused words to the totals at the end; lbits
// holds the first lds_range bits of the membership bitmap in LDS (with
// lds_range 0 every key reads the ordinal table, and lbits must still point
// at one readable LDS word).
template <int kMode>
__device__ __forceinline__ void probe_dense_tiles(const ProbeArgs& P, const uint32_t* lbits, uint64_t lds_range, int64_t b_begin,
    int64_t b_end, int64_t b_step)
{
    constexpr bool kDirect = kMode != kGeneric;
    const int lane = lane_id();
    const int wave = threadIdx.x / kWave;
    const uint64_t lt_mask = (1ull << lane) - 1ull;
    constexpr int kAcc = kMode == kDirectVal ? 1 : kMaxFused;
    FusedLane acc[kAcc];
#pragma unroll
    for (int f = 0; f < kAcc; ++f) {
        const int kind = f < P.nfused ? P.fused[f].kind : FK_COUNT_PAIRS;
        acc[f].count = 0;
        acc[f].lo = kind == FK_MIN ? 0x7fffffffffffffffull : kind == FK_MAX ? 0x8000000000000000ull : 0ull;
        acc[f].hi = 0;
    }
    unsigned long long v_pairs = 0;
    uint64_t v_lo = 0, v_hi = 0;
    int64_t v_min = 0x7fffffffffffffffll, v_max = (int64_t)0x8000000000000000ll;
    // Every thread of the block runs the same number of block tiles.
    for (int64_t b0 = b_begin; b0 < b_end; b0 += b_step) {
        const int64_t t0 = b0 + (int64_t)wave * kWaveTile;
        // Round 1: keys and the fused aggregates' probe column, unconditional
        // loads at clamped rows (the column is streamed with the keys: a
        // later round would wait for it after the ordinals).
        uint64_t key[kTileRows];
        int64_t val[kTileRows];
        uint32_t part = 0, live = 0;
#pragma unroll
        for (int r = 0; r < kTileRows; ++r) {
            const int64_t i = t0 + (int64_t)r * kWave + lane;
            const bool in = i < P.n;
            const int64_t ic = in ? i : 0;
            const uint8_t sb = P.sel ? P.sel[ic] : (uint8_t)1;
            const bool sel = in & (sb != 0);
            uint64_t k1[1] = {0};
            const bool kv = probe_row_key<1, kDirect>(P, ic, k1);
            key[r] = k1[0];
            val[r] = 0;
            if (kMode == kDirectVal) val[r] = __builtin_nontemporal_load((const int64_t*)P.cols.c[P.val_slot].data + ic);
            part |= sel ? 1u << r : 0u;
            live |= (sel & kv) ? 1u << r : 0u;
        }
        // Round 2: the LDS bitmap words. The rounds are kept apart: memory
        // counters are waited on in issue order.
        uint32_t ixs[kTileRows], wl[kTileRows];
        uint32_t inr = 0, cov = 0;
#pragma unroll
        for (int r = 0; r < kTileRows; ++r) {
            const uint64_t idx = key[r] - P.dmin;
            const bool in_range = ((live >> r) & 1u) & (idx < P.drange);
            const uint32_t ix = in_range ? (uint32_t)idx : 0u;
            const bool covered = ix < lds_range;
            ixs[r] = ix;
            inr |= in_range ? 1u << r : 0u;
            cov |= covered ? 1u << r : 0u;
            wl[r] = lbits[covered ? (ix >> 5) : 0u];
        }
        __builtin_amdgcn_sched_barrier(0);
        uint32_t need = 0;
#pragma unroll
        for (int r = 0; r < kTileRows; ++r) {
            const bool covered = (cov >> r) & 1u;
            const bool bit = ((wl[r] >> (ixs[r] & 31u)) & 1u) != 0u;
            need |= (((inr >> r) & 1u) & (!covered | bit)) ? 1u << r : 0u;
        }
        // Round 3: the ordinals.
        uint32_t head[kTileRows];
        uint32_t hit = need;
#pragma unroll
        for (int r = 0; r < kTileRows; ++r) head[r] = P.dtab[((need >> r) & 1u) ? ixs[r] : 0u];
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int r = 0; r < kTileRows; ++r) {
            if (head[r] == kAbsent) hit &= ~(1u << r);
        }
        const uint32_t emit = hit | (P.emit_unmatched ? part : 0u);
        uint32_t run = 0;
#pragma unroll
        for (int r = 0; r < kTileRows; ++r) run += (uint32_t)__popcll(__ballot((emit >> r) & 1u));
        unsigned long long pos0 = reserve_block_tile(P.pair_count, run, wave, lane);
        // Pairs in probe row order and the fused aggregate.
#pragma unroll
        for (int r = 0; r < kTileRows; ++r) {
            const uint64_t m = __ballot((emit >> r) & 1u);
            const unsigned long long pos = pos0 + (uint32_t)__popcll(m & lt_mask);
            pos0 += (uint32_t)__popcll(m);
            if (!((emit >> r) & 1u)) continue;
            const int64_t i = t0 + (int64_t)r * kWave + lane;
            const int64_t stream_id = (int64_t)((uint64_t)-1 << 32) + i;
            const bool h = (hit >> r) & 1u;
            const int64_t ord = h ? (int64_t)head[r] : -1;
            if (pos < P.pair_cap) {
                const ll2 v = P.probe_is_left ? ll2{stream_id, ord} : ll2{ord, stream_id};
                __builtin_nontemporal_store(v, (ll2*)(P.pairs + 2 * pos));
            }
            if (kMode == kDirectVal) {
                v_pairs += 1;
                const uint64_t x = (uint64_t)val[r];
                const uint64_t lo = v_lo + x;
                v_hi += (lo < v_lo ? 1ull : 0ull) + (val[r] < 0 ? ~0ull : 0ull);
                v_lo = lo;
                v_min = val[r] < v_min ? val[r] : v_min;
                v_max = val[r] > v_max ? val[r] : v_max;
                continue;
            }
#pragma unroll
            for (int f = 0; f < kAcc; ++f) {
                if (f >= P.nfused) continue;
                const FusedDesc d = P.fused[f];
                if (d.kind == FK_COUNT_PAIRS) {
                    acc[f].count += 1;
                } else if (!d.from_build) {
                    const DevCol& col = P.cols.c[d.slot];
                    if (col_valid(col, i)) fused_add(acc[f], d.kind, 1u, col_int(col, i));
                } else if (h) {
                    const int64_t b = (int64_t)P.ord_to_arena[head[r]];
                    if (P.payload_valid[f] && !P.payload_valid[f][b]) continue;
                    const char* pd = P.payload_data[f];
                    i128 x;
                    switch (P.payload_width[f]) {
                    case 1: x = ((const int8_t*)pd)[b]; break;
                    case 2: x = ((const int16_t*)pd)[b]; break;
                    case 4: x = ((const int32_t*)pd)[b]; break;
                    case 8: x = ((const int64_t*)pd)[b]; break;
                    default: {
                        const uint64_t* q = (const uint64_t*)(pd + 16 * (size_t)b);
                        x = (i128)(((u128)q[1] << 64) | (u128)q[0]);
                    }
                    }
                    fused_add(acc[f], d.kind, 1u, x);
                }
            }
        }
    }
    if (P.nfused) fused_flush<kAcc, kMode == kDirectVal>(P, acc, v_pairs, v_lo, v_hi, v_min, v_max, lane);
}

constexpr int64_t kProbeBlockTile = (int64_t)(kProbeBlock / kWave) * kWaveTile;

template <int kMode>
__global__ void __launch_bounds__(kProbeBlock) k_probe_dense(ProbeArgs P)
{
    extern __shared__ __align__(16) uint64_t lds64[];
    fill_lds_words((uint32_t*)lds64, P.dbits, P.lds_words);
    __syncthreads();
    probe_dense_tiles<kMode>(P, (const uint32_t*)lds64, (uint64_t)P.lds_words * 32u, (int64_t)blockIdx.x * kProbeBlockTile, P.n,
        (int64_t)gridDim.x * kProbeBlockTile);
}

__device__ __forceinline__ uint64_t uniform_u64(uint64_t x)
{
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)x);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(x >> 32));
    return ((uint64_t)hi << 32) | lo;
}

// The dense probe's common case, pipelined: a non-nullable 8-byte integer
// key, no probe selection, an inner join, and fused aggregates that count
// pairs or read one non-nullable 8-byte probe column (kVal); MIN/MAX words
// are kept only when a fused aggregate asks (kMinMax); kLeft: the probe
// side is the left one. A wave tile is kLeanRows rows per lane. While a
// tile's ordinals are gathered and its pairs placed and stored, the keys
// (and probe column) of the wave's next tile are already loading: they are
// issued right after the gathers, and memory returns in issue order, so
// the wait for the gathers does not wait for them. A tile loads through
// one address per lane with immediate offsets (the batch holds at least
// one wave tile: a tile that would run past the end is moved back to end
// at the last row, its rows before the tile start left out); pairs are
// stored at 32-bit offsets from the wave's output base; the pair count is
// per wave. Output order within a wave tile: row-major over r, as in
// k_probe_dense.
constexpr int kLeanRows = 8;
constexpr int kLeanWaveTile = kWave * kLeanRows;
constexpr uint64_t kLeanBlockTile = (uint64_t)(kProbeBlock / kWave) * kLeanWaveTile;

struct LeanState {
    unsigned long long pairs;   // this wave's pairs
    // SUM of the probe column: the low 32 bits of each value (unsigned) and
    // the high 32 bits (signed) summed apart, two adds per row without a
    // carry; exact for fewer than 2^32 rows per lane.
    uint64_t lo, hi;
    int64_t mn, mx;
};

// The start of the wave tile at t0, moved back to end at the last row when
// it would run past the end.
__device__ __forceinline__ uint64_t lean_base(uint64_t t0, uint64_t n)
{
    return t0 + kLeanWaveTile <= n ? t0 : n - kLeanWaveTile;
}

template <bool kVal>
__device__ __forceinline__ void lean_load(const ProbeArgs& P, uint64_t tl, int lane, uint64_t (&key)[kLeanRows],
    int64_t (&val)[kLeanRows])
{
    const uint64_t* kp = (const uint64_t*)P.cols.c[0].data + tl + lane;
    const int64_t* vp = kVal ? (const int64_t*)P.cols.c[P.val_slot].data + tl + lane : nullptr;
#pragma unroll
    for (int r = 0; r < kLeanRows; ++r) {
        key[r] = __builtin_nontemporal_load(kp + r * kWave);
        val[r] = kVal ? __builtin_nontemporal_load(vp + r * kWave) : 0;
    }
}

// One wave tile (rows tl + r * 64 + lane; those before t0 left out) whose
// keys and probe column are in key/val; loads the tile at ntl into
// nkey/nval on the way.
// The ordinal of dense index ix from the packed table: w bits at bit
// ix * w, read as two adjacent words (the table has one word of slack).
__device__ __forceinline__ uint32_t packed_ordinal(const uint32_t* pack, uint32_t w, uint32_t ix)
{
    const uint64_t bit = (uint64_t)ix * w;
    const uint64_t wi = bit >> 5;
    const uint64_t two = (uint64_t)pack[wi] | ((uint64_t)pack[wi + 1] << 32);
    const uint64_t mask = (1ull << w) - 1ull;
    const uint64_t v = (two >> (bit & 31u)) & mask;
    return v == mask ? kAbsent : (uint32_t)v;
}

template <bool kVal, bool kMinMax, bool kLeft, bool kPacked>
__device__ __forceinline__ void lean_tile(const ProbeArgs& P, const uint32_t* lbits, uint32_t lds_range, int wave, int lane,
    uint64_t t0, uint64_t tl, const uint64_t (&key)[kLeanRows], const int64_t (&val)[kLeanRows], uint64_t ntl,
    uint64_t (&nkey)[kLeanRows], int64_t (&nval)[kLeanRows], LeanState& st)
{
    const uint64_t n = (uint64_t)P.n;
    const uint32_t skip = (uint32_t)(t0 < n ? t0 - tl : kLeanWaveTile);
    // Round 2: the LDS bitmap words.
    uint32_t ixs[kLeanRows], wl[kLeanRows];
    bool inr[kLeanRows], cov[kLeanRows];
#pragma unroll
    for (int r = 0; r < kLeanRows; ++r) {
        const uint64_t idx = key[r] - P.dmin;
        inr[r] = (lane >= (int)skip - r * kWave) & (idx < P.drange);
        const uint32_t ix = inr[r] ? (uint32_t)idx : 0u;
        cov[r] = ix < lds_range;
        ixs[r] = ix;
        wl[r] = lbits[cov[r] ? (ix >> 5) : 0u];
    }
    __builtin_amdgcn_sched_barrier(0);
    // Round 3: the ordinals of the rows that may match; then the next
    // tile's loads.
    uint32_t head[kLeanRows];
    bool need[kLeanRows];
#pragma unroll
    for (int r = 0; r < kLeanRows; ++r) {
        const bool bit = ((wl[r] >> (ixs[r] & 31u)) & 1u) != 0u;
        need[r] = inr[r] & (!cov[r] | bit);
        // Only the lanes that may match issue the gather.
        head[r] = kAbsent;
        if (need[r]) head[r] = kPacked ? packed_ordinal(P.dpack, P.dpack_w, ixs[r]) : P.dtab[ixs[r]];
    }
    __builtin_amdgcn_sched_barrier(0);
    lean_load<kVal>(P, ntl, lane, nkey, nval);
    __builtin_amdgcn_sched_barrier(0);
    bool hit[kLeanRows];
#pragma unroll
    for (int r = 0; r < kLeanRows; ++r) hit[r] = need[r] & (head[r] != kAbsent);
    if (kVal) {
#pragma unroll
        for (int r = 0; r < kLeanRows; ++r) {
            const int64_t x = hit[r] ? val[r] : 0;
            st.lo += (uint64_t)(uint32_t)x;
            st.hi += (uint64_t)(x >> 32);
            if (kMinMax) {
                st.mn = hit[r] && x < st.mn ? x : st.mn;
                st.mx = hit[r] && x > st.mx ? x : st.mx;
            }
        }
    }
    uint64_t m[kLeanRows];
    uint32_t run = 0;
#pragma unroll
    for (int r = 0; r < kLeanRows; ++r) {
        m[r] = __ballot(hit[r]);
        run += (uint32_t)__popcll(m[r]);
    }
    const uint64_t pos0 = uniform_u64(reserve_block_tile(P.pair_count, run, wave, lane));
    st.pairs += run;
    // Pairs in probe row order.
    ll2* out = (ll2*)(P.pairs + 2 * pos0);
    const bool fits = pos0 + run <= P.pair_cap;
    const int64_t sid0 = (int64_t)(tl - (1ull << 32)) + lane;  // stream id of row tl + lane
    const uint64_t below = (1ull << lane) - 1ull;
    uint32_t at = 0;
#pragma unroll
    for (int r = 0; r < kLeanRows; ++r) {
        if (hit[r]) {
            const uint32_t o = at + (uint32_t)__popcll(m[r] & below);
            if (fits || pos0 + o < P.pair_cap) {
                const int64_t sid = sid0 + (int64_t)r * kWave;
                const int64_t ord = (int64_t)head[r];
                const ll2 v = kLeft ? ll2{sid, ord} : ll2{ord, sid};
                __builtin_nontemporal_store(v, out + o);
            }
        }
        at += (uint32_t)__popcll(m[r]);
    }
}

template <bool kVal, bool kMinMax, bool kLeft, bool kPacked>
__global__ void __launch_bounds__(kProbeBlock) k_probe_dense_lean(ProbeArgs P)
{
    extern __shared__ __align__(16) uint64_t lds64[];
    const uint32_t* lbits = (const uint32_t*)lds64;
    {
        uint32_t* b = (uint32_t*)lds64;
        fill_lds_words(b, P.dbits, P.lds_words);
    }
    __syncthreads();
    const uint32_t lds_range = P.lds_words * 32u;
    const int lane = lane_id();
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / kWave));
    const uint64_t n = (uint64_t)P.n;
    const uint64_t stride = (uint64_t)gridDim.x * kLeanBlockTile;
    LeanState st;
    st.pairs = 0;
    st.lo = 0;
    st.hi = 0;
    st.mn = 0x7fffffffffffffffll;
    st.mx = (int64_t)0x8000000000000000ll;
    uint64_t ka[kLeanRows], kb[kLeanRows];
    int64_t va[kLeanRows], vb[kLeanRows];
    // Every wave of the block runs the same number of tiles (b0 is the
    // block's); two per trip, the buffers alternating.
    uint64_t b0 = blockIdx.x * kLeanBlockTile;
    const uint64_t woff = (uint64_t)wave * kLeanWaveTile;
    lean_load<kVal>(P, lean_base(b0 + woff, n), lane, ka, va);
    while (b0 < n) {
        const uint64_t t0 = b0 + woff, t1 = b0 + stride + woff, t2 = b0 + 2 * stride + woff;
        lean_tile<kVal, kMinMax, kLeft, kPacked>(P, lbits, lds_range, wave, lane, t0, lean_base(t0, n), ka, va, lean_base(t1, n), kb, vb,
            st);
        b0 += stride;
        if (b0 >= n) break;
        lean_tile<kVal, kMinMax, kLeft, kPacked>(P, lbits, lds_range, wave, lane, t1, lean_base(t1, n), kb, vb, lean_base(t2, n), ka, va,
            st);
        b0 += stride;
    }
    if (P.nfused) {
        FusedLane none[1];
        none[0].count = 0;
        none[0].lo = 0;
        none[0].hi = 0;
        // The 128-bit sum: lo + hi * 2^32.
        const uint64_t sum_lo = st.lo + (st.hi << 32);
        const uint64_t sum_hi = (uint64_t)((int64_t)st.hi >> 32) + (sum_lo < st.lo ? 1ull : 0ull);
        fused_flush<1, true>(P, none, lane == 0 ? st.pairs : 0ull, sum_lo, sum_hi, st.mn, st.mx, lane);
    }
}

// probe_hash_tiles runs the block tiles b_begin, b_begin + b_step, ...
// below b_end over the slot table T (in LDS for kTabLds), behind the Bloom
// filter words in LDS at lds_bloom when P.bloom_words is not 0, and adds the
// block's fused words to the totals at the end.
template <int KW, int kTab, int kMode>
__device__ __forceinline__ void probe_hash_tiles(const ProbeArgs& P, const Slot<KW>* T, const uint32_t* lds_bloom, int64_t b_begin,
    int64_t b_end, int64_t b_step)
{
    constexpr bool kDirect = kMode != kGeneric;
    constexpr bool kLds = kTab == kTabLds;
    const int lane = lane_id();
    const uint64_t lt_mask = (1ull << lane) - 1ull;
    // Fused state per lane. kDirectVal: the pair count and the sum, min and
    // max of the one probe column, shared by every fused aggregate; else one
    // state per fused aggregate.
    constexpr int kAcc = kMode == kDirectVal ? 1 : kMaxFused;
    FusedLane acc[kAcc];
#pragma unroll
    for (int f = 0; f < kAcc; ++f) {
        const int kind = f < P.nfused ? P.fused[f].kind : FK_COUNT_PAIRS;
        acc[f].count = 0;
        acc[f].lo = kind == FK_MIN ? 0x7fffffffffffffffull : kind == FK_MAX ? 0x8000000000000000ull : 0ull;
        acc[f].hi = 0;
    }
    unsigned long long v_pairs = 0;
    uint64_t v_lo = 0, v_hi = 0;
    int64_t v_min = 0x7fffffffffffffffll, v_max = (int64_t)0x8000000000000000ll;
    const int wave = threadIdx.x / kWave;
    // Every thread of the block runs the same number of block tiles (two
    // barriers per tile place the block's pairs with one atomic).
    for (int64_t b0 = b_begin; b0 < b_end; b0 += b_step) {
        const int64_t t0 = b0 + (int64_t)wave * kWaveTile;
        // Per row of the tile: match count and slot head; and whether the
        // row is selected.
        uint32_t cnt[kTileRows], head[kTileRows];
        uint32_t part = 0;
#pragma unroll
        for (int q = 0; q < kSubRounds; ++q) {
            uint64_t key[kSub][KW];
            uint64_t s[kSub];
            int64_t val[kSub];
            bool live[kSub];
            // Every load below is unconditional, from a clamped index: a
            // load under a branch that merges with a default value makes
            // the compiler wait for it inside the branch, one row at a time.
            // Round 1: keys.
#pragma unroll
            for (int u = 0; u < kSub; ++u) {
                const int r = q * kSub + u;
                const int64_t i = t0 + (int64_t)r * kWave + lane;
                const bool in = i < P.n;
                const int64_t ic = in ? i : 0;
                const uint8_t sb = P.sel ? P.sel[ic] : (uint8_t)1;
                const bool sel = in & (sb != 0);
                part |= sel ? 1u << r : 0u;
#pragma unroll
                for (int w = 0; w < KW; ++w) key[u][w] = 0;
                const bool kv = probe_row_key<KW, kDirect>(P, ic, key[u]);
                live[u] = sel & kv;
                val[u] = 0;
                if (kMode == kDirectVal) val[u] = __builtin_nontemporal_load((const int64_t*)P.cols.c[P.val_slot].data + ic);
            }
            // Round 2: Bloom filter and the first slot (dense: the bitmap
            // and the ordinal).
            Slot<KW> sl[kSub];
            {
#pragma unroll
                for (int u = 0; u < kSub; ++u) {
                    const uint64_t h = key_place_hash<KW>(key[u]);
                    s[u] = h & P.cmask;
                    bool maybe = live[u];
                    if (!kLds && P.bloom_words) {
                        const uint32_t bits = bloom_bits(h);
                        const uint32_t bw = lds_bloom[bloom_word(h, P.bloom_words - 1u)];
                        maybe = maybe & ((bw & bits) == bits);
                    }
                    sl[u] = T[maybe ? s[u] : 0];
                    if (!maybe) sl[u].count = 0;
                }
            }
            // Round 3: compare (the walk continues only on a key miss).
#pragma unroll
            for (int u = 0; u < kSub; ++u) {
                const int r = q * kSub + u;
                head[r] = 0;
                cnt[r] = lookup_from<KW>(P, T, key[u], s[u], sl[u], &head[r]);
                // An unmatched row of an outer join is one output row too
                // (its build side NULL).
                const uint32_t times = cnt[r] ? cnt[r] : (((part >> r) & 1u) && P.emit_unmatched ? 1u : 0u);
                if (kMode == kDirectVal && times) {
                    v_pairs += times;
                    const u128 sum = ((((u128)v_hi) << 64) | (u128)v_lo) + (u128)(i128)val[u] * (u128)times;
                    v_lo = (uint64_t)sum;
                    v_hi = (uint64_t)(sum >> 64);
                    v_min = val[u] < v_min ? val[u] : v_min;
                    v_max = val[u] > v_max ? val[u] : v_max;
                }
            }
        }
        // Output positions, row-major across the wave: the tile's total
        // first, one atomic, then each row's offset.
        auto emits = [&](int r) -> uint32_t {
            return cnt[r] + (((part >> r) & 1u) && cnt[r] == 0u && P.emit_unmatched ? 1u : 0u);
        };
        uint32_t run = 0;
#pragma unroll
        for (int r = 0; r < kTileRows; ++r) {
            const uint32_t e = emits(r);
            if (__all(e <= 1u)) {
                run += (uint32_t)__popcll(__ballot(e != 0u));
            } else {
                uint32_t x = e;
#pragma unroll
                for (int o = 1; o < kWave; o <<= 1) x += (uint32_t)__shfl_xor((int)x, o, kWave);
                run += x;
            }
        }
        unsigned long long pos0 = reserve_block_tile(P.pair_count, run, wave, lane);
#pragma unroll
        for (int r = 0; r < kTileRows; ++r) {
            const uint32_t e = emits(r);
            unsigned long long pos;
            if (__all(e <= 1u)) {
                const uint64_t m = __ballot(e != 0u);
                pos = pos0 + (uint32_t)__popcll(m & lt_mask);
                pos0 += (uint32_t)__popcll(m);
            } else {
                uint32_t incl = e;
#pragma unroll
                for (int o = 1; o < kWave; o <<= 1) {
                    const uint32_t y = (uint32_t)__shfl_up((int)incl, o, kWave);
                    if (lane >= o) incl += y;
                }
                pos = pos0 + incl - e;
                pos0 += (uint32_t)__shfl((int)incl, kWave - 1, kWave);
            }
            if (e == 0u) continue;
            const int64_t i = t0 + (int64_t)r * kWave + lane;
            const int64_t stream_id = (int64_t)((uint64_t)-1 << 32) + i;
            const uint32_t c = cnt[r];
            if (c == 0u && pos < P.pair_cap) {
                const ll2 v = P.probe_is_left ? ll2{stream_id, -1} : ll2{-1, stream_id};
                __builtin_nontemporal_store(v, (ll2*)(P.pairs + 2 * pos));
            }
            for (uint32_t k = 0; k < c; ++k, ++pos) {
                const int64_t ord = c == 1u ? (int64_t)head[r] : P.ordinal[head[r] + k];
                if (pos < P.pair_cap) {
                    const ll2 v = P.probe_is_left ? ll2{stream_id, ord} : ll2{ord, stream_id};
                    __builtin_nontemporal_store(v, (ll2*)(P.pairs + 2 * pos));
                }
            }
            // Fused aggregates of this row's e output rows (kDirectVal:
            // applied above); an unmatched row has a NULL build side.
#pragma unroll
            for (int f = 0; f < kAcc; ++f) {
                if (kMode == kDirectVal || f >= P.nfused) continue;
                const FusedDesc d = P.fused[f];
                if (d.kind == FK_COUNT_PAIRS) {
                    acc[f].count += e;
                } else if (!d.from_build) {
                    const DevCol& col = P.cols.c[d.slot];
                    if (col_valid(col, i)) fused_add(acc[f], d.kind, e, col_int(col, i));
                } else {
                    const uint8_t* vv = P.payload_valid[f];
                    const char* pd = P.payload_data[f];
                    for (uint32_t k = 0; k < c; ++k) {
                        const int64_t b = c == 1u ? (int64_t)P.ord_to_arena[head[r]] : (int64_t)(head[r] + k);
                        if (vv && !vv[b]) continue;
                        i128 x;
                        switch (P.payload_width[f]) {
                        case 1: x = ((const int8_t*)pd)[b]; break;
                        case 2: x = ((const int16_t*)pd)[b]; break;
                        case 4: x = ((const int32_t*)pd)[b]; break;
                        case 8: x = ((const int64_t*)pd)[b]; break;
                        default: {
                            const uint64_t* q = (const uint64_t*)(pd + 16 * (size_t)b);
                            x = (i128)(((u128)q[1] << 64) | (u128)q[0]);
                        }
                        }
                        fused_add(acc[f], d.kind, 1u, x);
                    }
                }
            }
        }
    }
    if (P.nfused) fused_flush<kAcc, kMode == kDirectVal>(P, acc, v_pairs, v_lo, v_hi, v_min, v_max, lane);
}

template <int KW, int kTab, int kMode>
__global__ void __launch_bounds__(kProbeBlock) k_probe(ProbeArgs P)
{
    constexpr bool kLds = kTab == kTabLds;
    extern __shared__ __align__(16) uint64_t lds64[];
    if (kLds) {
        const uint64_t words = (P.cmask + 1) * (sizeof(Slot<KW>) / 8);
        fill_lds_words((uint32_t*)lds64, (const uint32_t*)P.table, words * 2);
    } else if (P.bloom_words) {
        fill_lds_words((uint32_t*)lds64, P.bloom, P.bloom_words);
    }
    __syncthreads();
    const Slot<KW>* T = kLds ? (const Slot<KW>*)lds64 : (const Slot<KW>*)P.table;
    probe_hash_tiles<KW, kTab, kMode>(P, T, (const uint32_t*)lds64, (int64_t)blockIdx.x * kProbeBlockTile, P.n,
        (int64_t)gridDim.x * kProbeBlockTile);
}

// Batched probe: many joins' probes (of one compiled join, one-word keys) in
// ONE persistent launch. Each workgroup takes items (instance, block tile
// range) from a device queue and probes that range of the instance's batch
// with its table: the dense table reading the ordinal table for every key
// (no bitmap in LDS), or the hash table without the Bloom filter (nothing is
// staged into LDS per item). Pairs and fused words go to each instance's own
// pair buffer and counters. Every workgroup leaves the loop when the queue
// is drained.
struct ProbeItem {
    int32_t inst;
    int32_t pad;
    int64_t b_begin;   // first row of the item (a multiple of the block tile)
    int64_t b_end;     // one past its last block tile's first row
};

// ctl[0] = instances, ctl[1] = items (device words, so that a captured graph
// serves any count up to its buffers').
template <int kMode>
__global__ void __launch_bounds__(kProbeBlock) k_probe_batched(const ProbeArgs* __restrict__ insts,
    const ProbeItem* __restrict__ items, const int* __restrict__ ctl, unsigned int* head)
{
    const int nitems = ctl[1];
    __shared__ int s_item;
    __shared__ uint32_t no_bits[4];
    for (;;) {
        if (threadIdx.x == 0) s_item = (int)atomicAdd(head, 1u);
        __syncthreads();
        const int it = s_item;
        __syncthreads();
        if (it >= nitems) break;
        const ProbeItem w = items[it];
        const ProbeArgs& P = insts[w.inst];
        if (P.dtab) probe_dense_tiles<kMode>(P, no_bits, 0, w.b_begin, w.b_end, kProbeBlockTile);
        else probe_hash_tiles<1, kTabHash, kMode>(P, (const Slot<1>*)P.table, no_bits, w.b_begin, w.b_end, kProbeBlockTile);
    }
}

// Every instance's pair counter and fused words at their identities.
__global__ void k_probe_init_batched(const ProbeArgs* __restrict__ insts, const int* __restrict__ ctl, int words)
{
    const int n = ctl[0];
    for (int i = blockIdx.x; i < n; i += gridDim.x) probe_counters_init(insts[i].pair_count, insts[i].fused_acc, words, (int)threadIdx.x);
}

// Every instance's pair count and fused words into out[i * stride ...].
__global__ void k_probe_collect_batched(const ProbeArgs* __restrict__ insts, const int* __restrict__ ctl, int words,
    unsigned long long* out)
{
    const int n = ctl[0];
    const int stride = 1 + words;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < (int64_t)n * stride; t += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(t / stride), k = (int)(t % stride);
        out[t] = k == 0 ? insts[i].pair_count[0] : insts[i].fused_acc[k - 1];
    }
}

// Order-independent checksum of n pairs: the sum and the xor of a 64-bit
// mix of every pair (out[0], out[1], added to), so that two pair sets can be
// compared without moving them off the device.
__device__ __forceinline__ uint64_t pair_mix(uint64_t x)
{
    x ^= x >> 30;
    x *= 0xbf58476d1ce4e5b9ull;
    x ^= x >> 27;
    x *= 0x94d049bb133111ebull;
    x ^= x >> 31;
    return x;
}

__global__ void __launch_bounds__(kBlock) k_pairs_checksum(const int64_t* pairs, int64_t n, unsigned long long* out)
{
    unsigned long long sum = 0, x = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t h = pair_mix((uint64_t)pairs[2 * i] ^ pair_mix((uint64_t)pairs[2 * i + 1] + 0x9e3779b97f4a7c15ull));
        sum += h;
        x ^= h;
    }
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        sum += shfl_xor_u64(sum, o);
        x ^= shfl_xor_u64(x, o);
    }
    if (lane_id() == 0) {
        atomicAdd(&out[0], sum);
        atomicXor(&out[1], x);
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

// A captured probe pipeline (the detached batched probe): device and pinned
// buffers of fixed capacity and the graph that reads them -- reset the
// queue, set every instance's counters, the persistent probe, collect the
// counts and fused words, copy them to the host. A call writes the
// instances' descriptors into the buffers and replays the graph.
struct ProbePipeline {
    int mode = 0;
    int words = 0;
    static constexpr int kMaxInst = 256;
    static constexpr int kMaxItems = 16384;
    static constexpr int kStride = 2 + kMaxFused * kAccWords;   // counter words per instance
    char* d_buf = nullptr;       // [ctl 16 B | insts | items]
    char* h_stage = nullptr;     // pinned copy of d_buf's layout
    unsigned int* d_head = nullptr;
    unsigned long long* d_counters = nullptr;   // kMaxInst x kStride
    unsigned long long* d_out = nullptr;        // kMaxInst x (1 + words)
    unsigned long long* h_out = nullptr;        // pinned
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;

    static size_t insts_off() { return 16; }
    static size_t items_off() { return 16 + (size_t)kMaxInst * sizeof(ProbeArgs); }
    static size_t buf_bytes() { return items_off() + (size_t)kMaxItems * sizeof(ProbeItem); }
    ~ProbePipeline()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        dfree(d_buf);
        dfree(d_head);
        dfree(d_counters);
        dfree(d_out);
        hfree_pinned(h_stage);
        hfree_pinned(h_out);
    }
};

struct JoinPlan {
    ~JoinPlan()
    {
        for (auto& per_stream : pipes) {
            for (ProbePipeline*& pp : per_stream) {
                delete pp;
                pp = nullptr;
            }
        }
    }
    std::mutex pipe_mu;
    ProbePipeline* pipes[kMaxStreams][3] = {};   // per stream and kernel mode
    QdbRowJoinSpec spec;
    int KW = 1;
    JKeyDesc build_key;   // slots index build-side DevCols
    JKeyDesc probe_key;   // slots index probe-side DevCols
    int build_nslots = 0, build_slot_col[kMaxDevCols];
    int probe_nslots = 0, probe_slot_col[kMaxDevCols];
    int build_key_slots = 0;  // build slots used by keys (payload slots follow)
    bool dense_ok = false;    // one integer key of at most 8 bytes: the dense table may serve
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
    // probe structures (after Build): the slot table and the Bloom filter,
    // or the dense direct table
    void* table = nullptr;
    uint64_t capacity = 0;
    uint32_t* bloom = nullptr;
    bool dense = false;
    uint64_t dmin = 0, drange = 0;
    uint32_t* dbits = nullptr;
    uint32_t* dtab = nullptr;
    uint32_t* dpack = nullptr;    // dense: the ordinals at dpack_w bits (large ranges)
    uint32_t dpack_w = 0;
    uint32_t bloom_words = 0;
    int64_t keys_distinct = 0;
    int64_t max_key_rows = 0;   // most build rows of one key (a probe row's most matches)
    bool built = false;
    // The arena's row count stays on the device while batches are appended
    // (rows_dev); `rows` is exact once read back (rows_known), `rows_bound`
    // bounds it meanwhile (the arena is sized for it).
    unsigned long long* rows_dev = nullptr;
    int64_t rows_bound = 0;
    bool rows_known = true;
    // One-word keys: the smallest and largest build key, when read back.
    bool have_range = false;
    uint64_t kmin = 0, kmax = 0;
    // device: the pair counter (word 0) and then the fused aggregate words
    // of the last probe, read back together
    unsigned long long* pair_count = nullptr;
    unsigned long long* fused_acc = nullptr;    // pair_count + 2
    // host: that readback; fused_ready when it holds a probe's fused words
    // not yet added to the totals
    unsigned long long* readback = nullptr;     // pinned, kReadbackWords
    bool fused_ready = false;
    // the pair counter and the fused words hold their identities (set by
    // the build for the first probe)
    bool counters_ready = false;
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
    {
        const QdbValueType& bt = spec.Build.Types[spec.BuildKeyColumns[0]];
        P->dense_ok = spec.NumKeys == 1 && bt.Kind != QDB_KIND_FLOAT && bt.Kind != QDB_KIND_BOOL && key_bytes(bt) <= 8
            && spec.TableMode == QDB_JOIN_TABLE_AUTO;
    }
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
    dfree(J->bloom);
    dfree(J->dbits);
    dfree(J->dtab);
    dfree(J->dpack);
    dfree(J->pair_count);
    dfree(J->rows_dev);
    hfree_pinned(J->readback);
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
    // Sized for the bound on the rows (the exact count may still be on the
    // device); growing copies the bound's rows.
    const int64_t have = J->rows_bound;
    if (have + more_rows > J->cap) {
        int64_t cap = J->cap ? J->cap : 1024;
        while (cap < have + more_rows) cap *= 2;
        if (!grow_buf(&J->keys, have * P->KW, cap * P->KW, 8)) return false;
        if (!grow_buf(&J->ordinal, have, cap, 8)) return false;
        for (int p = 0; p < P->npayload; ++p) {
            const QdbValueType& t = P->spec.Build.Types[p];
            const size_t w = t.Kind == QDB_KIND_BOOL ? 1 : (size_t)t.Width;
            if (!grow_buf(&J->pdata[(size_t)p], have, cap, w)) return false;
            if (t.Nullable && !grow_buf(&J->pvalid[(size_t)p], have, cap, 1)) return false;
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
    // Per block: the kept rows (then the first arena row) and, for one-word
    // keys, the key range.
    const size_t range_off = ((size_t)nblocks * 4 + 15) / 16 * 16;
    char* work = (char*)dalloc(range_off + (KW == 1 ? (size_t)nblocks * 16 : 0) + 16);
    if (!work) return false;
    unsigned int* counts = (unsigned int*)work;
    unsigned long long* brange = KW == 1 ? (unsigned long long*)(work + range_off) : nullptr;
    hipLaunchKernelGGL((k_append_count<KW>), dim3(nblocks), dim3(kBlock), 0, stream(), cols, P->build_key, n, batch->Selection, counts,
        brange);
    // Each block's first arena row; the arena's row count and key range grow
    // here (the first append sets every arena word).
    hipLaunchKernelGGL(k_scan_counts, dim3(1), dim3(1024), 0, stream(), counts, (unsigned int)nblocks, J->rows_dev,
        (const unsigned long long*)brange, J->appended == 0 ? 1 : 0);
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
            batch->Selection, counts, J->appended, J->keys, J->ordinal, J->ord_to_arena, pcols, np, AC);
        ok = hip_ok(hipGetLastError(), "append write");
        if (P->npayload == 0) break;
    }
    // The kept rows are counted on the device; the host reads the count
    // back at Build (with the dense table's key range).
    dfree(work);
    if (!ok) return false;
    J->rows_bound += n;
    J->rows_known = false;
    J->appended += n;
    return true;
}

// The dense direct table when the keys are distinct and their range is
// small; false (and no error) when the hash table must serve instead.
// The arena's exact row count on the host (one readback when appends left
// it on the device).
bool read_rows(JoinBuild* J)
{
    if (J->rows_known) return true;
    unsigned long long* w = J->readback;
    J->fused_ready = false;
    if (!hip_ok(hipMemcpyAsync(w, J->rows_dev, kArenaWords * 8, hipMemcpyDeviceToHost, stream()), "arena rows") || !sync("arena rows")) {
        return false;
    }
    J->rows = (int64_t)w[0];
    J->rows_known = true;
    // One-word keys: the appends kept their range.
    if (J->plan->KW == 1 && J->rows > 0) {
        J->have_range = true;
        J->kmin = ~w[2];
        J->kmax = w[3];
    }
    return true;
}

void dense_pack(JoinBuild* J);

// Frees the dense table's buffers.
void dense_drop(JoinBuild* J)
{
    dfree(J->dbits);
    dfree(J->dtab);
    J->dbits = nullptr;
    J->dtab = nullptr;
}

// The widest key range the dense table takes for `rows` build rows.
__host__ __device__ inline uint64_t dense_limit(uint64_t rows)
{
    return rows * 8 > (1ull << 16) ? rows * 8 : (1ull << 16);
}

// The dense table's first half: its buffers (sized for the bound on the
// rows: the appends kept the key range on the device, the kernels read it
// there), its kernels, and the readback of the row count, the range and the
// flags (repeated key, range too wide), queued on the stream. False (nothing
// queued, no error) when the dense table cannot serve.
bool dense_build_launch(JoinBuild* J)
{
    const int64_t bound = J->rows_known ? J->rows : J->rows_bound;
    const uint64_t limit_cap = dense_limit((uint64_t)bound);
    if (limit_cap >= (1ull << 32)) return false;
    const size_t words_cap = (size_t)((limit_cap + 31) / 32);
    J->dbits = (uint32_t*)dalloc(words_cap * 4);
    J->dtab = (uint32_t*)dalloc((size_t)limit_cap * 4);
    if (!J->dbits || !J->dtab) {
        dense_drop(J);
        clear_error();
        return false;
    }
    {
        TimedRegion timer("join_build");
        hipLaunchKernelGGL(k_dense_prepare, dim3(grid_for((int64_t)limit_cap, kBlock, 8)), dim3(kBlock), 0, stream(), J->rows_dev,
            J->dbits, J->dtab);
        hipLaunchKernelGGL(k_dense_fill, dim3(grid_for(bound, kBlock, 8)), dim3(kBlock), 0, stream(), J->keys, J->ordinal,
            J->rows_dev, J->dbits, J->dtab, J->pair_count, J->fused_acc, J->plan->spec.NumFused * kAccWords);
    }
    J->fused_ready = false;
    if (!hip_ok(hipGetLastError(), "dense fill") ||
        !hip_ok(hipMemcpyAsync(J->readback, J->rows_dev, kArenaWords * 8, hipMemcpyDeviceToHost, stream()), "dense info")) {
        dense_drop(J);
        clear_error();
        return false;
    }
    return true;
}

// The dense table's second half, once the stream has synchronised (`ok`
// false: it failed): the readback decides. False (buffers dropped, no
// error) when the hash table must serve.
bool dense_build_finish(JoinBuild* J, bool ok)
{
    const unsigned long long* h = J->readback;
    if (ok) {
        J->rows = (int64_t)h[0];
        J->rows_known = true;
        J->have_range = J->rows > 0;
        J->kmin = ~h[2];
        J->kmax = h[3];
    }
    const unsigned int* h_flags = (const unsigned int*)(h + 1);
    const uint64_t range = J->kmax - J->kmin + 1;
    // Repeated keys or a key range too wide: the hash table serves.
    if (!ok || J->rows <= 0 || h_flags[0] || h_flags[1] || J->kmax < J->kmin || range == 0 ||
        range > dense_limit((uint64_t)J->rows)) {
        dense_drop(J);
        clear_error();
        return false;
    }
    J->dense = true;
    J->counters_ready = true;
    J->dmin = J->kmin;
    J->drange = range;
    J->keys_distinct = J->rows;
    J->max_key_rows = J->rows > 0 ? 1 : 0;
    J->built = true;
    dense_pack(J);
    return true;
}

// Ordinal tables larger than this are packed for the lean probe: at
// ceil(log2(appended rows + 1)) bits per key the table of the 1.5M-key
// range of 1.5M appended build rows takes 3.9 MB instead of 6 MB, within
// one XCD's 4 MB L2 (QDB_GPU_JOIN_PACK=0: never).
constexpr uint64_t kPackMinTableBytes = (uint64_t)2 << 20;

void dense_pack(JoinBuild* J)
{
    const char* e = getenv("QDB_GPU_JOIN_PACK");
    if (e && e[0] == '0') return;
    if (J->drange * 4 <= kPackMinTableBytes) return;
    // Ordinals are appended-row indices (rows a selection left out count
    // too): w bits hold every ordinal below J->appended and the all-ones
    // absent marker.
    uint32_t w = 1;
    while (w < 32 && ((uint64_t)1 << w) <= (uint64_t)J->appended) ++w;
    if (w >= 25) return;  // too little to gain
    const uint64_t nwords = (J->drange * w + 31) / 32 + 1;
    J->dpack = (uint32_t*)dalloc(nwords * 4);
    if (!J->dpack) {
        clear_error();
        return;
    }
    J->dpack_w = w;
    hipLaunchKernelGGL(k_dense_pack, dim3(grid_for((int64_t)nwords, kBlock, 4)), dim3(kBlock), 0, stream(), J->dtab, J->drange, w,
        J->dpack, nwords);
    if (hipGetLastError() != hipSuccess) {
        dfree(J->dpack);
        J->dpack = nullptr;
        J->dpack_w = 0;
    }
}

bool try_dense_build(JoinBuild* J)
{
    if (!dense_build_launch(J)) return false;
    return dense_build_finish(J, sync("dense fill"));
}

// Whether J's build tries the dense table first.
bool dense_wanted(const JoinBuild* J)
{
    const int64_t bound = J->rows_known ? J->rows : J->rows_bound;
    return J->plan->KW == 1 && J->plan->dense_ok && bound > 0 && bound < (int64_t)0x7fffffff;
}

// Builds J's table: the dense table when it serves (unless dense_tried:
// the dense attempt was already made), else the hash table.
template <int KW>
bool do_build(JoinBuild* J, bool dense_tried = false)
{
    JoinPlan* P = J->plan;
    if (J->appended > (int64_t)0xffffffff) return fail("join builds are limited to 2^32 appended rows");
    if (!dense_tried && dense_wanted(J) && try_dense_build(J)) return true;
    if (!read_rows(J)) return false;
    const int64_t rows = J->rows;
    if (rows >= (int64_t)0x7fffffff) return fail("join builds are limited to 2^31 rows");
    uint64_t capacity = 1024;
    int cap_bits = 10;
    while (capacity < (uint64_t)rows * 2) {
        capacity *= 2;
        ++cap_bits;
    }
    J->capacity = capacity;
    // Bloom filter: about 16 bits per row, between 8 KiB and 128 KiB.
    uint32_t words = 2048;
    while (words < 32768u && (size_t)words * 8 <= max_lds_per_block() && (uint64_t)words * 32 < (uint64_t)rows * 16) words *= 2;
    J->bloom_words = words;
    const size_t nrows = (size_t)(rows > 0 ? rows : 1);
    J->table = dalloc(capacity * sizeof(Slot<KW>));
    J->bloom = (uint32_t*)dalloc((size_t)words * 4);
    Slot<KW>* table = (Slot<KW>*)J->table;
    uint32_t* row_slot = (uint32_t*)dalloc(nrows * 4);
    uint32_t* counts = (uint32_t*)dalloc(capacity * 4);
    uint32_t* start = (uint32_t*)dalloc(capacity * 4);
    uint32_t* perm = (uint32_t*)dalloc(nrows * 4);
    // [0] unused, [1] duplicate flag, [2] distinct keys, [3] most rows of one key
    unsigned long long* stats = (unsigned long long*)dalloc(32);
    const int cgrid = grid_for((int64_t)capacity, kBlock, 8);
    unsigned long long* dpart = (unsigned long long*)dalloc((size_t)cgrid * 16);
    uint64_t* new_keys = (uint64_t*)dalloc(nrows * KW * 8);
    int64_t* new_ord = (int64_t*)dalloc(nrows * 8);
    uint32_t* sorted_slot = nullptr;
    uint32_t* iota = nullptr;
    auto release = [&]() {
        dfree(dpart);
        dfree(row_slot);
        dfree(counts);
        dfree(start);
        dfree(perm);
        dfree(stats);
        dfree(sorted_slot);
        dfree(iota);
    };
    if (!table || !J->bloom || !row_slot || !counts || !start || !perm || !stats || !dpart || !new_keys || !new_ord) {
        release();
        return false;
    }
    // One-word keys: their range (read back here when no dense attempt
    // read it), and an empty key word outside it.
    if (KW == 1 && !J->have_range && rows > 0) {
        const int mgrid = grid_for(rows, kBlock, 2);
        unsigned long long* mm = (unsigned long long*)dalloc(16);
        unsigned long long* mpart = (unsigned long long*)dalloc((size_t)mgrid * 16);
        unsigned long long h_mm[2] = {0, 0};
        bool got = mm && mpart;
        if (got) {
            hipLaunchKernelGGL(k_key_minmax, dim3(mgrid), dim3(kBlock), 0, stream(), J->keys, (const unsigned long long*)J->rows_dev,
                mpart);
            ReduceOps mops;
            memset(&mops, 0, sizeof(mops));
            mops.op[0] = 3;  // unsigned min
            mops.op[1] = 4;  // unsigned max
            hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(kBlock), 0, stream(), mpart, mgrid, 2, 2, mops, mm);
            got = hip_ok(hipGetLastError(), "key range") && hip_ok(hipMemcpyAsync(h_mm, mm, 16, hipMemcpyDeviceToHost, stream()), "key range")
                && sync("key range");
        }
        dfree(mm);
        dfree(mpart);
        if (got) {
            J->have_range = true;
            J->kmin = h_mm[0];
            J->kmax = h_mm[1];
        }
        clear_error();
    }
    int use_empty = 0;
    uint64_t empty = 0;
    if (KW == 1 && J->have_range) {
        if (J->kmax != ~0ull) {
            use_empty = 1;
            empty = J->kmax + 1;
        } else if (J->kmin != 0ull) {
            use_empty = 1;
            empty = J->kmin - 1;
        }
    }
    bool ok = zero_fill(J->bloom, (size_t)words * 4) && zero_fill(stats, 32);
    if (ok && use_empty) {
        hipLaunchKernelGGL(k_fill_slots, dim3(grid_for((int64_t)capacity, kBlock, 8)), dim3(kBlock), 0, stream(),
            (Slot<1>*)table, capacity, empty);
        ok = hip_ok(hipGetLastError(), "slot fill");
    } else if (ok) {
        ok = zero_fill(table, capacity * sizeof(Slot<KW>));
    }
    if (ok) {
        TimedRegion timer("join_build");
        if (rows > 0) {
            hipLaunchKernelGGL((k_build_insert<KW>), dim3(grid_for(rows, kBlock, 8)), dim3(kBlock), 0, stream(), table,
                capacity - 1, J->keys, rows, row_slot, (unsigned int*)(stats + 1), use_empty, empty);
        }
        hipLaunchKernelGGL((k_slot_counts<KW>), dim3(cgrid), dim3(kBlock), 0, stream(), table, capacity, counts, J->bloom,
            words - 1u, dpart);
        ReduceOps ops;
        memset(&ops, 0, sizeof(ops));
        ops.op[1] = 4;  // unsigned max
        hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(kBlock), 0, stream(), dpart, cgrid, 2, 2, ops, stats + 2);
        ok = hip_ok(hipGetLastError(), "build insert") && exclusive_scan_u32(counts, start, (int64_t)capacity);
    }
    unsigned long long h_stats[4] = {0, 0, 0, 0};
    ok = ok && hip_ok(hipGetLastError(), "build table")
        && hip_ok(hipMemcpyAsync(h_stats, stats, 32, hipMemcpyDeviceToHost, stream()), "build stats") && sync("build");
    if (!ok) {
        release();
        dfree(new_keys);
        dfree(new_ord);
        return false;
    }
    J->keys_distinct = (int64_t)h_stats[2];
    J->max_key_rows = (int64_t)h_stats[3];
    const bool dups = (h_stats[1] & 0xffffffffull) != 0;
    if (rows > 0) {
        TimedRegion timer("join_build_place");
        const int g = grid_for(rows, kBlock, 8);
        if (!dups) {
            hipLaunchKernelGGL(k_place_unique, dim3(g), dim3(kBlock), 0, stream(), row_slot, start, rows, perm);
        } else {
            // Duplicate keys: a stable sort by slot keeps each key's rows in
            // build order.
            sorted_slot = (uint32_t*)dalloc(nrows * 4);
            iota = (uint32_t*)dalloc(nrows * 4);
            ok = sorted_slot && iota;
            if (ok) {
                hipLaunchKernelGGL(k_iota, dim3(g), dim3(kBlock), 0, stream(), iota, rows);
                ok = sort_pairs_u32(row_slot, sorted_slot, iota, perm, rows, cap_bits);
            }
        }
        if (ok) {
            hipLaunchKernelGGL((k_build_csr<KW>), dim3(g), dim3(kBlock), 0, stream(), perm, rows, J->keys, J->ordinal,
                new_keys, new_ord, J->ord_to_arena);
            hipLaunchKernelGGL((k_finish_slots<KW>), dim3(grid_for((int64_t)capacity, kBlock, 8)), dim3(kBlock), 0, stream(),
                table, capacity, start, new_ord);
            // Payload columns in the same order.
            for (int p = 0; ok && p < P->npayload; ++p) {
                const QdbValueType& t = P->spec.Build.Types[p];
                const int w = t.Kind == QDB_KIND_BOOL ? 1 : t.Width;
                char* nd = (char*)dalloc((size_t)J->cap * (size_t)w);
                uint8_t* nv = t.Nullable ? (uint8_t*)dalloc((size_t)J->cap) : nullptr;
                if (!nd || (t.Nullable && !nv)) {
                    ok = false;
                    break;
                }
                hipLaunchKernelGGL(k_permute_column, dim3(g), dim3(kBlock), 0, stream(), perm, rows, J->pdata[(size_t)p], nd,
                    w, t.Nullable ? J->pvalid[(size_t)p] : nullptr, nv);
                dfree(J->pdata[(size_t)p]);
                J->pdata[(size_t)p] = nd;
                if (t.Nullable) {
                    dfree(J->pvalid[(size_t)p]);
                    J->pvalid[(size_t)p] = nv;
                }
            }
            ok = ok && hip_ok(hipGetLastError(), "build place");
        }
    }
    ok = ok && sync("build place");
    release();
    if (!ok) {
        dfree(new_keys);
        dfree(new_ord);
        return false;
    }
    dfree(J->keys);
    dfree(J->ordinal);
    J->keys = new_keys;
    J->ordinal = new_ord;
    J->cap = rows > 0 ? rows : 1;
    J->built = true;
    return true;
}

// The kernel's own (static) LDS bytes; the dynamic allocation of a launch
// must fit beside it within the device's per-workgroup limit.
size_t static_lds(const void* fn)
{
    hipFuncAttributes a;
    if (hipFuncGetAttributes(&a, fn) != hipSuccess) {
        (void)hipGetLastError();
        return 16 * 1024;  // conservative
    }
    return a.sharedSizeBytes;
}

template <int KW, int kTab, int kMode>
void launch_probe_one(int grid, size_t lds, const ProbeArgs& A)
{
    static bool attr_set = false;
    const void* fn = (const void*)k_probe<KW, kTab, kMode>;
    static const size_t s_lds = static_lds(fn);  // the kernel's static LDS, read once
    if (!attr_set) {
        // Room for the kernel's static LDS beside the dynamic allocation; a
        // refused attribute must not leave an error for the launch check.
        (void)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)(max_lds_per_block() - s_lds));
        (void)hipGetLastError();
        attr_set = true;
    }
    if (lds + s_lds > max_lds_per_block()) {
        fail("probe: %zu bytes of LDS exceed the workgroup limit", lds + s_lds);
        return;
    }
    hipLaunchKernelGGL((k_probe<KW, kTab, kMode>), dim3(grid), dim3(kProbeBlock), lds, stream(), A);
}

template <int KW, int kTab>
void launch_probe_tab(int mode, int grid, size_t lds, const ProbeArgs& A)
{
    if constexpr (KW == 1) {
        if (mode == kDirectVal) return launch_probe_one<KW, kTab, kDirectVal>(grid, lds, A);
        if (mode == kDirectKey) return launch_probe_one<KW, kTab, kDirectKey>(grid, lds, A);
    }
    launch_probe_one<KW, kTab, kGeneric>(grid, lds, A);
}

template <int kMode>
void launch_dense_one(int grid, size_t lds, const ProbeArgs& A)
{
    static bool attr_set = false;
    const void* fn = (const void*)k_probe_dense<kMode>;
    static const size_t s_lds = static_lds(fn);  // the kernel's static LDS, read once
    if (!attr_set) {
        (void)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)(max_lds_per_block() - s_lds));
        (void)hipGetLastError();
        attr_set = true;
    }
    if (lds + s_lds > max_lds_per_block()) {
        fail("dense probe: %zu bytes of LDS exceed the workgroup limit", lds + s_lds);
        return;
    }
    hipLaunchKernelGGL((k_probe_dense<kMode>), dim3(grid), dim3(kProbeBlock), lds, stream(), A);
}

template <bool kVal, bool kMinMax, bool kLeft, bool kPacked>
void launch_lean_one(int grid, size_t lds, const ProbeArgs& A)
{
    static bool attr_set = false;
    const void* fn = (const void*)k_probe_dense_lean<kVal, kMinMax, kLeft, kPacked>;
    static const size_t s_lds = static_lds(fn);  // the kernel's static LDS, read once
    if (!attr_set) {
        (void)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)(max_lds_per_block() - s_lds));
        (void)hipGetLastError();
        attr_set = true;
    }
    if (lds + s_lds > max_lds_per_block()) {
        fail("dense probe: %zu bytes of LDS exceed the workgroup limit", lds + s_lds);
        return;
    }
    hipLaunchKernelGGL((k_probe_dense_lean<kVal, kMinMax, kLeft, kPacked>), dim3(grid), dim3(kProbeBlock), lds, stream(), A);
}

template <bool kVal, bool kMinMax>
void launch_lean_left(bool left, int grid, size_t lds, const ProbeArgs& A)
{
    if (A.dpack) {
        if (left) launch_lean_one<kVal, kMinMax, true, true>(grid, lds, A);
        else launch_lean_one<kVal, kMinMax, false, true>(grid, lds, A);
        return;
    }
    if (left) launch_lean_one<kVal, kMinMax, true, false>(grid, lds, A);
    else launch_lean_one<kVal, kMinMax, false, false>(grid, lds, A);
}

void launch_dense_lean(bool val, bool minmax, bool left, int grid, size_t lds, const ProbeArgs& A)
{
    if (val) {
        if (minmax) launch_lean_left<true, true>(left, grid, lds, A);
        else launch_lean_left<true, false>(left, grid, lds, A);
    } else {
        launch_lean_left<false, false>(left, grid, lds, A);
    }
}

// Dynamic LDS the dense probe of a mode may use: the device limit less the
// kernel's static LDS.
template <int kMode>
size_t dense_lds_budget()
{
    static size_t cached = 0;
    if (cached) return cached;
    size_t s = static_lds((const void*)k_probe_dense<kMode>);
    // The lean forms (same LDS use) fit the same budget.
    const void* lean[] = {(const void*)k_probe_dense_lean<true, true, true, false>, (const void*)k_probe_dense_lean<true, true, false, false>,
        (const void*)k_probe_dense_lean<true, false, true, false>, (const void*)k_probe_dense_lean<true, false, false, false>,
        (const void*)k_probe_dense_lean<false, false, true, false>, (const void*)k_probe_dense_lean<false, false, false, false>,
        (const void*)k_probe_dense_lean<true, true, true, true>, (const void*)k_probe_dense_lean<true, true, false, true>,
        (const void*)k_probe_dense_lean<true, false, true, true>, (const void*)k_probe_dense_lean<true, false, false, true>,
        (const void*)k_probe_dense_lean<false, false, true, true>, (const void*)k_probe_dense_lean<false, false, false, true>};
    for (const void* f : lean) s = s > static_lds(f) ? s : static_lds(f);
    cached = max_lds_per_block() > s ? max_lds_per_block() - s : 0;
    return cached;
}

size_t dense_lds_budget_of(int mode)
{
    return mode == kDirectVal ? dense_lds_budget<kDirectVal>() : mode == kDirectKey ? dense_lds_budget<kDirectKey>()
                                                                                       : dense_lds_budget<kGeneric>();
}

// Direct modes and the dense table exist for one-word keys only.
template <int KW>
void launch_probe(int mode, int tab, int grid, size_t lds, const ProbeArgs& A)
{
    if constexpr (KW == 1) {
        if (tab == kTabDense) {
            if (mode == kDirectVal) return launch_dense_one<kDirectVal>(grid, lds, A);
            if (mode == kDirectKey) return launch_dense_one<kDirectKey>(grid, lds, A);
            return launch_dense_one<kGeneric>(grid, lds, A);
        }
    }
    if (tab == kTabLds) return launch_probe_tab<KW, kTabLds>(mode, grid, lds, A);
    launch_probe_tab<KW, kTabHash>(mode, grid, lds, A);
}

// The probe's counters before a launch: the pair count 0 and the first
// `words` fused words at their identities (sums 0, MIN INT64_MAX, MAX
// INT64_MIN).
__global__ void k_probe_init(unsigned long long* pair_count, unsigned long long* acc, int words)
{
    probe_counters_init(pair_count, acc, words, (int)threadIdx.x);
}

// The probe's arguments for one batch against J's table (everything but
// the LDS staging, the pair buffer and the counters) and the kernel mode.
template <int KW>
bool probe_args_of(JoinBuild* J, const TRowSet* batch, ProbeArgs& A, int* mode_out)
{
    JoinPlan* P = J->plan;
    const int64_t n = batch->RowCount;
    memset(&A, 0, sizeof(A));
    if (!bind_side(P->probe_slot_col, P->probe_nslots, P->spec.Probe, batch, &A.cols)) return false;
    A.K = P->probe_key;
    {
        const QdbValueType& kt = P->spec.Probe.Types[P->probe_slot_col[0]];
        A.key_i64 = KW == 1 && P->probe_key.nkeys == 1 && kt.Width == 8 && !kt.Nullable && kt.Kind != QDB_KIND_FLOAT
            && kt.Kind != QDB_KIND_BOOL && P->probe_key.f[0].slot == 0;
    }
    A.n = n;
    A.sel = batch->Selection;
    A.table = J->table;
    A.cmask = J->capacity - 1;
    A.ordinal = J->ordinal;
    A.ord_to_arena = J->ord_to_arena;
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
    A.fused_part = nullptr;
    // Kernel mode: direct key loads, and the fused aggregates' one probe
    // column streamed beside the key when it is a non-nullable 8-byte
    // integer.
    int mode = A.key_i64 ? kDirectKey : kGeneric;
    if (mode == kDirectKey) {
        int vslot = -1;
        bool one = true;
        for (int f = 0; f < A.nfused; ++f) {
            const FusedDesc& d = P->fused[f];
            if (d.from_build) one = false;
            if (d.kind == FK_COUNT_PAIRS || d.from_build) continue;
            if (vslot >= 0 && vslot != d.slot) one = false;
            vslot = d.slot;
        }
        if (one && vslot >= 0) {
            const QdbValueType& vt = P->spec.Probe.Types[P->probe_slot_col[vslot]];
            if (vt.Width == 8 && !vt.Nullable && vt.Kind != QDB_KIND_FLOAT && vt.Kind != QDB_KIND_BOOL) {
                mode = kDirectVal;
                A.val_slot = vslot;
            }
        }
    }
    *mode_out = mode;
    return true;
}

// Pairs a probe of n rows against J can emit, and the capacity its first
// attempt allocates: a probe row has at most max_key_rows matches (one pair
// when unmatched in an outer join), and the first attempt takes that bound
// up to kFirstPairBudget pairs (the kernels count every pair whatever the
// capacity, so a larger result is probed again at its exact size).
unsigned long long first_pair_capacity(const JoinBuild* J, int64_t n)
{
    const bool outer = J->plan->spec.JoinType != QDB_JOIN_INNER;
    const u128 bound = (u128)(uint64_t)n * (u128)(uint64_t)(J->max_key_rows > 1 ? J->max_key_rows : 1)
        + (outer ? (u128)(uint64_t)n : (u128)0);
    const unsigned long long budget = (unsigned long long)n > kFirstPairBudget ? (unsigned long long)n : kFirstPairBudget;
    return bound < (u128)budget ? (unsigned long long)bound : budget;
}

// A pair buffer of at least `cap` pairs.
bool reserve_pairs(QdbPairBuffer* pairs, unsigned long long cap)
{
    if ((unsigned long long)pairs->Capacity >= cap) return true;
    if (cap > (unsigned long long)(SIZE_MAX / 16)) return fail("probe: %llu pairs do not fit in memory", cap);
    dfree(pairs->Data);
    pairs->Capacity = 0;
    pairs->Data = (int64_t*)dalloc((size_t)cap * 16);
    if (!pairs->Data) return false;
    pairs->Capacity = (int64_t)cap;
    return true;
}

template <int KW>
// Probes one batch. With `scratch_part` (the allocation-free variant) no
// pair is written and nothing is allocated: the fused aggregates' partials
// go to that device scratch, the pair count to *pairs_seen.
bool do_probe(JoinBuild* J, const TRowSet* batch, QdbPairBuffer* pairs, unsigned long long* scratch_part = nullptr,
    unsigned long long* pairs_seen = nullptr)
{
    const int64_t n = batch->RowCount;
    pairs->Count = 0;
    J->last_probe = *batch;
    J->fused_ready = false;
    if (n <= 0) return true;
    if (n >= (int64_t)1 << 32) return fail("probe batches are limited to 2^32 rows");
    ProbeArgs A;
    int mode = kGeneric;
    if (!probe_args_of<KW>(J, batch, A, &mode)) return false;
    // The dense probe's lean form: direct key, no probe selection, inner
    // join, fused aggregates that count pairs or read the one direct probe
    // column (QDB_GPU_JOIN_LEAN=0 keeps the general form).
    bool lean = J->dense && A.sel == nullptr && !A.emit_unmatched && (mode == kDirectVal || mode == kDirectKey);
    bool minmax = false;
    for (int f = 0; f < A.nfused; ++f) {
        const int kind = A.fused[f].kind;
        if (mode == kDirectKey && kind != FK_COUNT_PAIRS) lean = false;
        if (kind == FK_MIN || kind == FK_MAX) minmax = true;
    }
    lean = lean && n >= kLeanWaveTile;  // every wave tile lies inside the batch
    if (const char* e = getenv("QDB_GPU_JOIN_LEAN")) lean = lean && e[0] != '0';
    // The slot table in LDS when it fits in 64 KiB; else the Bloom filter.
    // Dense: as much of the membership bitmap as LDS holds.
    const size_t table_bytes = J->capacity * sizeof(Slot<KW>);
    const bool lds_table = !J->dense && table_bytes <= 64 * 1024;
    size_t lds = lds_table ? table_bytes : (size_t)J->bloom_words * 4;
    int tab = lds_table ? kTabLds : kTabHash;
    if (J->dense) {
        tab = kTabDense;
        const uint64_t words = (J->drange + 31) / 32;
        const uint64_t fit = dense_lds_budget_of(mode) / 4;  // beside the kernels' static LDS
        A.lds_words = (uint32_t)(words < fit ? words : fit);
        A.dmin = J->dmin;
        A.drange = J->drange;
        A.dbits = J->dbits;
        A.dtab = J->dtab;
        A.dpack = J->dpack;
        A.dpack_w = J->dpack_w;
        lds = (size_t)A.lds_words * 4;
    }
    unsigned long long cap = first_pair_capacity(J, n);
    for (int attempt = 0; attempt < 2; ++attempt) {
        if (!scratch_part && !reserve_pairs(pairs, cap)) return false;
        A.pairs = scratch_part ? nullptr : pairs->Data;
        A.pair_cap = scratch_part ? 0ull : (unsigned long long)pairs->Capacity;
        A.pair_count = J->pair_count;
        if (!J->counters_ready || attempt > 0) {
            hipLaunchKernelGGL(k_probe_init, dim3(1), dim3(64), 0, stream(), J->pair_count, J->fused_acc, A.nfused * kAccWords);
            if (!hip_ok(hipGetLastError(), "probe init")) return false;
        }
        J->counters_ready = false;
        // Fused accumulators are per probe call (k_reduce_partials writes
        // every word); they are summed on the host.
        const int64_t tiles = (n + (int64_t)kProbeBlock / kWave * kWaveTile - 1) / ((int64_t)kProbeBlock / kWave * kWaveTile);
        const int per_cu = lds_table ? 2 : 1;
        const int grid = (int)(tiles < (int64_t)num_cus() * per_cu ? tiles : (int64_t)num_cus() * per_cu);
        // Fused totals: the workgroups add their words to the join's
        // accumulators with atomics; the allocation-free variant writes
        // per-workgroup partials into its scratch and reduces them.
        unsigned long long* part = nullptr;
        A.fused_part = nullptr;
        A.fused_acc = J->fused_acc;
        if (A.nfused && scratch_part) {
            const size_t part_bytes = (size_t)grid * kMaxFused * kAccWords * 8;
            if (part_bytes > kNoAllocScratchBytes) {
                return fail("no-alloc probe: %zu bytes of partials exceed the %zu-byte scratch", part_bytes,
                    (size_t)kNoAllocScratchBytes);
            }
            part = scratch_part;
            A.fused_part = part;
        }
        {
            TimedRegion timer("join_probe");
            if (lean) launch_dense_lean(mode == kDirectVal, minmax, A.probe_is_left != 0, grid, lds, A);
            else launch_probe<KW>(mode, tab, grid, lds, A);
            if (part) {
                ReduceOps ops;
                for (int k = 0; k < kMaxFused * kAccWords; ++k) {
                    const int w = k % kAccWords;
                    ops.op[k] = w < 4 ? 0 : w == 4 ? 1 : 2;
                }
                hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(kBlock), 0, stream(), part, grid, kMaxFused * kAccWords,
                    A.nfused * kAccWords, ops, J->fused_acc);
            }
        }
        if (!last_error_text().empty()) return false;
        if (!hip_ok(hipGetLastError(), "probe launch")) return false;
        // The pair count and the fused words in one readback.
        const size_t rb = 16 + (size_t)A.nfused * kAccWords * 8;
        if (!hip_ok(hipMemcpyAsync(J->readback, J->pair_count, rb, hipMemcpyDeviceToHost, stream()), "probe readback") || !sync("probe"))
            return false;
        const unsigned long long count = J->readback[0];
        if (scratch_part) {
            *pairs_seen = count;
            J->fused_ready = A.nfused > 0;
            return true;
        }
        if (count <= A.pair_cap) {
            pairs->Count = (int64_t)count;
            J->fused_ready = A.nfused > 0;
            return true;
        }
        cap = count;  // more pairs than the first attempt's buffer: again, at the exact size
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

// Folds one probe's fused words (the readback after the pair count) into
// the results r[0..nf).
// The words of a probe that saw no row (MIN/MAX words at their identities).
const unsigned long long kZeroFusedWords[kMaxFused * kAccWords] = {
    0, 0, 0, 0, 0x7fffffffffffffffull, 0x8000000000000000ull, 0, 0, 0, 0, 0x7fffffffffffffffull, 0x8000000000000000ull,
    0, 0, 0, 0, 0x7fffffffffffffffull, 0x8000000000000000ull, 0, 0, 0, 0, 0x7fffffffffffffffull, 0x8000000000000000ull};

void merge_fused_words(const JoinPlan* P, const unsigned long long* w, QdbFusedResult* res)
{
    const int nf = P->spec.NumFused;
    for (int f = 0; f < nf; ++f) {
        const unsigned long long* a = &w[(size_t)f * kAccWords];
        QdbFusedResult& r = res[f];
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
}

bool collect_fused(JoinBuild* J)
{
    JoinPlan* P = J->plan;
    const int nf = P->spec.NumFused;
    if (!nf || !J->fused_ready) return true;
    J->fused_ready = false;
    std::lock_guard<std::mutex> lock(g_fused_mu);
    FusedTotals& T = fused_totals()[J];
    if (T.r.empty()) {
        T.r.resize((size_t)nf);
        for (int f = 0; f < nf; ++f) {
            memset(&T.r[(size_t)f], 0, sizeof(QdbFusedResult));
        }
    }
    merge_fused_words(P, J->readback + 2, T.r.data());
    return true;
}

// The blobs' fields after a Build.
void publish_build(RowArenaBlob* AB, RowTableBlob* TB, const JoinBuild* J)
{
    AB->Rows = J->rows;
    TB->entries = J->table;
    TB->Capacity = (int64_t)J->capacity;
    TB->Mask = J->capacity - 1;
    int shift = 64;
    for (uint64_t c = J->capacity; c > 1; c >>= 1) --shift;
    TB->Shift = shift;
    TB->Rows = J->rows;
    TB->Keys = J->keys_distinct;
}

void register_pairs(const void* pairs, JoinBuild* J)
{
    std::lock_guard<std::mutex> lock(g_pairs_mu);
    pair_owner()[pairs] = J;
}

// Build of every join arenas[i] (one compiled join): the dense tables'
// kernels and readbacks of all of them are queued before ONE
// synchronisation; a join whose keys the dense table does not take gets
// the hash table as by its own Build.
bool batched_build(JoinPlan* P, void* const* arenas, void* const* tables, int64_t n)
{
    std::vector<char> queued((size_t)n, 0);
    bool any = false;
    for (int64_t i = 0; i < n; ++i) {
        JoinBuild* J = ((RowArenaBlob*)arenas[i])->build;
        if (!J) return fail("batched build of join %lld before Init", (long long)i);
        if (J->plan != P) return fail("join %lld belongs to another compiled join", (long long)i);
        if (J->built) return fail("join %lld is already built", (long long)i);
        if (J->appended > (int64_t)0xffffffff) return fail("join builds are limited to 2^32 appended rows");
    }
    for (int64_t i = 0; i < n; ++i) {
        JoinBuild* J = ((RowArenaBlob*)arenas[i])->build;
        if (dense_wanted(J)) {
            queued[(size_t)i] = dense_build_launch(J) ? 1 : 0;
            any = any || queued[(size_t)i];
        }
    }
    const bool synced = !any || sync("batched dense fill");
    bool ok = true;
    for (int64_t i = 0; i < n; ++i) {
        JoinBuild* J = ((RowArenaBlob*)arenas[i])->build;
        bool built = queued[(size_t)i] && dense_build_finish(J, synced);
        if (!built && ok) {
            ok = P->KW == 1 ? do_build<1>(J, queued[(size_t)i] != 0) : do_build<2>(J, queued[(size_t)i] != 0);
            built = ok;
        }
        if (built) publish_build((RowArenaBlob*)arenas[i], (RowTableBlob*)tables[i], J);
    }
    return ok;
}

void launch_probe_batched(int mode, int grid, const ProbeArgs* insts, const ProbeItem* items, const int* ctl, unsigned int* head)
{
    if (mode == kDirectVal) hipLaunchKernelGGL(k_probe_batched<kDirectVal>, dim3(grid), dim3(kProbeBlock), 0, stream(), insts, items, ctl, head);
    else if (mode == kDirectKey) hipLaunchKernelGGL(k_probe_batched<kDirectKey>, dim3(grid), dim3(kProbeBlock), 0, stream(), insts, items, ctl, head);
    else hipLaunchKernelGGL(k_probe_batched<kGeneric>, dim3(grid), dim3(kProbeBlock), 0, stream(), insts, items, ctl, head);
}

// Rows per work item of the batched probe: four block tiles.
constexpr int64_t kProbeItemRows = 4 * kProbeBlockTile;

// Probe of batches[i] against join arenas[i] into pairs[i] for every i (one
// compiled join with a one-word key), in ONE persistent launch
// (k_probe_batched) between one counter-initialising launch and one launch
// that collects every instance's pair count and fused words for a single
// readback. A probe whose pairs outgrow the first capacity is probed again
// on its own at the exact size. Each pair buffer and the fused results are
// then as after jr_dispatch(Probe) of that batch.
bool batched_probe(JoinPlan* P, void* const* arenas, const TRowSet* const* batches, void* const* pairs, int64_t n)
{
    for (int64_t i = 0; i < n; ++i) {
        JoinBuild* J = ((RowArenaBlob*)arenas[i])->build;
        if (!J || !J->built) return fail("batched probe of join %lld before Build", (long long)i);
        if (J->plan != P) return fail("join %lld belongs to another compiled join", (long long)i);
        if (!batches[i] || !pairs[i]) return fail("batched probe %lld: null batch or pair buffer", (long long)i);
        if (batches[i]->RowCount >= (int64_t)1 << 32) return fail("probe batches are limited to 2^32 rows");
    }
    if (P->KW != 1) {
        // Two-word keys: one probe at a time.
        for (int64_t i = 0; i < n; ++i) {
            JoinBuild* J = ((RowArenaBlob*)arenas[i])->build;
            if (!do_probe<2>(J, batches[i], (QdbPairBuffer*)pairs[i]) || !collect_fused(J)) return false;
            register_pairs(pairs[i], J);
        }
        return true;
    }
    std::vector<ProbeArgs> insts;
    std::vector<ProbeItem> items;
    std::vector<int64_t> which;
    int mode = -1;
    for (int64_t i = 0; i < n; ++i) {
        JoinBuild* J = ((RowArenaBlob*)arenas[i])->build;
        QdbPairBuffer* PB = (QdbPairBuffer*)pairs[i];
        const TRowSet* b = batches[i];
        PB->Count = 0;
        J->last_probe = *b;
        J->fused_ready = false;
        register_pairs(PB, J);
        if (b->RowCount <= 0) continue;
        ProbeArgs A;
        int m = kGeneric;
        if (!probe_args_of<1>(J, b, A, &m)) return false;
        if (mode >= 0 && m != mode) return fail("batched probe: kernel modes differ");
        mode = m;
        // Nothing staged in LDS per item: the dense table reads its ordinal
        // table for every key, the hash table has no Bloom filter.
        A.lds_words = 0;
        A.bloom_words = 0;
        if (J->dense) {
            A.dmin = J->dmin;
            A.drange = J->drange;
            A.dbits = J->dbits;
            A.dtab = J->dtab;
        }
        if (!reserve_pairs(PB, first_pair_capacity(J, b->RowCount))) return false;
        A.pairs = PB->Data;
        A.pair_cap = (unsigned long long)PB->Capacity;
        A.pair_count = J->pair_count;
        A.fused_acc = J->fused_acc;
        A.fused_part = nullptr;
        const int k = (int)insts.size();
        insts.push_back(A);
        which.push_back(i);
        for (int64_t r = 0; r < b->RowCount; r += kProbeItemRows) {
            ProbeItem it;
            it.inst = k;
            it.pad = 0;
            it.b_begin = r;
            it.b_end = r + kProbeItemRows < b->RowCount ? r + kProbeItemRows : b->RowCount;
            items.push_back(it);
        }
    }
    if (insts.empty()) return true;
    const int nk = (int)insts.size();
    const int nitems = (int)items.size();
    const int words = P->spec.NumFused * kAccWords;
    const int stride = 1 + words;
    const int ctl[4] = {nk, nitems, 0, 0};
    ProbeArgs* d_insts = (ProbeArgs*)upload(insts.data(), insts.size() * sizeof(ProbeArgs));
    ProbeItem* d_items = (ProbeItem*)upload(items.data(), items.size() * sizeof(ProbeItem));
    int* d_ctl = (int*)upload(ctl, sizeof(ctl));
    unsigned int* d_head = (unsigned int*)dalloc(16);
    unsigned long long* d_out = (unsigned long long*)dalloc((size_t)nk * stride * 8);
    unsigned long long* h_out = (unsigned long long*)halloc_pinned((size_t)nk * stride * 8);
    auto release = [&]() {
        dfree(d_insts);
        dfree(d_items);
        dfree(d_ctl);
        dfree(d_head);
        dfree(d_out);
        hfree_pinned(h_out);
    };
    bool ok = d_insts && d_items && d_ctl && d_head && d_out && h_out && zero_fill(d_head, 16);
    if (ok) {
        hipLaunchKernelGGL(k_probe_init_batched, dim3(nk < 1024 ? nk : 1024), dim3(64), 0, stream(), d_insts, d_ctl, words);
        const int grid = nitems < num_cus() ? nitems : num_cus();
        {
            TimedRegion timer("join_probe_batched");
            launch_probe_batched(mode, grid, d_insts, d_items, d_ctl, d_head);
        }
        hipLaunchKernelGGL(k_probe_collect_batched, dim3(grid_for((int64_t)nk * stride, kBlock, 1)), dim3(kBlock), 0, stream(), d_insts,
            d_ctl, words, d_out);
        ok = hip_ok(hipGetLastError(), "batched probe launch") &&
            hip_ok(hipMemcpyAsync(h_out, d_out, (size_t)nk * stride * 8, hipMemcpyDeviceToHost, stream()), "batched probe readback") &&
            sync("batched probe");
    }
    for (int k = 0; ok && k < nk; ++k) {
        const int64_t i = which[(size_t)k];
        JoinBuild* J = ((RowArenaBlob*)arenas[i])->build;
        QdbPairBuffer* PB = (QdbPairBuffer*)pairs[i];
        const unsigned long long* w = h_out + (size_t)k * stride;
        J->counters_ready = false;
        if (w[0] > insts[(size_t)k].pair_cap) {
            // More pairs than the first capacity: this probe again, alone.
            ok = do_probe<1>(J, batches[i], PB) && collect_fused(J);
            continue;
        }
        PB->Count = (int64_t)w[0];
        memcpy(J->readback + 2, w + 1, (size_t)words * 8);
        J->fused_ready = words > 0;
        ok = collect_fused(J);
    }
    release();
    return ok;
}

// The calling thread's stream's pipeline for `mode`, made and captured at
// first use.
ProbePipeline* probe_pipeline(JoinPlan* P, int mode)
{
    const int si = thread_stream_index();
    {
        std::lock_guard<std::mutex> lock(P->pipe_mu);
        if (P->pipes[si][mode]) return P->pipes[si][mode];
    }
    ProbePipeline* pp = new ProbePipeline();
    pp->mode = mode;
    pp->words = P->spec.NumFused * kAccWords;
    pp->d_buf = (char*)dalloc(ProbePipeline::buf_bytes());
    pp->h_stage = (char*)halloc_pinned(ProbePipeline::buf_bytes());
    pp->d_head = (unsigned int*)dalloc(16);
    pp->d_counters = (unsigned long long*)dalloc((size_t)ProbePipeline::kMaxInst * ProbePipeline::kStride * 8);
    const size_t out_bytes = (size_t)ProbePipeline::kMaxInst * (1 + pp->words) * 8;
    pp->d_out = (unsigned long long*)dalloc(out_bytes);
    pp->h_out = (unsigned long long*)halloc_pinned(out_bytes);
    if (!pp->d_buf || !pp->h_stage || !pp->d_head || !pp->d_counters || !pp->d_out || !pp->h_out || !sync("probe pipeline")) {
        delete pp;
        return nullptr;
    }
    const ProbeArgs* d_insts = (const ProbeArgs*)(pp->d_buf + ProbePipeline::insts_off());
    const ProbeItem* d_items = (const ProbeItem*)(pp->d_buf + ProbePipeline::items_off());
    const int* d_ctl = (const int*)pp->d_buf;
    hipStream_t s = stream();
    // Capture: kernels and one copy; the queue head is cleared by a kernel
    // (no memset node).
    // No other thread allocates or frees device memory meanwhile (either can
    // invalidate the capture); a capture that fails anyway is made again.
    bool ok = false;
    for (int attempt = 0; attempt < 3 && !ok; ++attempt) {
        std::unique_lock<std::shared_mutex> g(capture_mutex());
        if (attempt > 0) {
            (void)hipGetLastError();
            clear_error();
            if (pp->graph) (void)hipGraphDestroy(pp->graph);
            pp->graph = nullptr;
        }
        if (!hip_ok(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "probe pipeline capture")) continue;
        zero_fill(pp->d_head, 16);
        hipLaunchKernelGGL(k_probe_init_batched, dim3(ProbePipeline::kMaxInst), dim3(64), 0, s, d_insts, d_ctl, pp->words);
        launch_probe_batched(mode, num_cus(), d_insts, d_items, d_ctl, pp->d_head);
        hipLaunchKernelGGL(k_probe_collect_batched, dim3(64), dim3(kBlock), 0, s, d_insts, d_ctl, pp->words, pp->d_out);
        (void)hipMemcpyAsync(pp->h_out, pp->d_out, out_bytes, hipMemcpyDeviceToHost, s);
        const hipError_t e = hipStreamEndCapture(s, &pp->graph);
        ok = hip_ok(e, "probe pipeline capture") && pp->graph &&
            hip_ok(hipGraphInstantiate(&pp->exec, pp->graph, nullptr, nullptr, 0), "probe pipeline instantiate");
    }
    if (!ok) {
        delete pp;
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(P->pipe_mu);
    if (P->pipes[si][mode]) {
        delete pp;  // another caller on this stream made one meanwhile
        return P->pipes[si][mode];
    }
    P->pipes[si][mode] = pp;
    return pp;
}

// Detached batched probe: batches[i] (i < n) against ONE built join J,
// each into its own pair buffer pairs[i] (NULL: pairs counted, not written),
// counts[i] and fused[i * NumFused + f] returned per batch; J's own counters,
// pair buffer and fused totals are not touched, so calls on different
// threads (each on its own stream) may probe one resident table together.
// Runs as replays of the thread's captured pipeline, up to 256 batches and
// 16384 items per replay.
bool detached_probe(JoinPlan* P, JoinBuild* J, const TRowSet* const* batches, QdbPairBuffer* const* pairs, int64_t n,
    int64_t* counts, QdbFusedResult* fused)
{
    const int nf = P->spec.NumFused;
    for (int64_t i = 0; i < n; ++i) {
        counts[i] = 0;
        if (fused && nf) {
            // The results of a batch without rows (COUNT 0, the rest NULL).
            for (int f = 0; f < nf; ++f) memset(&fused[i * nf + f], 0, sizeof(QdbFusedResult));
            merge_fused_words(P, kZeroFusedWords, &fused[i * nf]);
        }
        if (!batches[i]) return fail("detached probe %lld: null batch", (long long)i);
        if (batches[i]->RowCount >= (int64_t)1 << 32) return fail("probe batches are limited to 2^32 rows");
    }
    std::vector<int64_t> todo;
    for (int64_t i = 0; i < n; ++i) {
        if (batches[i]->RowCount > 0) todo.push_back(i);
    }
    std::vector<unsigned long long> caps((size_t)n, 0);
    ProbePipeline* pp = nullptr;
    int pp_mode = -1;
    for (int pass = 0; pass < 2 && !todo.empty(); ++pass) {
        std::vector<int64_t> again;
        size_t at = 0;
        while (at < todo.size()) {
            // One replay: as many batches as the buffers hold.
            std::vector<ProbeArgs> insts;
            std::vector<ProbeItem> items;
            std::vector<int64_t> which;
            while (at < todo.size() && (int)insts.size() < ProbePipeline::kMaxInst) {
                const int64_t i = todo[at];
                const TRowSet* b = batches[i];
                const int64_t nitems_i = (b->RowCount + kProbeItemRows - 1) / kProbeItemRows;
                if (!insts.empty() && (int64_t)items.size() + nitems_i > ProbePipeline::kMaxItems) break;
                if (nitems_i > ProbePipeline::kMaxItems) return fail("detached probe: batch %lld has too many rows", (long long)i);
                ProbeArgs A;
                int m = kGeneric;
                if (!probe_args_of<1>(J, b, A, &m)) return false;
                if (pp_mode >= 0 && m != pp_mode) return fail("detached probe: kernel modes differ");
                if (!pp) {
                    pp = probe_pipeline(P, m);
                    if (!pp) return false;
                    pp_mode = m;
                }
                A.lds_words = 0;
                A.bloom_words = 0;
                if (J->dense) {
                    A.dmin = J->dmin;
                    A.drange = J->drange;
                    A.dbits = J->dbits;
                    A.dtab = J->dtab;
                }
                QdbPairBuffer* PB = pairs ? pairs[i] : nullptr;
                if (PB) {
                    PB->Count = 0;
                    const unsigned long long want = pass == 0 ? first_pair_capacity(J, b->RowCount) : caps[(size_t)i];
                    if (!reserve_pairs(PB, want)) return false;
                    A.pairs = PB->Data;
                    A.pair_cap = (unsigned long long)PB->Capacity;
                } else {
                    A.pairs = nullptr;
                    A.pair_cap = 0;
                }
                const int k = (int)insts.size();
                A.pair_count = pp->d_counters + (size_t)k * ProbePipeline::kStride;
                A.fused_acc = A.pair_count + 2;
                A.fused_part = nullptr;
                insts.push_back(A);
                which.push_back(i);
                for (int64_t r = 0; r < b->RowCount; r += kProbeItemRows) {
                    ProbeItem it;
                    it.inst = k;
                    it.pad = 0;
                    it.b_begin = r;
                    it.b_end = r + kProbeItemRows < b->RowCount ? r + kProbeItemRows : b->RowCount;
                    items.push_back(it);
                }
                ++at;
            }
            int* ctl = (int*)pp->h_stage;
            ctl[0] = (int)insts.size();
            ctl[1] = (int)items.size();
            ctl[2] = ctl[3] = 0;
            memcpy(pp->h_stage + ProbePipeline::insts_off(), insts.data(), insts.size() * sizeof(ProbeArgs));
            memcpy(pp->h_stage + ProbePipeline::items_off(), items.data(), items.size() * sizeof(ProbeItem));
            const size_t head_bytes = ProbePipeline::insts_off() + insts.size() * sizeof(ProbeArgs);
            bool ok = hip_ok(hipMemcpyAsync(pp->d_buf, pp->h_stage, head_bytes, hipMemcpyHostToDevice, stream()), "pipeline descriptors") &&
                hip_ok(hipMemcpyAsync(pp->d_buf + ProbePipeline::items_off(), pp->h_stage + ProbePipeline::items_off(),
                           items.size() * sizeof(ProbeItem), hipMemcpyHostToDevice, stream()),
                    "pipeline items");
            {
                TimedRegion timer("join_probe_detached");
                ok = ok && hip_ok(hipGraphLaunch(pp->exec, stream()), "pipeline launch");
            }
            ok = ok && sync("detached probe");
            if (!ok) return false;
            const int stride = 1 + pp->words;
            for (size_t k = 0; k < which.size(); ++k) {
                const int64_t i = which[k];
                const unsigned long long* w = pp->h_out + k * stride;
                if (pairs && pairs[i] && w[0] > insts[k].pair_cap) {
                    // More pairs than the first capacity: probed again at the exact size.
                    caps[(size_t)i] = w[0];
                    again.push_back(i);
                    continue;
                }
                counts[i] = (int64_t)w[0];
                if (pairs && pairs[i]) pairs[i]->Count = (int64_t)w[0];
                if (fused && nf) {
                    for (int f = 0; f < nf; ++f) memset(&fused[i * nf + f], 0, sizeof(QdbFusedResult));
                    merge_fused_words(P, w + 1, &fused[i * nf]);
                }
            }
        }
        todo.swap(again);
    }
    if (!todo.empty()) return fail("detached probe: pair buffer overflow");
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
        J->pair_count = (unsigned long long*)dalloc(16 + (size_t)kMaxFused * kAccWords * 8);
        J->fused_acc = J->pair_count ? J->pair_count + 2 : nullptr;
        // The arena words are set by the first append (an arena with no
        // append has no rows, known on the host).
        J->rows_dev = (unsigned long long*)dalloc(kArenaWords * 8);
        J->readback = (unsigned long long*)halloc_pinned(sizeof(unsigned long long) * kReadbackWords);
        if (!J->pair_count || !J->rows_dev || !J->readback) {
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
        publish_build(AB, TB, J);
        return true;
    }
    case QDB_JR_OP_PROBE: {
        JoinBuild* J = AB->build;
        if (!J || !J->built) return fail("join probe before Build");
        if (!batch) return fail("join probe without a batch");
        bool ok = P->KW == 1 ? do_probe<1>(J, batch, PB) : do_probe<2>(J, batch, PB);
        if (ok) ok = collect_fused(J);
        if (ok) register_pairs(pairs, J);
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
    if (!publish_rowset(out)) return 0;
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

// ---------------------------------------------------------------------------
// The probe that allocates nothing (qdb_gpu_join_probe_fused_noalloc). Its
// scratch, the fused aggregates' per-workgroup partials, is this file-scope
// device array: every device the library runs on holds one copy until the
// process ends. Host code never uses the array's name as an address (on the
// host it only names the symbol): the device address comes from
// hipGetSymbolAddress, once per device, with that device current.
// ---------------------------------------------------------------------------

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {

constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxScratchDevices] = {};

// The current device's address of `scratch` (looked up at the first call on
// that device), or nullptr with the error text.
unsigned long long* noalloc_scratch()
{
    int dev = 0;
    if (!hip_ok(hipGetDevice(&dev), "hipGetDevice")) return nullptr;
    if (dev < 0 || dev >= kMaxScratchDevices) return fail("device %d: no scratch slot", dev), nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        if (!hip_ok(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)")) return nullptr;
        g_scratch_ptr[dev] = ptr;
    }
    return (unsigned long long*)g_scratch_ptr[dev];
}

}  // namespace

extern "C" int qdb_gpu_join_build_batched(const QdbRowJoinKernels* kernels, void* const* arenas, void* const* tables, int64_t n)
{
    clear_error();
    if (!kernels || (n > 0 && (!arenas || !tables))) return fail("null argument"), -1;
    if (kernels->Slot < 0 || kernels->Slot >= kSlots) return fail("bad join slot %d", kernels->Slot), -1;
    JoinPlan* P = join_slots().get(kernels->Slot);
    if (!P) return fail("join kernels are not compiled"), -1;
    for (int64_t i = 0; i < n; ++i) {
        if (!arenas[i] || !tables[i]) return fail("null arena or table %lld", (long long)i), -1;
    }
    return batched_build(P, arenas, tables, n) ? 0 : -1;
}

extern "C" int qdb_gpu_join_probe_batched(const QdbRowJoinKernels* kernels, void* const* arenas, const TRowSet* const* batches,
    void* const* pairs, int64_t n)
{
    clear_error();
    if (!kernels || (n > 0 && (!arenas || !batches || !pairs))) return fail("null argument"), -1;
    if (kernels->Slot < 0 || kernels->Slot >= kSlots) return fail("bad join slot %d", kernels->Slot), -1;
    JoinPlan* P = join_slots().get(kernels->Slot);
    if (!P) return fail("join kernels are not compiled"), -1;
    for (int64_t i = 0; i < n; ++i) {
        if (!arenas[i]) return fail("null arena %lld", (long long)i), -1;
    }
    return batched_probe(P, arenas, batches, pairs, n) ? 0 : -1;
}

extern "C" int qdb_gpu_pairs_checksum(const int64_t* pairs, int64_t n, uint64_t* out)
{
    clear_error();
    if (!out || (n > 0 && !pairs) || n < 0) return fail("bad argument"), -1;
    if (!ensure_init()) return -1;
    unsigned long long* d = (unsigned long long*)dalloc(16);
    bool ok = d && zero_fill(d, 16);
    if (ok && n > 0) {
        hipLaunchKernelGGL(k_pairs_checksum, dim3(grid_for(n, kBlock, 4)), dim3(kBlock), 0, stream(), pairs, n, d);
        ok = hip_ok(hipGetLastError(), "pairs checksum");
    }
    ok = ok && hip_ok(hipMemcpyAsync(out, d, 16, hipMemcpyDeviceToHost, stream()), "pairs checksum") && sync("pairs checksum");
    dfree(d);
    return ok ? 0 : -1;
}

extern "C" int qdb_gpu_join_probe_detached(const QdbRowJoinKernels* kernels, void* arena, const TRowSet* const* batches,
    QdbPairBuffer* const* pairs, int64_t n, int64_t* counts, QdbFusedResult* fused)
{
    clear_error();
    if (!kernels || !arena || (n > 0 && (!batches || !counts))) return fail("null argument"), -1;
    if (kernels->Slot < 0 || kernels->Slot >= kSlots) return fail("bad join slot %d", kernels->Slot), -1;
    JoinPlan* P = join_slots().get(kernels->Slot);
    if (!P) return fail("join kernels are not compiled"), -1;
    JoinBuild* J = ((RowArenaBlob*)arena)->build;
    if (!J || !J->built) return fail("detached probe before Build"), -1;
    if (J->plan != P) return fail("the join belongs to another compiled join"), -1;
    if (P->KW != 1) return fail("detached probes take one-word join keys"), -1;
    return detached_probe(P, J, batches, pairs, n, counts, fused) ? 0 : -1;
}

extern "C" int64_t qdb_gpu_join_table_info(void* arena, int64_t* out, int64_t n)
{
    clear_error();
    JoinBuild* J = arena ? ((RowArenaBlob*)arena)->build : nullptr;
    if (!J || !out) return fail("join table info: no join"), -1;
    const int64_t v[6] = {J->built ? 1 : 0, J->dense ? 1 : 0, (int64_t)J->drange, (int64_t)J->dpack_w, J->rows, J->appended};
    int64_t k = 0;
    for (; k < n && k < 6; ++k) out[k] = v[k];
    return k;
}

extern "C" int qdb_gpu_join_probe_fused_noalloc(void* arena, const TRowSet* batch, QdbFusedResult* out, int64_t max_out,
    int64_t* pairs_out)
{
    clear_error();
    if (!arena || !batch || !out || !pairs_out) return fail("null argument"), -1;
    if (!ensure_init()) return -1;
    RowArenaBlob* AB = (RowArenaBlob*)arena;
    JoinBuild* J = AB->build;
    if (!J || !J->built) return fail("no-alloc probe before Build"), -1;
    JoinPlan* P = J->plan;
    const int nf = P->spec.NumFused;
    if (nf < 1) return fail("no-alloc probe: the join has no fused aggregate"), -1;
    if (max_out < nf) return fail("no-alloc probe: %lld results requested, the join has %d", (long long)max_out, nf), -1;
    unsigned long long* part = noalloc_scratch();
    if (!part) return -1;
    QdbPairBuffer none = {0, 0, nullptr};
    unsigned long long seen = 0;
    const bool ok = P->KW == 1 ? do_probe<1>(J, batch, &none, part, &seen) : do_probe<2>(J, batch, &none, part, &seen);
    if (!ok) return -1;
    memset(out, 0, (size_t)nf * sizeof(QdbFusedResult));
    if (J->fused_ready) merge_fused_words(P, J->readback + 2, out);
    else merge_fused_words(P, kZeroFusedWords, out);
    J->fused_ready = false;
    *pairs_out = (int64_t)seen;
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
