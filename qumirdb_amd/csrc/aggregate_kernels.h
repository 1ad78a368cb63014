// Aggregate kernels and launch arguments, shared by the per-key-width
// translation units (aggregate_kw*.hip) and the host code (aggregate.hip).
#pragma once

#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>

#include "device.h"
#include "runtime.h"
#else
#include "device.h"
#endif

namespace qdbgpu {
namespace agg {

constexpr int kMaxSrc = QDB_AGG_MAX_AGGS;
constexpr int kMaxWords = 32;
constexpr int kBlock = 256;

enum WordOp : uint8_t {
    OP_ADD = 0,     // wrapping u64 add
    OP_FADD = 1,    // f64 add
    OP_MIN_I64 = 2,
    OP_MAX_I64 = 3,
    OP_MIN_U64 = 4,
    OP_MAX_U64 = 5,
    // i128 MIN/MAX: the LO word carries the op, then the HI word and a lock
    // word; applied under the lock, never pre-aggregated.
    OP_PAIR_MIN = 6,
    OP_PAIR_MAX = 7,
    OP_PAIR_HI = 8,
    OP_LOCK = 9,
};

enum WordKind : uint8_t {
    WK_ROWS = 0,       // 1 per row
    WK_VALID = 1,      // 1 if the source is valid
    WK_SUM64 = 2,      // i64 value
    WK_I128_P0 = 3,    // bits [0,32) of the i128 value
    WK_I128_P1 = 4,    // bits [32,64)
    WK_I128_P2 = 5,    // bits [64,128)
    WK_LIMB = 6,       // limb `part` of the f64 value
    WK_OUTSIDE = 7,    // f64 value outside the limb window
    WK_MIN_I64 = 8,
    WK_MAX_I64 = 9,
    WK_MIN_F64 = 10,
    WK_MAX_F64 = 11,
    WK_PAIR_LO = 12,
    WK_PAIR_HI = 13,
    WK_LOCK = 14,
};

struct WordDesc {
    int32_t kind;
    int32_t op;
    int32_t src;
    int32_t part;
};

struct KeyFieldDesc {
    int32_t slot;      // device column slot
    int32_t kind;
    int32_t width;     // bytes of the packed value (16 for i128)
    int32_t nullable;
    int32_t word;      // packed word index
    int32_t shift;     // bit offset inside the word
    int32_t null_word;
    int32_t null_bit;
};

struct SrcDesc {
    int32_t slot;      // device column slot, or -1
    int32_t prog;      // computed program index, or -1
};

// Everything an update launch needs, passed by value.
struct AggArgs {
    DevCols cols;
    int64_t nrows;
    int64_t row_base;          // first row of this launch
    const uint8_t* sel;
    const uint32_t* row_index; // replay: rows to apply (row_base-relative)
    const DevProgram* pred;
    const DevProgram* progs;
    int nkeys;
    KeyFieldDesc keys[QDB_AGG_MAX_KEYS];
    int nsrc;
    SrcDesc src[kMaxSrc];
    int nwords;
    WordDesc words[kMaxWords];
    int lds_slots;             // power of two
    int lds_groups;            // max groups in LDS
    // Copies of the LDS table's state words (1: none): lane l applies its
    // row to copy l % lds_copies, so that the lanes of a wave whose rows
    // share a group spread their atomics over that many addresses; the
    // copies are combined when the table is flushed. Copy c >= 1 of group
    // g's word w is at extra + (c - 1) * lds_copy_stride + g * (nwords | 1)
    // + w (an odd number of words per group and per copy: one word of
    // different groups or copies falls in different LDS banks).
    int lds_copies;
    int lds_copy_stride;
    int match_rounds;
    int has_pairs;             // i128 MIN/MAX words present
    // HBM table: one record per slot, `rstride` words apart (KW key words,
    // then the state words, set to their identities when the table is
    // made). One-word keys are published by a 64-bit CAS on the key word
    // (kEmptyKey = empty slot; the key equal to kEmptyKey has the extra
    // record at slot cmask + 1); wider keys through a control word per slot.
    unsigned long long* ctrl;  // KW > 1: 0 empty, 1 being filled, tag | 2 ready
    unsigned int* occ;         // KW == 1: one bit per slot that holds a group (the finalize scans it)
    uint64_t cmask;
    uint64_t* recs;
    int rstride;
    unsigned int* counters;    // [0] groups, [1] overflow rows, [2] overflow records, [3] extra record used
    // overflow
    uint32_t* ov_rows;
    uint64_t* ov_recs;
    unsigned int ov_rec_cap;
    // Partial records (null: off). A workgroup whose LDS table ends with at
    // most part_cap groups writes them as records (key words, state words)
    // to its own slot, part_recs + blockIdx.x * part_cap records, and their
    // count to part_n[blockIdx.x] (0 when it moved them into the HBM table
    // itself); a second launch combines the slots. Device-scope atomics on
    // one cache line serialize at about 10 ns each, so a few groups' words
    // updated by every workgroup of a large grid would take longer than
    // the rows.
    uint64_t* part_recs;
    unsigned int* part_n;
    int part_cap;
};

__device__ __forceinline__ uint64_t identity_of(uint8_t op)
{
    switch (op) {
    case OP_MIN_I64: return 0x7fffffffffffffffull;
    case OP_MAX_I64: return 0x8000000000000000ull;
    case OP_MIN_U64: return ~0ull;
    case OP_PAIR_MIN: return ~0ull;
    case OP_PAIR_HI: return 0ull;  // set per reducer at insert (see pair_hi_identity)
    default: return 0ull;
    }
}

// Initial value of a new group's word.
__device__ __forceinline__ uint64_t word_init(const WordDesc& d)
{
    if (d.kind == WK_PAIR_LO) return d.part ? 0ull : ~0ull;
    if (d.kind == WK_PAIR_HI) return d.part ? 0x8000000000000000ull : 0x7fffffffffffffffull;
    if (d.kind == WK_LOCK) return 0ull;
    return identity_of(d.op);
}

__device__ __forceinline__ uint64_t combine(uint8_t op, uint64_t a, uint64_t b)
{
    switch (op) {
    case OP_ADD: return a + b;
    case OP_FADD: return (uint64_t)__double_as_longlong(__longlong_as_double((long long)a) + __longlong_as_double((long long)b));
    case OP_MIN_I64: return (int64_t)a < (int64_t)b ? a : b;
    case OP_MAX_I64: return (int64_t)a > (int64_t)b ? a : b;
    case OP_MIN_U64: return a < b ? a : b;
    default: return a > b ? a : b;
    }
}

__device__ __forceinline__ void atomic_apply_lds(uint8_t op, uint64_t* p, uint64_t v)
{
    if (v == identity_of(op) && op != OP_FADD) return;
    switch (op) {
    case OP_ADD: atomicAdd((unsigned long long*)p, (unsigned long long)v); break;
    case OP_FADD:
        if (v != 0) atomicAdd((double*)p, __longlong_as_double((long long)v));
        break;
    case OP_MIN_I64: atomicMin((long long*)p, (long long)v); break;
    case OP_MAX_I64: atomicMax((long long*)p, (long long)v); break;
    case OP_MIN_U64: atomicMin((unsigned long long*)p, (unsigned long long)v); break;
    default: atomicMax((unsigned long long*)p, (unsigned long long)v); break;
    }
}

__device__ __forceinline__ void atomic_apply_global(uint8_t op, uint64_t* p, uint64_t v)
{
    // Same operations; device-scope atomics on HBM. A MIN/MAX word only
    // moves one way, so a plain read that already beats v (even a stale one,
    // which can only be further from the final value) makes the atomic
    // unnecessary; the record's line is in cache from the key lookup.
    if (op == OP_MIN_I64 || op == OP_MAX_I64 || op == OP_MIN_U64 || op == OP_MAX_U64) {
        const uint64_t cur = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        if (combine(op, cur, v) == cur) return;
    }
    atomic_apply_lds(op, p, v);
}

template <int KW>
__device__ __forceinline__ bool keys_equal(const uint64_t (&a)[KW], const uint64_t (&b)[KW])
{
    bool eq = true;
#pragma unroll
    for (int w = 0; w < KW; ++w) eq = eq && a[w] == b[w];
    return eq;
}

// ORs key field `kf` of a row into the packed key.
template <int KW>
__device__ __forceinline__ void pack_key_field(const KeyFieldDesc& kf, const DevCol& c, int64_t row, uint64_t (&key)[KW])
{
    const bool valid = col_valid(c, row);
    if (kf.nullable && !valid) {
#pragma unroll
        for (int w = 0; w < KW; ++w) {
            if (w == kf.null_word) key[w] |= 1ull << kf.null_bit;
        }
        return;
    }
    uint64_t lo, hi = 0;
    if (kf.kind == QDB_KIND_FLOAT) {
        lo = kf.width == 4 ? (uint64_t)__float_as_uint(((const float*)c.data)[row]) : canonical_f64_bits(((const double*)c.data)[row]);
        if (kf.width == 4) {
            uint32_t b = (uint32_t)lo;
            if ((b & 0x7fffffffu) == 0) b = 0;
            else if ((b & 0x7f800000u) == 0x7f800000u && (b & 0x007fffffu) != 0) b = 0x7fc00000u;
            lo = b;
        }
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

// The packed key of a row.
template <int KW>
__device__ __forceinline__ void row_key(const AggArgs& A, int64_t row, uint64_t (&key)[KW])
{
#pragma unroll
    for (int w = 0; w < KW; ++w) key[w] = 0;
    QDB_DESC_LOOP
    for (int f = 0; f < QDB_AGG_MAX_KEYS; ++f) {
        if (f >= A.nkeys) break;
        const KeyFieldDesc kf = A.keys[f];
        pack_key_field<KW>(kf, A.cols.c[kf.slot], row, key);
    }
}

// The exact-summation limbs of an f64 value (all 0 and *outside set when it
// falls outside the limb window).
__device__ __forceinline__ void f64_limbs(double f, int64_t (&limb)[kLimbs], bool* outside)
{
#pragma unroll
    for (int k = 0; k < kLimbs; ++k) limb[k] = f64_limb(f, k, outside);
    if (*outside) {
#pragma unroll
        for (int k = 0; k < kLimbs; ++k) limb[k] = 0;
    }
}

// The contribution of a valid argument value to state word `d`.
__device__ __forceinline__ uint64_t word_contrib(const WordDesc& d, const DVal& v, const int64_t (&limb)[kLimbs], bool outside)
{
    const u128 u = (u128)v.i;
    switch (d.kind) {
    case WK_VALID: return 1;
    case WK_SUM64:
    case WK_MIN_I64:
    case WK_MAX_I64: return (uint64_t)u;
    case WK_I128_P0: return (uint64_t)u & 0xffffffffull;
    case WK_I128_P1: return (uint64_t)u >> 32;
    case WK_I128_P2: return (uint64_t)(u >> 64);
    case WK_LIMB: {
        // Constant indices only: a limb picked by a run-time index would put
        // the array in scratch memory.
        static_assert(kLimbs == 4, "limb selection below");
        const int64_t l = d.part == 0 ? limb[0] : d.part == 1 ? limb[1] : d.part == 2 ? limb[2] : limb[3];
        return (uint64_t)l;
    }
    case WK_OUTSIDE: return outside ? (uint64_t)__double_as_longlong(v.f) : 0ull;
    case WK_MIN_F64: return f64_min_key(v.f);
    case WK_MAX_F64: return f64_max_key(v.f);
    case WK_PAIR_LO: return (uint64_t)u;
    case WK_PAIR_HI: return (uint64_t)(u >> 64);
    default: return 0;
    }
}

// A column value whose kind and width are constants (kernels specialized at
// run time). The load is unconditional: a NULL row's slot is read and not
// used, so no wait sits inside a branch.
template <int kKind, int kWidth>
__device__ __forceinline__ DVal read_col_t(const DevCol& c, int64_t row)
{
    DevCol k = c;
    k.kind = kKind;
    k.width = kWidth;
    DVal v;
    v.valid = col_valid(k, row);
    v.i = 0;
    v.f = 0.0;
    if (kKind == QDB_KIND_FLOAT) {
        v.cls = QDB_REG_F64;
        v.f = col_f64(k, row);
    } else {
        v.cls = kKind == QDB_KIND_BOOL ? QDB_REG_BOOL : QDB_REG_INT;
        v.i = col_int(k, row);
    }
    return v;
}

// A row's contribution to every state word; false when the row does not
// take part (not selected, or rejected by the fused predicate). Each
// argument is read (or computed) once, by one copy of the column reader and
// the interpreter, and stored into the words that take it; the word loop is
// unrolled so that `c` stays in registers.
template <int KW, int S>
__device__ __forceinline__ bool row_contrib(const AggArgs& A, int64_t row, uint64_t (&key)[KW], uint64_t (&c)[S])
{
    if (A.sel && !A.sel[row]) return false;
    if (A.pred) {
        DVal p = eval_program<true>(A.pred, A.cols, row);
        if (!(p.valid && p.i != 0)) return false;
    }
#pragma unroll
    for (int w = 0; w < S; ++w) {
        c[w] = 0;
        if (w < A.nwords) c[w] = A.words[w].kind == WK_ROWS ? 1ull : word_init(A.words[w]);
    }
    QDB_DESC_LOOP
    for (int j = 0; j < kMaxSrc; ++j) {
        if (j >= A.nsrc) break;
        const SrcDesc sd = A.src[j];
        const DVal v = sd.prog >= 0 ? eval_program<true>(A.progs + sd.prog, A.cols, row) : read_col(A.cols.c[sd.slot], row);
        if (!v.valid) continue;  // the words keep their identities
        int64_t limb[kLimbs];
        bool outside = false;
#pragma unroll
        for (int k = 0; k < kLimbs; ++k) limb[k] = 0;
        if (v.cls == QDB_REG_F64) f64_limbs(v.f, limb, &outside);
#pragma unroll
        for (int w = 0; w < S; ++w) {
            if (w >= A.nwords) continue;
            const WordDesc d = A.words[w];
            if (d.src != j) continue;
            if (d.kind == WK_ROWS || d.kind == WK_LOCK) continue;
            c[w] = word_contrib(d, v, limb, outside);
        }
    }
    row_key<KW>(A, row, key);
    return true;
}

// The update path reads the state-word layout and the row's contribution
// through a policy: GenericPolicy reads the descriptors in the launch
// arguments; a policy generated for one aggregate (at run time) gives them
// as constants and computes the contribution with straight-line code.
struct GenericPolicy {
    __device__ static __forceinline__ int nwords(const AggArgs& A) { return A.nwords; }
    __device__ static __forceinline__ uint8_t op(const AggArgs& A, int w) { return (uint8_t)A.words[w].op; }
    __device__ static __forceinline__ uint64_t init(const AggArgs& A, int w) { return word_init(A.words[w]); }
    __device__ static __forceinline__ bool has_pairs(const AggArgs& A) { return A.has_pairs != 0; }
    __device__ static __forceinline__ int match_rounds(const AggArgs& A) { return A.match_rounds; }
    template <int KW, int S>
    __device__ static __forceinline__ bool contrib(const AggArgs& A, int64_t row, uint64_t (&key)[KW], uint64_t (&c)[S])
    {
        return row_contrib<KW, S>(A, row, key, c);
    }
};

// ---------------------------------------------------------------------------
// HBM table
// ---------------------------------------------------------------------------

// Lookups never wait for another lane. A slot that another lane is still
// filling makes the lookup return kBusy, and the caller retries in a loop
// that the whole wave runs together (locate, apply_pairs). A lane that spun
// inside the lookup could wait forever on a lane of its own wave: the
// compiler may move the code that publishes an insert after the loop in
// which the other lanes wait.
constexpr int64_t kNoRoom = -1;
constexpr int64_t kBusy = -2;

// The empty marker of a one-word key slot.
constexpr unsigned long long kEmptyKey = 0x8000000000000000ull;
// Slots a lookup walks before it gives up (the row or record is then kept
// for the replay after the table has grown).
constexpr int kMaxProbe = 256;

struct LdsTable {
    unsigned int* ctrl;
    unsigned int* count;
    uint64_t* recs;
    uint64_t* extra;          // copies 1.. of the state words (lds_copies > 1)
    unsigned int* inserted;   // HBM groups this workgroup inserted
    int fresh;                // this thread's last HBM lookup inserted its key
};

// Finds or inserts `key` in the HBM table: the slot of its record, kNoRoom
// when kMaxProbe slots hold other keys, or kBusy. Every state word starts at
// its identity, so an inserted key needs no further publication: every
// update is an atomic combine. One-word keys: the key word itself is the
// CAS target, and keys never change once set, so relaxed loads suffice (a
// stale empty marker only sends the lane to the CAS, which sees the truth).
template <int KW, int S>
__device__ __forceinline__ int64_t global_find_or_insert(const AggArgs& A, LdsTable& T, const uint64_t (&key)[KW])
{
    const uint64_t h = key_place_hash<KW>(key);
    uint64_t slot = h & A.cmask;
    if constexpr (KW == 1) {
        if (key[0] == kEmptyKey) {
            if (atomicCAS(&A.counters[3], 0u, 1u) == 0u) atomicAdd(T.inserted, 1u);
            return (int64_t)(A.cmask + 1);
        }
        for (int p = 0; p < kMaxProbe; ++p) {
            unsigned long long* kp = (unsigned long long*)(A.recs + slot * (uint64_t)A.rstride);
            unsigned long long c = __hip_atomic_load(kp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (c == kEmptyKey) {
                c = atomicCAS(kp, kEmptyKey, (unsigned long long)key[0]);
                if (c == kEmptyKey) {
                    atomicAdd(T.inserted, 1u);
                    atomicOr(&A.occ[slot >> 5], 1u << (slot & 31u));
                    T.fresh = 1;
                    return (int64_t)slot;
                }
            }
            if (c == (unsigned long long)key[0]) return (int64_t)slot;
            slot = (slot + 1) & A.cmask;
        }
        return kNoRoom;
    } else {
    const unsigned long long tag = (unsigned long long)(h >> 32) << 32;
    for (int p = 0; p < kMaxProbe; ++p) {
        unsigned long long c = __hip_atomic_load(&A.ctrl[slot], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        uint64_t* r = A.recs + slot * (uint64_t)A.rstride;
        if (c == 0ull) {
            c = atomicCAS(&A.ctrl[slot], 0ull, 1ull);
            if (c == 0ull) {
#pragma unroll
                for (int w = 0; w < KW; ++w) __hip_atomic_store(&r[w], key[w], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_store(&A.ctrl[slot], tag | 2ull, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
                atomicAdd(T.inserted, 1u);
                T.fresh = 1;
                return (int64_t)slot;
            }
        }
        if (c == 1ull) return kBusy;  // being filled
        if ((c & 0xffffffff00000000ull) == tag) {
            // The key words were written before the control word: read them
            // after an acquire of it (only on a tag match).
            (void)__hip_atomic_load(&A.ctrl[slot], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
            uint64_t k2[KW];
#pragma unroll
            for (int w = 0; w < KW; ++w) k2[w] = __hip_atomic_load(&r[w], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (keys_equal<KW>(k2, key)) return (int64_t)slot;
        }
        slot = (slot + 1) & A.cmask;
    }
    return kNoRoom;
    }
}

// One attempt at an i128 MIN/MAX under the group's lock word; false when
// the lock is taken.
__device__ __forceinline__ bool pair_try(uint64_t* lo, bool is_max, uint64_t vlo, uint64_t vhi)
{
    uint64_t* hi = lo + 1;
    unsigned long long* lock = (unsigned long long*)(lo + 2);
    if (atomicCAS(lock, 0ull, 1ull) != 0ull) return false;
    i128 cur = (i128)(((u128)__hip_atomic_load(hi, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) << 64)
        | (u128)__hip_atomic_load(lo, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    i128 v = (i128)(((u128)vhi << 64) | (u128)vlo);
    if (is_max ? v > cur : v < cur) {
        __hip_atomic_store(lo, vlo, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(hi, vhi, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __hip_atomic_store(lock, 0ull, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    return true;
}

// ---------------------------------------------------------------------------
// LDS table: ctrl[lds_slots] (u32: 0 empty, 1 being filled, else lgid + 2),
// a group counter, then dense records of KW + nwords words.
// ---------------------------------------------------------------------------

// The LDS group of `key`, kNoRoom when it is absent and the table is full,
// or kBusy.
template <int KW, int S, class P>
__device__ __forceinline__ int lds_find_or_insert(const AggArgs& A, LdsTable& T, const uint64_t (&key)[KW])
{
    const uint64_t h = key_place_hash<KW>(key);
    const unsigned int mask = (unsigned int)A.lds_slots - 1u;
    unsigned int slot = (unsigned int)(h >> 40) & mask;
    const int R = KW + P::nwords(A);
    for (unsigned int probes = 0; probes <= mask; ++probes) {
        unsigned int c = __hip_atomic_load(&T.ctrl[slot], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
        if (c == 0u) {
            if (__hip_atomic_load(T.count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) >= (unsigned int)A.lds_groups) {
                return (int)kNoRoom;
            }
            c = atomicCAS(&T.ctrl[slot], 0u, 1u);
            if (c == 0u) {
                unsigned int lg = atomicAdd(T.count, 1u);
                if (lg >= (unsigned int)A.lds_groups) {
                    __hip_atomic_store(&T.ctrl[slot], 0u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
                    return (int)kNoRoom;
                }
                uint64_t* r = T.recs + (size_t)lg * R;
#pragma unroll
                for (int w = 0; w < KW; ++w) r[w] = key[w];
                for (int w = 0; w < P::nwords(A); ++w) r[KW + w] = P::init(A, w);
                for (int c = 1; c < A.lds_copies; ++c) {
                    uint64_t* e = T.extra + (size_t)(c - 1) * A.lds_copy_stride + (size_t)lg * (P::nwords(A) | 1);
                    for (int w = 0; w < P::nwords(A); ++w) e[w] = P::init(A, w);
                }
                __hip_atomic_store(&T.ctrl[slot], lg + 2u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
                return (int)lg;
            }
        }
        if (c == 1u) return (int)kBusy;
        const unsigned int lg = c - 2u;
        const uint64_t* r = T.recs + (size_t)lg * R;
        bool eq = true;
#pragma unroll
        for (int w = 0; w < KW; ++w) eq = eq && r[w] == key[w];
        if (eq) return (int)lg;
        slot = (slot + 1u) & mask;
    }
    return (int)kNoRoom;
}

enum Where : int { AT_NONE = 0, AT_LDS = 1, AT_HBM = 2, AT_BUSY = 3 };

// One lookup attempt: the workgroup's LDS table first, then the HBM table.
template <int KW, int S, class P>
__device__ __forceinline__ int locate_once(const AggArgs& A, LdsTable& T, bool use_lds, const uint64_t (&key)[KW],
    int64_t* id)
{
    if (use_lds) {
        const int lg = lds_find_or_insert<KW, S, P>(A, T, key);
        if (lg >= 0) {
            *id = lg;
            return AT_LDS;
        }
        if (lg == (int)kBusy) return AT_BUSY;
    }
    const int64_t gid = global_find_or_insert<KW, S>(A, T, key);
    if (gid >= 0) {
        *id = gid;
        return AT_HBM;
    }
    return gid == kBusy ? AT_BUSY : AT_NONE;
}

// Wave-uniform: every lane of the wave calls it, lanes with `want` false
// only take part. Returns where the key's group is (AT_NONE: no room).
template <int KW, int S, class P>
__device__ __forceinline__ int locate(const AggArgs& A, LdsTable& T, bool use_lds, bool want, const uint64_t (&key)[KW],
    int64_t* id)
{
    int where = want ? AT_BUSY : AT_NONE;
    if constexpr (KW == 1) {
        // One-word keys never see a slot being filled in HBM; the LDS table
        // may still report one.
        if (!use_lds) {
            if (want) where = locate_once<KW, S, P>(A, T, false, key, id);
            return where;
        }
    }
    while (__any(where == AT_BUSY)) {
        if (where == AT_BUSY) where = locate_once<KW, S, P>(A, T, use_lds, key, id);
    }
    return where;
}

// Combines a contribution into the group's words, except the i128 MIN/MAX
// pairs (apply_pairs).
template <int KW, int S, class P>
__device__ __forceinline__ void apply_words(const AggArgs& A, LdsTable& T, int where, int64_t id, const uint64_t (&c)[S])
{
    // Separate paths for the two tables: LDS atomics for the workgroup's
    // table, global atomics for the HBM record (a pointer that could be
    // either would make every access a flat one).
    if (where == AT_LDS) {
        uint64_t* r = T.recs + (size_t)id * (KW + P::nwords(A)) + KW;
        if (A.lds_copies > 1) {
            const int c = lane_id() % A.lds_copies;
            if (c) r = T.extra + (size_t)(c - 1) * A.lds_copy_stride + (size_t)id * (P::nwords(A) | 1);
        }
#pragma unroll
        for (int w = 0; w < S; ++w) {
            if (w >= P::nwords(A)) continue;
            const int op = P::op(A, w);
            if (op == OP_PAIR_MIN || op == OP_PAIR_MAX || op == OP_PAIR_HI || op == OP_LOCK) continue;
            // A MIN/MAX word only moves one way: a plain read that already
            // beats the contribution (even a stale one) makes the atomic
            // unnecessary -- after a group's first rows, most rows.
            if (op == OP_MIN_I64 || op == OP_MAX_I64 || op == OP_MIN_U64 || op == OP_MAX_U64) {
                const uint64_t cur = __hip_atomic_load(&r[w], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                if (combine((uint8_t)op, cur, c[w]) == cur) continue;
            }
            atomic_apply_lds((uint8_t)op, &r[w], c[w]);
        }
    } else {
        uint64_t* r = A.recs + (uint64_t)id * (uint64_t)A.rstride + KW;
#pragma unroll
        for (int w = 0; w < S; ++w) {
            if (w >= P::nwords(A)) continue;
            const int op = P::op(A, w);
            if (op == OP_PAIR_MIN || op == OP_PAIR_MAX || op == OP_PAIR_HI || op == OP_LOCK) continue;
            atomic_apply_global((uint8_t)op, &r[w], c[w]);
        }
    }
}

// Wave-uniform: the i128 MIN/MAX pairs of HBM group `gid` for the lanes
// with `want`, each under the group's lock (pairs never live in LDS).
template <int KW, int S, class P>
__device__ __forceinline__ void apply_pairs(const AggArgs& A, bool want, int64_t gid, const uint64_t (&c)[S])
{
#pragma unroll
    for (int w = 0; w + 1 < S; ++w) {
        if (w >= P::nwords(A)) continue;
        const int op = P::op(A, w);
        if (op != OP_PAIR_MIN && op != OP_PAIR_MAX) continue;
        bool pending = want;
        while (__any(pending)) {
            if (pending) {
                uint64_t* r = A.recs + (size_t)gid * (uint64_t)A.rstride + KW;
                if (pair_try(&r[w], op == OP_PAIR_MAX, c[w], c[w + 1])) pending = false;
            }
        }
    }
}

// Records a row for the replay after the HBM table has grown.
__device__ __forceinline__ void spill_row(const AggArgs& A, int64_t row)
{
    const unsigned int pos = atomicAdd(&A.counters[1], 1u);
    A.ov_rows[pos] = (uint32_t)row;
}

// Wave-uniform: applies a contribution (the lanes with `want`) to the LDS
// table or the HBM table; false for a lane whose group found no room.
template <int KW, int S, class P>
__device__ __forceinline__ bool apply_contrib(const AggArgs& A, LdsTable& T, bool use_lds, bool want,
    const uint64_t (&key)[KW], const uint64_t (&c)[S])
{
    int64_t id = 0;
    const int where = locate<KW, S, P>(A, T, use_lds, want, key, &id);
    if (want && where != AT_NONE) apply_words<KW, S, P>(A, T, where, id, c);
    if (P::has_pairs(A)) apply_pairs<KW, S, P>(A, want && where == AT_HBM, id, c);
    return !want || where != AT_NONE;
}

// Wave-uniform: applies group records (key + state words) to the HBM table,
// or keeps them for the replay after the table has grown.
template <int KW, int S, class P>
__device__ __forceinline__ void global_upsert_or_spill(const AggArgs& A, LdsTable& T, bool want, const uint64_t (&key)[KW],
    const uint64_t (&c)[S])
{
    if (apply_contrib<KW, S, P>(A, T, false, want, key, c)) return;
    unsigned int pos = atomicAdd(&A.counters[2], 1u);
    if (pos < A.ov_rec_cap) {
        uint64_t* r = A.ov_recs + (size_t)pos * (KW + P::nwords(A));
#pragma unroll
        for (int w = 0; w < KW; ++w) r[w] = key[w];
#pragma unroll
        for (int w = 0; w < S; ++w) {
            if (w < P::nwords(A)) r[KW + w] = c[w];
        }
    }
}

// Butterfly reduction of one word over the lanes in `in` (others give the
// identity); every lane ends with the combined value.
__device__ __forceinline__ uint64_t wave_reduce(uint8_t op, uint64_t v, bool in)
{
    uint64_t x = in ? v : identity_of(op);
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
        x = combine(op, x, shfl_xor_u64(x, off));
    }
    return x;
}

// The workgroup's LDS table laid out for A (records, then the control
// words, then the group count).
template <int KW, class P>
__device__ __forceinline__ LdsTable lds_table_at(const AggArgs& A, uint64_t* lds_raw, unsigned int* inserted)
{
    LdsTable T;
    T.recs = lds_raw;
    T.ctrl = (unsigned int*)(lds_raw + (size_t)A.lds_groups * (KW + P::nwords(A)));
    T.count = T.ctrl + A.lds_slots;
    // 8-byte aligned past the counter (lds_slots is even).
    T.extra = (uint64_t*)(T.ctrl + A.lds_slots + 2);
    T.inserted = inserted;
    T.fresh = 0;
    return T;
}

// Clears the LDS table and the inserted-group counter (every thread; ends
// with a barrier).
template <int KW, class P>
__device__ __forceinline__ void lds_table_reset(const AggArgs& A, const LdsTable& T, bool use_lds)
{
    if (threadIdx.x == 0) *T.inserted = 0u;
    if (use_lds) {
        for (int i = threadIdx.x; i < A.lds_slots; i += blockDim.x) T.ctrl[i] = 0u;
        if (threadIdx.x == 0) *T.count = 0u;
    }
    __syncthreads();
}

// Rows idx in [wave_first + lane, end) stepping by `step` (row = row_base +
// idx, or row_index[idx]), pre-aggregated in the LDS table and then in the
// HBM table. Every lane of a wave runs the same number of iterations: the
// ballots and the retry loops inside see the whole wave.
template <int KW, int S, class P>
__device__ __forceinline__ void agg_rows(const AggArgs& A, LdsTable& T, bool use_lds, int64_t wave_first, int64_t end, int64_t step)
{
    const int lane = lane_id();
    for (int64_t base = wave_first; base < end; base += step) {
        const int64_t idx = base + lane;
        bool valid = idx < end;
        int64_t row = 0;
        if (valid) row = A.row_index ? (int64_t)A.row_index[idx] : A.row_base + idx;
        uint64_t key[KW];
        uint64_t c[S];
        if (valid) valid = P::template contrib<KW, S>(A, row, key, c);
        if (!valid) {
#pragma unroll
            for (int w = 0; w < KW; ++w) key[w] = 0;
#pragma unroll
            for (int w = 0; w < S; ++w) c[w] = 0;
        }
        // Pre-aggregation: the lanes that share the first remaining lane's
        // key are combined and applied once by that lane.
        uint64_t active = __ballot(valid);
        for (int round = 0; round < P::match_rounds(A) && active; ++round) {
            const int leader = __ffsll((unsigned long long)active) - 1;
            uint64_t lk[KW];
#pragma unroll
            for (int w = 0; w < KW; ++w) lk[w] = shfl_u64(key[w], leader);
            const bool same = valid && ((active >> lane) & 1ull) && keys_equal<KW>(key, lk);
            const uint64_t m = __ballot(same);
            uint64_t red[S];
            if (__popcll(m) > 1) {
#pragma unroll
                for (int w = 0; w < S; ++w) {
                    red[w] = w < P::nwords(A) ? wave_reduce(P::op(A, w), c[w], same) : 0;
                }
            } else {
#pragma unroll
                for (int w = 0; w < S; ++w) red[w] = c[w];
            }
            const bool ok = apply_contrib<KW, S, P>(A, T, use_lds, lane == leader, lk, red);
            // The combined contribution found no room: every row of the
            // match is replayed on its own.
            if (!__shfl((int)ok, leader) && same) spill_row(A, row);
            active &= ~m;
            if (same) valid = false;
        }
        if (!apply_contrib<KW, S, P>(A, T, use_lds, valid, key, c)) spill_row(A, row);
    }
}

// Moves the workgroup's LDS groups into the HBM table and adds the groups
// the workgroup inserted to the table's count (every thread; barriers
// inside, the last one before the table may be reused).
template <int KW, int S, class P>
__device__ __forceinline__ void lds_table_flush(const AggArgs& A, LdsTable& T, bool use_lds)
{
    __syncthreads();
    if (use_lds) {
        // Every thread runs the same number of iterations.
        const unsigned int n = *T.count < (unsigned int)A.lds_groups ? *T.count : (unsigned int)A.lds_groups;
        const int R = KW + P::nwords(A);
        const bool to_parts = A.part_recs != nullptr && n <= (unsigned int)A.part_cap;  // uniform
        if (A.part_recs != nullptr && threadIdx.x == 0) A.part_n[blockIdx.x] = to_parts ? n : 0u;
        for (unsigned int b = 0; b < n; b += blockDim.x) {
            const unsigned int lg = b + threadIdx.x;
            const bool want = lg < n;
            uint64_t key[KW];
            uint64_t c[S];
            const uint64_t* r = T.recs + (size_t)(want ? lg : 0u) * R;
#pragma unroll
            for (int w = 0; w < KW; ++w) key[w] = r[w];
#pragma unroll
            for (int w = 0; w < S; ++w) c[w] = w < P::nwords(A) ? r[KW + w] : 0;
            // The copies' words combined into the first copy's.
            for (int k = 1; k < A.lds_copies; ++k) {
                const uint64_t* e = T.extra + (size_t)(k - 1) * A.lds_copy_stride + (size_t)(want ? lg : 0u) * (P::nwords(A) | 1);
#pragma unroll
                for (int w = 0; w < S; ++w) {
                    if (w < P::nwords(A)) c[w] = combine((uint8_t)P::op(A, w), c[w], e[w]);
                }
            }
            if (to_parts) {
                if (want) {
                    uint64_t* r = A.part_recs + ((size_t)blockIdx.x * (size_t)A.part_cap + lg) * R;
#pragma unroll
                    for (int w = 0; w < KW; ++w) r[w] = key[w];
#pragma unroll
                    for (int w = 0; w < S; ++w) {
                        if (w < P::nwords(A)) r[KW + w] = c[w];
                    }
                }
            } else {
                global_upsert_or_spill<KW, S, P>(A, T, want, key, c);
            }
        }
        __syncthreads();
    } else if (A.part_recs != nullptr && threadIdx.x == 0) {
        A.part_n[blockIdx.x] = 0u;
    }
    // One atomic per workgroup for the groups it inserted.
    if (threadIdx.x == 0 && *T.inserted) atomicAdd(&A.counters[0], *T.inserted);
    __syncthreads();
}

// The update of one launch: every selected row of [row_base, row_base +
// nrows) (or of row_index), pre-aggregated in the workgroup's LDS table and
// then in the HBM table. The generic kernel below reads the descriptors
// from its arguments; kernels specialized at run time call it with the
// descriptors set to constants.
template <int KW, int S, class P>
__device__ __forceinline__ void agg_update_body(const AggArgs& A)
{
    extern __shared__ uint64_t lds_raw[];
    __shared__ unsigned int s_inserted;
    const bool use_lds = A.lds_slots > 0 && A.row_index == nullptr;
    LdsTable T = lds_table_at<KW, P>(A, lds_raw, &s_inserted);
    lds_table_reset<KW, P>(A, T, use_lds);
    const int64_t start = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    agg_rows<KW, S, P>(A, T, use_lds, start - lane_id(), A.nrows, (int64_t)gridDim.x * blockDim.x);
    lds_table_flush<KW, S, P>(A, T, use_lds);
}

constexpr int kMaxPartsPerReducer = 64;  // slots one workgroup of the combining launch takes at most

// The second launch of an update with partial records: workgroup b
// combines the records of the update's workgroups [b * per, (b + 1) * per)
// (slot p holds part_n[p] records of KW + nwords words, part_cap records
// apart) in its LDS table and moves its groups into the HBM table; a record
// that finds no room there is kept for the replay after growth. A is the
// update's, without part_recs.
template <int KW, int S, class P>
__device__ __forceinline__ void agg_reduce_parts_body(const AggArgs& A, const uint64_t* parts, const unsigned int* part_n, int nparts,
    int per)
{
    extern __shared__ uint64_t lds_raw[];
    __shared__ unsigned int s_inserted;
    __shared__ unsigned int s_first[kMaxPartsPerReducer + 1];  // exclusive prefix of the slots' record counts
    const bool use_lds = A.lds_slots > 0;
    LdsTable T = lds_table_at<KW, P>(A, lds_raw, &s_inserted);
    lds_table_reset<KW, P>(A, T, use_lds);
    const int R = KW + P::nwords(A);
    const int p0 = (int)blockIdx.x * per;
    const int np = p0 + per < nparts ? per : (nparts > p0 ? nparts - p0 : 0);
    if (threadIdx.x == 0) {
        unsigned int t = 0;
        for (int q = 0; q < np; ++q) {
            s_first[q] = t;
            t += part_n[p0 + q];
        }
        s_first[np] = t;
    }
    __syncthreads();
    const unsigned int total = s_first[np];
    const uint64_t cap = (uint64_t)A.part_cap;
    const int lane = lane_id();
    // Only the records the slots hold, one per thread; every lane of a wave
    // runs the same number of iterations (the lookups inside are
    // wave-uniform).
    for (unsigned int base = threadIdx.x - (unsigned int)lane; base < total; base += blockDim.x) {
        const unsigned int k = base + (unsigned int)lane;
        const bool want = k < total;
        // The slot of record k: the last q with s_first[q] <= k.
        int lo = 0, hi = np - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (s_first[mid] <= k) lo = mid;
            else hi = mid - 1;
        }
        const unsigned int j = want ? k - s_first[lo] : 0u;
        const uint64_t* r = parts + ((uint64_t)(p0 + lo) * cap + j) * (uint64_t)R;
        uint64_t key[KW];
        uint64_t c[S];
#pragma unroll
        for (int w = 0; w < KW; ++w) key[w] = want ? r[w] : 0ull;
#pragma unroll
        for (int w = 0; w < S; ++w) c[w] = (want && w < P::nwords(A)) ? r[KW + w] : 0ull;
        if (!apply_contrib<KW, S, P>(A, T, use_lds, want, key, c)) {
            const unsigned int pos = atomicAdd(&A.counters[2], 1u);
            if (pos < A.ov_rec_cap) {
                uint64_t* o = A.ov_recs + (size_t)pos * R;
#pragma unroll
                for (int w = 0; w < KW; ++w) o[w] = key[w];
#pragma unroll
                for (int w = 0; w < S; ++w) {
                    if (w < P::nwords(A)) o[KW + w] = c[w];
                }
            }
        }
    }
    lds_table_flush<KW, S, P>(A, T, use_lds);
}

// One work item of a batched update: rows [row0, row1) of instance `inst`.
struct BatchItem {
    int32_t inst;
    int32_t pad;
    int64_t row0;
    int64_t row1;
};

// Many aggregate instances of one layout in one launch: persistent
// workgroups take work items from a device queue (one atomic per item);
// a workgroup keeps one instance's groups in its LDS table while its items
// stay on that instance and moves them to that instance's HBM table when
// it changes instance and at the end. Items come instance by instance.
template <int KW, int S, class P>
__device__ __forceinline__ void agg_batch_body(const AggArgs* __restrict__ insts, const BatchItem* __restrict__ items, int nitems,
    unsigned int* head)
{
    extern __shared__ uint64_t lds_raw[];
    __shared__ unsigned int s_inserted;
    __shared__ int s_item;
    LdsTable T = lds_table_at<KW, P>(insts[0], lds_raw, &s_inserted);
    int cur = -1;
    bool use_lds = false;
    for (;;) {
        if (threadIdx.x == 0) s_item = (int)atomicAdd(head, 1u);
        __syncthreads();
        const int it = __builtin_amdgcn_readfirstlane(s_item);
        __syncthreads();  // every thread has read s_item before it is written again
        if (it >= nitems) break;
        const BatchItem w = items[it];
        const int inst = __builtin_amdgcn_readfirstlane(w.inst);
        if (inst != cur) {
            if (cur >= 0) lds_table_flush<KW, S, P>(insts[cur], T, use_lds);
            cur = inst;
            use_lds = insts[cur].lds_slots > 0;
            T = lds_table_at<KW, P>(insts[cur], lds_raw, &s_inserted);
            lds_table_reset<KW, P>(insts[cur], T, use_lds);
        }
        agg_rows<KW, S, P>(insts[cur], T, use_lds, w.row0 + (int64_t)threadIdx.x - lane_id(), w.row1, (int64_t)blockDim.x);
    }
    if (cur >= 0) lds_table_flush<KW, S, P>(insts[cur], T, use_lds);
}

// ---------------------------------------------------------------------------
// Bucketed path for many groups: rows are staged (key and argument words)
// and partitioned by bucket (top bits of the key hash), each bucket is aggregated
// in one workgroup's LDS table, and the groups are merged into the HBM
// table -- with plain stores when one workgroup holds the whole bucket, as
// no other workgroup can hold one of its keys.
// ---------------------------------------------------------------------------

// A run of bucket-sorted staged rows aggregated by one workgroup; `owned`:
// the run is its bucket's only run.
struct SortedItem {
    int64_t row0;
    int64_t row1;
    int32_t owned;
    int32_t pad;
};

constexpr unsigned int kSharedItem = 0x80000000u;  // count flag: an item's records may repeat a key
constexpr int kStageUnroll = 4;                    // rows per thread per step in the bucketed kernels

// Bucket partition of the staged rows in two passes over the same
// contiguous row ranges (workgroup k: rows [k * chunk, (k + 1) * chunk)).
// Counting: per-bucket counts of the range into hist[bucket * gridDim.x +
// k]. After an exclusive scan of hist, scattering: each row's staged words
// at the next position of its bucket's run for the range. A row that does
// not take part counts in the last bucket and is not written.
// Hot keys of a bucketed update (one-word keys): the rows of the keys that a
// sample of the batch shows most often are aggregated in the scatter's LDS
// instead of being staged, sorted and merged (Zipf-like inputs put a large
// share of the rows on a few hundred keys). Each scatter workgroup writes
// its states of the hot keys as records to a slot of its own, and one wave
// per hot key combines the slots into that key's record. No other record
// of the update holds a hot key.
constexpr int kHotMax = 512;                 // hot keys at most
constexpr int kHotSlots = 4096;              // LDS hash slots of the hot set (a power of two; 1/8 full at most)
constexpr int kHotSamples = 32768;           // rows sampled, evenly spaced
constexpr int kHotSampleSlots = 8192;        // LDS hash slots of the sample's counts
constexpr unsigned int kHotMinCount = 4;     // sample count that makes a key hot
constexpr int kHotSampleBlock = 1024;        // threads of the sampling workgroup
constexpr size_t kHotSetBytes = (size_t)kHotSlots * 8 + (size_t)kHotSlots * 2;

// The LDS hot set: keys by hash slot (kEmptyKey: free) and their indexes in
// the hot list.
struct HotSet {
    unsigned long long* keys;
    unsigned short* idx;
};

__device__ __forceinline__ unsigned int hot_slot(uint64_t h)
{
    return (unsigned int)(h >> 20) & (unsigned int)(kHotSlots - 1);
}

// The hot set at `base` (8-byte aligned) filled from hot_keys[0, n) (every
// thread; ends with a barrier). The keys are distinct and none is kEmptyKey.
__device__ __forceinline__ HotSet hot_set_load(void* base, const unsigned long long* hot_keys, unsigned int n)
{
    HotSet H;
    H.keys = (unsigned long long*)base;
    H.idx = (unsigned short*)(H.keys + kHotSlots);
    for (int i = threadIdx.x; i < kHotSlots; i += blockDim.x) H.keys[i] = kEmptyKey;
    __syncthreads();
    for (unsigned int i = threadIdx.x; i < n; i += blockDim.x) {
        const unsigned long long k = hot_keys[i];
        uint64_t key1[1] = {k};
        unsigned int slot = hot_slot(key_place_hash<1>(key1));
        for (int p = 0; p < kHotSlots; ++p) {
            if (atomicCAS(&H.keys[slot], kEmptyKey, k) == kEmptyKey) {
                H.idx[slot] = (unsigned short)i;
                break;
            }
            slot = (slot + 1u) & (unsigned int)(kHotSlots - 1);
        }
    }
    __syncthreads();
    return H;
}

// The hot index of `key` (hash h), or -1. kEmptyKey marks free slots and
// is never hot.
__device__ __forceinline__ int hot_find(const HotSet& H, uint64_t key, uint64_t h)
{
    if (key == kEmptyKey) return -1;
    unsigned int slot = hot_slot(h);
    for (int p = 0; p < kHotSlots; ++p) {
        const unsigned long long k = H.keys[slot];
        if (k == key) return (int)H.idx[slot];
        if (k == kEmptyKey) return -1;
        slot = (slot + 1u) & (unsigned int)(kHotSlots - 1);
    }
    return -1;
}

// One workgroup of kHotSampleBlock threads: the keys of kHotSamples evenly
// spaced rows of [row_base, row_base + n) counted in an LDS table; every key
// seen at least kHotMinCount times goes to hot_keys (at most kHotMax), and
// their number to *hot_n. A key that finds no slot within 8 probes is not
// counted: the table fills up with a skewed sample's rare keys, and a hot
// key claims its slot among the first samples.
template <int KW, class P>
__device__ __forceinline__ void agg_hot_sample_body(const AggArgs& A, int64_t n, unsigned long long* hot_keys, unsigned int* hot_n)
{
    __shared__ unsigned long long sk[kHotSampleSlots];
    __shared__ unsigned int sc[kHotSampleSlots];
    __shared__ unsigned int s_n;
    for (int i = threadIdx.x; i < kHotSampleSlots; i += blockDim.x) {
        sk[i] = kEmptyKey;
        sc[i] = 0u;
    }
    if (threadIdx.x == 0) s_n = 0u;
    __syncthreads();
    if constexpr (KW == 1) {
        constexpr int kU = 8;  // rows per thread whose loads are in flight together
        const int per = kHotSamples / kHotSampleBlock;
        for (int k0 = 0; k0 < per; k0 += kU) {
            uint64_t key[kU][1];
            bool ok[kU];
#pragma unroll
            for (int u = 0; u < kU; ++u) {
                const int64_t sidx = (int64_t)threadIdx.x + (int64_t)(k0 + u) * kHotSampleBlock;
                const int64_t row = sidx * n / kHotSamples;
                ok[u] = P::template stage_key<KW>(A, A.row_base + row, key[u]);
            }
            // Adds `cnt` to the count of key k: a plain read finds a key
            // already in its slot, a CAS claims a free one.
            auto add = [&](unsigned long long k, unsigned int cnt) {
                uint64_t k1[1] = {k};
                unsigned int slot = (unsigned int)(key_place_hash<1>(k1) >> 40) & (unsigned int)(kHotSampleSlots - 1);
                for (int p = 0; p < 8; ++p) {
                    unsigned long long cur = sk[slot];
                    if (cur == kEmptyKey) cur = atomicCAS(&sk[slot], kEmptyKey, k);
                    if (cur == kEmptyKey || cur == k) {
                        atomicAdd(&sc[slot], cnt);
                        break;
                    }
                    slot = (slot + 1u) & (unsigned int)(kHotSampleSlots - 1);
                }
            };
            const int lane = lane_id();
#pragma unroll
            for (int u = 0; u < kU; ++u) {
                bool want = ok[u] && key[u][0] != kEmptyKey;
                // The lanes that share the first wanting lane's key count once
                // (the most frequent keys fill whole waves of the sample).
                const uint64_t wm = __ballot(want);
                if (wm) {
                    const int leader = __ffsll((unsigned long long)wm) - 1;
                    const unsigned long long lk = shfl_u64(key[u][0], leader);
                    const bool same = want && key[u][0] == lk;
                    const uint64_t sm = __ballot(same);
                    if (same) {
                        want = false;
                        if (lane == leader) add(lk, (unsigned int)__popcll(sm));
                    }
                }
                if (want) add(key[u][0], 1u);
            }
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kHotSampleSlots; i += blockDim.x) {
        if (sk[i] != kEmptyKey && sc[i] >= kHotMinCount) {
            const unsigned int pos = atomicAdd(&s_n, 1u);
            if (pos < (unsigned int)kHotMax) hot_keys[pos] = sk[i];
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) *hot_n = s_n < (unsigned int)kHotMax ? s_n : (unsigned int)kHotMax;
}

// One wave per hot key h < *hot_n: the state words of the key's records in
// the scatter workgroups' slots (parts + (b * kHotMax + h) * R, b < nblk)
// combined into out + h * R (key word first).
template <int KW, class P>
__device__ __forceinline__ void agg_hot_reduce_body(const AggArgs& A, const unsigned long long* parts, int nblk,
    const unsigned int* hot_n, unsigned long long* out)
{
    const int R = KW + P::nwords(A);
    const int lane = lane_id();
    const int nh = (int)*hot_n;
    const int waves = (int)(gridDim.x * blockDim.x / kWave);
    for (int h = (int)((blockIdx.x * blockDim.x + threadIdx.x) / kWave); h < nh; h += waves) {
        for (int w = 0; w < P::nwords(A); ++w) {
            const uint8_t op = P::op(A, w);
            uint64_t acc = identity_of(op);
            for (int b = lane; b < nblk; b += kWave) acc = combine(op, acc, parts[((size_t)b * kHotMax + h) * R + KW + w]);
#pragma unroll
            for (int o = 1; o < kWave; o <<= 1) acc = combine(op, acc, shfl_xor_u64(acc, o));
            if (lane == 0) out[(size_t)h * R + KW + w] = acc;
        }
        if (lane == 0) {
#pragma unroll
            for (int w = 0; w < KW; ++w) out[(size_t)h * R + w] = parts[(size_t)h * R + w];
        }
    }
}

// Bucket partition of the staged rows in two passes over the same
// contiguous row ranges (workgroup k: rows [k * chunk, (k + 1) * chunk)).
// Counting: per-bucket counts of the range into hist[bucket * gridDim.x +
// k]. After an exclusive scan of hist, scattering: each row's staged words
// at the next position of its bucket's run for the range. A row that does
// not take part, or whose key is hot (hot_keys non-null), counts in the
// last bucket and is not written; the scatter aggregates the hot rows in
// LDS and writes its states of the hot keys to hot_parts + blockIdx.x *
// kHotMax * (KW + nwords).
template <int KW, int W, class P>
__device__ __forceinline__ void agg_count_body(const AggArgs& A, unsigned int* hist, int shift, unsigned int nb, int64_t chunk,
    const unsigned long long* hot_keys, const unsigned int* hot_n)
{
    extern __shared__ unsigned int lh[];
    for (unsigned int b = threadIdx.x; b < nb; b += blockDim.x) lh[b] = 0u;
    const unsigned int nhot = (KW == 1 && hot_keys) ? *hot_n : 0u;
    HotSet H;
    H.keys = nullptr;
    H.idx = nullptr;
    if (nhot) H = hot_set_load(lh + ((nb + 1u) & ~1u), hot_keys, nhot);  // barrier inside
    __syncthreads();
    const int64_t r0 = blockIdx.x * chunk;
    const int64_t r1 = r0 + chunk < A.nrows ? r0 + chunk : A.nrows;
    // kStageUnroll rows per thread per step, staged (loads issued) before
    // any is used, so that their memory latencies overlap.
    for (int64_t i0 = r0 + threadIdx.x; i0 < r1; i0 += (int64_t)blockDim.x * kStageUnroll) {
        // Only the key (and predicate) columns are read here.
        uint64_t key[kStageUnroll][KW];
        bool ok[kStageUnroll];
#pragma unroll
        for (int u = 0; u < kStageUnroll; ++u) {
            const int64_t i = i0 + (int64_t)u * blockDim.x;
            ok[u] = i < r1 && P::template stage_key<KW>(A, A.row_base + (i < r1 ? i : r0), key[u]);
        }
#pragma unroll
        for (int u = 0; u < kStageUnroll; ++u) {
            if (i0 + (int64_t)u * blockDim.x >= r1) continue;
            const uint64_t h = key_place_hash<KW>(key[u]);
            bool keep = ok[u];
            if (keep && nhot) keep = hot_find(H, key[u][0], h) < 0;
            const unsigned int b = keep ? (unsigned int)(h >> shift) : nb - 1u;
            atomicAdd(&lh[b], 1u);
        }
    }
    __syncthreads();
    for (unsigned int b = threadIdx.x; b < nb; b += blockDim.x) hist[(uint64_t)b * gridDim.x + blockIdx.x] = lh[b];
}

template <int KW, int S, int W, class P>
__device__ __forceinline__ void agg_scatter_body(const AggArgs& A, const unsigned int* offs, unsigned long long* staged, int shift,
    unsigned int nb, int64_t chunk, const unsigned long long* hot_keys, const unsigned int* hot_n, unsigned long long* hot_parts)
{
    extern __shared__ unsigned int cur[];
    for (unsigned int b = threadIdx.x; b < nb; b += blockDim.x) cur[b] = offs[(uint64_t)b * gridDim.x + blockIdx.x];
    const unsigned int nhot = (KW == 1 && hot_keys) ? *hot_n : 0u;
    const int R = KW + P::nwords(A);
    HotSet H;
    H.keys = nullptr;
    H.idx = nullptr;
    struct {
        uint64_t* recs;
    } HT = {nullptr};  // the hot keys' states as LDS records (key word, state words)
    if (nhot) {
        void* hb = cur + ((nb + 1u) & ~1u);
        HT.recs = (uint64_t*)((char*)hb + kHotSetBytes);
        for (unsigned int k = threadIdx.x; k < nhot * (unsigned int)R; k += blockDim.x) {
            const unsigned int h = k / (unsigned int)R, w = k % (unsigned int)R;
            HT.recs[k] = w < (unsigned int)KW ? hot_keys[h] : P::init(A, (int)w - KW);
        }
        H = hot_set_load(hb, hot_keys, nhot);  // barrier inside
    }
    __syncthreads();
    const int64_t r0 = blockIdx.x * chunk;
    const int64_t r1 = r0 + chunk < A.nrows ? r0 + chunk : A.nrows;
    for (int64_t i0 = r0 + threadIdx.x; i0 < r1; i0 += (int64_t)blockDim.x * kStageUnroll) {
        uint64_t st[kStageUnroll][W];
        bool ok[kStageUnroll];
#pragma unroll
        for (int u = 0; u < kStageUnroll; ++u) {
            const int64_t i = i0 + (int64_t)u * blockDim.x;
            ok[u] = i < r1 && P::template stage<KW>(A, A.row_base + (i < r1 ? i : r0), st[u]);
        }
#pragma unroll
        for (int u = 0; u < kStageUnroll; ++u) {
            if (!ok[u]) continue;
            uint64_t key[KW];
#pragma unroll
            for (int w = 0; w < KW; ++w) key[w] = st[u][w];
            const uint64_t h = key_place_hash<KW>(key);
            const int hi = nhot ? hot_find(H, key[0], h) : -1;
            if (hi >= 0) {
                uint64_t k2[KW];
                uint64_t c[S];
                P::template unstage<KW, S>(A, st[u], k2, c);
                // The hot key's LDS record (the bucketed path has no i128
                // MIN/MAX pairs): atomics, a MIN/MAX word skipped when a plain
                // read already dominates. (apply_words inlined here crashes
                // ROCm 7.2's SimplifyCFG.)
                uint64_t* hr = HT.recs + (size_t)hi * R + KW;
#pragma unroll
                for (int w = 0; w < S; ++w) {
                    if (w < P::nwords(A)) {
                        const int op = P::op(A, w);
                        const bool mm = op == OP_MIN_I64 || op == OP_MAX_I64 || op == OP_MIN_U64 || op == OP_MAX_U64;
                        const uint64_t now = mm ? __hip_atomic_load(&hr[w], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) : 0ull;
                        if (!mm || combine((uint8_t)op, now, c[w]) != now) atomic_apply_lds((uint8_t)op, &hr[w], c[w]);
                    }
                }
            } else {
                const unsigned int b = (unsigned int)(h >> shift);
                const unsigned int pos = atomicAdd(&cur[b], 1u);
#pragma unroll
                for (int w = 0; w < W; ++w) staged[(uint64_t)pos * W + w] = st[u][w];
            }
        }
    }
    if (nhot) {
        __syncthreads();
        unsigned long long* dst = hot_parts + (size_t)blockIdx.x * kHotMax * R;
        for (unsigned int k = threadIdx.x; k < nhot * (unsigned int)R; k += blockDim.x) dst[k] = HT.recs[k];
    }
}

// Aggregates items of bucket-sorted staged rows in the workgroup's LDS table,
// one item at a time (from a device queue); writes the item's groups as
// records (key words, state words) from recs[row0 * R] on and their count to
// counts[item]. A row whose group finds no room in the LDS table becomes a
// record of its own, and the count carries kSharedItem.
template <int KW, int S, int W, class P>
__device__ __forceinline__ void agg_sorted_body(const AggArgs& A, const unsigned long long* staged, const SortedItem* items, int nitems,
    unsigned int* head, unsigned long long* recs, unsigned int* counts)
{
    extern __shared__ uint64_t lds_raw[];
    __shared__ unsigned int s_unused;
    __shared__ unsigned int s_nrec;
    __shared__ unsigned int s_shared;
    __shared__ int s_item;
    LdsTable T = lds_table_at<KW, P>(A, lds_raw, &s_unused);
    const int R = KW + P::nwords(A);
    const int lane = lane_id();
    for (;;) {
        if (threadIdx.x == 0) s_item = (int)atomicAdd(head, 1u);
        __syncthreads();
        const int it = __builtin_amdgcn_readfirstlane(s_item);
        __syncthreads();
        if (it >= nitems) break;
        const SortedItem w = items[it];
        if (threadIdx.x == 0) {
            s_nrec = 0u;
            s_shared = 0u;
        }
        lds_table_reset<KW, P>(A, T, true);
        unsigned long long* out = recs + (uint64_t)w.row0 * (uint64_t)R;
        for (int64_t base0 = w.row0 + (int64_t)threadIdx.x - lane; base0 < w.row1; base0 += (int64_t)blockDim.x * kStageUnroll) {
          // kStageUnroll rows per lane: every load issued before the first
          // row is applied.
          uint64_t stu[kStageUnroll][W];
#pragma unroll
          for (int u = 0; u < kStageUnroll; ++u) {
              const int64_t i = base0 + (int64_t)u * blockDim.x + lane;
#pragma unroll
              for (int k = 0; k < W; ++k) stu[u][k] = staged[(uint64_t)(i < w.row1 ? i : w.row0) * W + k];
          }
#pragma unroll
          for (int u = 0; u < kStageUnroll; ++u) {
            const int64_t base = base0 + (int64_t)u * blockDim.x;
            if (base >= w.row1) break;  // wave-uniform
            const int64_t i = base + lane;
            const bool valid = i < w.row1;
            uint64_t key[KW];
            uint64_t c[S];
            P::template unstage<KW, S>(A, stu[u], key, c);
            int where = valid ? AT_BUSY : AT_NONE;
            int lg = 0;
            while (__any(where == AT_BUSY)) {
                if (where == AT_BUSY) {
                    const int r = lds_find_or_insert<KW, S, P>(A, T, key);
                    where = r >= 0 ? AT_LDS : r == (int)kBusy ? AT_BUSY : AT_NONE;
                    if (r >= 0) lg = r;
                }
            }
            if (valid && where == AT_LDS) {
                apply_words<KW, S, P>(A, T, AT_LDS, lg, c);
            } else if (valid) {
                // No room in LDS: the row is a record of its own.
                const unsigned int pos = atomicAdd(&s_nrec, 1u);
                unsigned long long* r = out + (uint64_t)pos * R;
#pragma unroll
                for (int k = 0; k < KW; ++k) r[k] = key[k];
#pragma unroll
                for (int k = 0; k < S; ++k) {
                    if (k < P::nwords(A)) r[KW + k] = c[k];
                }
                s_shared = 1u;
            }
          }
        }
        __syncthreads();
        const unsigned int n = *T.count < (unsigned int)A.lds_groups ? *T.count : (unsigned int)A.lds_groups;
        const unsigned int first = s_nrec;
        // The LDS records are dense in the output layout: one flat copy.
        unsigned long long* dst = out + (uint64_t)first * R;
        for (unsigned int k = threadIdx.x; k < n * (unsigned int)R; k += blockDim.x) dst[k] = T.recs[k];
        if (threadIdx.x == 0) counts[it] = (first + n) | (s_shared ? kSharedItem : 0u);
        __syncthreads();
    }
}

// Merges one record into the HBM table. Owned: its key occurs once in the
// item and in no other item, so a key this thread inserts takes the
// record's words with plain (coherent) stores; otherwise (a key already in
// the table, or a bucket split over items) the words are combined with
// atomics. A record that finds no room is kept for the replay after growth.
// One lookup serves every case.
template <int KW, int S, class P>
__device__ __forceinline__ void merge_record(const AggArgs& A, LdsTable& T, bool want, bool owned, const uint64_t (&key)[KW],
    const uint64_t (&c)[S])
{
    int64_t id = 0;
    T.fresh = 0;
    const int where = locate<KW, S, P>(A, T, false, want, key, &id);
    if (want && where == AT_HBM) {
        // A key this thread just inserted into an owned bucket's slot: its
        // words still hold their identities and no one else can update
        // them, so the record's words are stored as they are.
        const bool store = owned && T.fresh;
        uint64_t* r = A.recs + (uint64_t)id * (uint64_t)A.rstride + KW;
#pragma unroll
        for (int w = 0; w < S; ++w) {
            if (w >= P::nwords(A)) continue;
            const uint8_t op = P::op(A, w);
            if (c[w] == identity_of(op) && op != OP_FADD) continue;
            if (store) __hip_atomic_store(&r[w], c[w], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            else atomic_apply_global(op, &r[w], c[w]);
        }
    } else if (want) {
        const unsigned int pos = atomicAdd(&A.counters[2], 1u);
        if (pos < A.ov_rec_cap) {
            uint64_t* r = A.ov_recs + (size_t)pos * (KW + P::nwords(A));
#pragma unroll
            for (int w = 0; w < KW; ++w) r[w] = key[w];
#pragma unroll
            for (int w = 0; w < S; ++w) {
                if (w < P::nwords(A)) r[KW + w] = c[w];
            }
        }
    }
}

// Merges every item's records into the HBM table (items from a device
// queue), with plain stores for owned items and atomics for the others.
template <int KW, int S, class P>
__device__ __forceinline__ void agg_merge_body(const AggArgs& A, const unsigned long long* recs, const SortedItem* items,
    const unsigned int* counts, int nitems, unsigned int* head)
{
    __shared__ unsigned int s_inserted;
    __shared__ int s_item;
    LdsTable T;
    T.recs = nullptr;
    T.ctrl = nullptr;
    T.count = nullptr;
    T.fresh = 0;
    T.inserted = &s_inserted;
    if (threadIdx.x == 0) s_inserted = 0u;
    __syncthreads();
    const int R = KW + P::nwords(A);
    for (;;) {
        if (threadIdx.x == 0) s_item = (int)atomicAdd(head, 1u);
        __syncthreads();
        const int it = __builtin_amdgcn_readfirstlane(s_item);
        __syncthreads();
        if (it >= nitems) break;
        const SortedItem w = items[it];
        const unsigned int cw = counts[it];
        const unsigned int n = cw & ~kSharedItem;
        const bool owned = w.owned && !(cw & kSharedItem);
        const unsigned long long* in = recs + (uint64_t)w.row0 * (uint64_t)R;
        for (unsigned int b = 0; b < n; b += blockDim.x) {
            const unsigned int g = b + threadIdx.x;
            const bool want = g < n;
            const unsigned long long* r = in + (uint64_t)(want ? g : 0u) * R;
            uint64_t key[KW];
            uint64_t c[S];
#pragma unroll
            for (int k = 0; k < KW; ++k) key[k] = r[k];
#pragma unroll
            for (int k = 0; k < S; ++k) c[k] = k < P::nwords(A) ? r[KW + k] : 0;
            merge_record<KW, S, P>(A, T, want, owned, key, c);
        }
    }
    __syncthreads();
    if (threadIdx.x == 0 && s_inserted) atomicAdd(&A.counters[0], s_inserted);
}

template <int KW, int S>
__global__ void __launch_bounds__(kBlock) k_agg_update(AggArgs A)
{
    agg_update_body<KW, S, GenericPolicy>(A);
}

template <int KW, int S>
__global__ void __launch_bounds__(kBlock) k_agg_batch(const AggArgs* insts, const BatchItem* items, int nitems, unsigned int* head)
{
    agg_batch_body<KW, S, GenericPolicy>(insts, items, nitems, head);
}

// Replays spilled records (key + words) into the grown HBM table.
template <int KW, int S>
__global__ void __launch_bounds__(kBlock) k_agg_apply_records(AggArgs A, const uint64_t* recs, unsigned int n)
{
    using P = GenericPolicy;
    __shared__ unsigned int s_inserted;
    LdsTable T;
    T.recs = nullptr;
    T.ctrl = nullptr;
    T.count = nullptr;
    T.fresh = 0;
    T.inserted = &s_inserted;
    if (threadIdx.x == 0) s_inserted = 0u;
    __syncthreads();
    const int R = KW + P::nwords(A);
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t b = (uint64_t)blockIdx.x * blockDim.x; b < n; b += stride) {
        const uint64_t i = b + threadIdx.x;
        const bool want = i < n;
        const uint64_t* r = recs + (size_t)(want ? i : 0) * R;
        uint64_t key[KW];
        uint64_t c[S];
#pragma unroll
        for (int w = 0; w < KW; ++w) key[w] = r[w];
#pragma unroll
        for (int w = 0; w < S; ++w) c[w] = w < P::nwords(A) ? r[KW + w] : 0;
        global_upsert_or_spill<KW, S, P>(A, T, want, key, c);
    }
    __syncthreads();
    if (threadIdx.x == 0 && s_inserted) atomicAdd(&A.counters[0], s_inserted);
}

// Whether slot s of a table holds a group (the extra record excluded).
template <int KW>
__device__ __forceinline__ bool slot_used(const uint64_t* recs, const unsigned long long* ctrl, int rstride, uint64_t s)
{
    if (KW == 1) return recs[s * (uint64_t)rstride] != kEmptyKey;
    return ctrl[s] >= 2ull;
}

// Moves every group of the old table (slots [0, old_slots)) into the grown
// table of A: keys are distinct, so a slot is claimed with one CAS and the
// record copied with plain stores. The extra record is copied by the host.
template <int KW>
__global__ void __launch_bounds__(kBlock) k_agg_rehash(AggArgs A, const uint64_t* old_recs, const unsigned long long* old_ctrl,
    uint64_t old_slots)
{
    const int R = KW + A.nwords;
    for (uint64_t s = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; s < old_slots; s += (uint64_t)gridDim.x * blockDim.x) {
        if (!slot_used<KW>(old_recs, old_ctrl, A.rstride, s)) continue;
        const uint64_t* src = old_recs + s * (uint64_t)A.rstride;
        uint64_t key[KW];
#pragma unroll
        for (int w = 0; w < KW; ++w) key[w] = src[w];
        const uint64_t h = key_place_hash<KW>(key);
        uint64_t slot = h & A.cmask;
        if (KW == 1) {
            while (atomicCAS((unsigned long long*)(A.recs + slot * (uint64_t)A.rstride), kEmptyKey, (unsigned long long)key[0]) != kEmptyKey) {
                slot = (slot + 1) & A.cmask;
            }
            atomicOr(&A.occ[slot >> 5], 1u << (slot & 31u));
        } else {
            const unsigned long long v = ((unsigned long long)(h >> 32) << 32) | 2ull;
            while (atomicCAS(&A.ctrl[slot], 0ull, v) != 0ull) slot = (slot + 1) & A.cmask;
        }
        uint64_t* dst = A.recs + slot * (uint64_t)A.rstride;
        for (int w = 1; w < R; ++w) dst[w] = src[w];
        if (KW > 1) dst[0] = src[0];
    }
}

// ---------------------------------------------------------------------------
// Finalize: records -> output columns.
// ---------------------------------------------------------------------------

enum RedKind : uint8_t { RK_COUNT_ROWS, RK_COUNT_VALID, RK_SUM64, RK_SUM128, RK_SUMF64, RK_MINMAX_I64, RK_MINMAX_F64, RK_MINMAX_I128 };

struct RedOut {
    int32_t kind;
    int32_t out_width;    // bytes per output value
    int32_t nullable_out;
    int32_t is_max;
    int32_t w0;           // first state word
    int32_t wvalid;       // valid-count word, or -1 (then the rows word)
};

struct FinArgs {
    const uint64_t* recs;
    const uint32_t* index;    // slot of output group g - pend_n (null: itself)
    // Output groups [0, pend_n) come from record runs of R words each: run
    // k is pend_items[k].row0 on, pend_prefix[k + 1] - pend_prefix[k]
    // records; the groups after them from recs / index.
    const uint64_t* pend_recs;
    const SortedItem* pend_items;
    const uint32_t* pend_prefix;
    int pend_runs;
    unsigned int pend_n;
    int stride;               // words between records
    int R;
    int KW;
    unsigned int groups;
    int nkeys;
    KeyFieldDesc keys[QDB_AGG_MAX_KEYS];
    int32_t key_out_width[QDB_AGG_MAX_KEYS];
    int nred;
    RedOut red[QDB_AGG_MAX_AGGS];
    int rows_word;
    char* out_data[QDB_AGG_MAX_KEYS + QDB_AGG_MAX_AGGS];
    uint8_t* out_mask[QDB_AGG_MAX_KEYS + QDB_AGG_MAX_AGGS];
};


#ifndef __HIPCC_RTC__
// Launchers, one translation unit per key width (aggregate_kw<KW>.hip).
void launch_update(int KW, int S, const AggArgs& A, int grid, size_t lds);
void launch_batch(int KW, int S, const AggArgs* insts, const BatchItem* items, int nitems, unsigned int* head, int grid, size_t lds);
void launch_records(int KW, int S, const AggArgs& A, const uint64_t* recs, unsigned int n);
void launch_rehash(int KW, const AggArgs& A, const uint64_t* old_recs, const unsigned long long* old_ctrl, uint64_t old_slots);
#endif

}  // namespace agg
}  // namespace qdbgpu
