// Device-side helpers shared by the filter, project, aggregate and join
// kernels: column reads, key hashing, the expression interpreter, i128 and
// ordered-float encodings, wave64 primitives.
#pragma once

#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/qdb_gpu.h"
#else
#include "qdb_gpu.h"
#endif

// Loops over descriptors (key fields, sources, program instructions) are
// kept rolled in the precompiled generic kernels and fully unrolled in the
// kernels specialized at run time, where the descriptors are constants.
#ifdef QDB_SPECIALIZED
#define QDB_DESC_LOOP _Pragma("unroll")
#else
#define QDB_DESC_LOOP _Pragma("unroll 1")
#endif

namespace qdbgpu {

typedef __int128 i128;
typedef unsigned __int128 u128;

constexpr int kWave = 64;
constexpr int kMaxDevCols = 16;

// Every descriptor a kernel reads from its arguments has 4-byte fields: a
// walk over an array of byte fields can make the compiler address a scalar
// load from an odd base, and scalar loads ignore the low address bits.

// One input column as a kernel sees it: the TColumn fields plus the value
// type, so that a launch passes its columns by value.
struct DevCol {
    const char* data;
    const uint8_t* mask;
    int32_t mask_off;
    int32_t bit_off;
    int32_t kind;
    int32_t width;
    int32_t nullable;
    int32_t pad;
};

// An expression instruction as the kernels read it (QdbExprInsn with 4-byte
// fields; programs are converted when they are uploaded).
struct DevInsn {
    int32_t op;
    int32_t dst;
    int32_t a;
    int32_t b;
    int32_t cls;
    int32_t column;
    int64_t lo;
    int64_t hi;
};

struct DevProgram {
    int32_t n;
    int32_t result;
    DevInsn insns[QDB_EXPR_MAX_INSNS];
};

inline DevProgram to_dev_program(const QdbExprProgram& p)
{
    DevProgram d;
    d.n = p.NumInsns;
    d.result = p.Result;
    for (int k = 0; k < QDB_EXPR_MAX_INSNS; ++k) {
        const QdbExprInsn& s = p.Insns[k];
        d.insns[k].op = s.Op;
        d.insns[k].dst = s.Dst;
        d.insns[k].a = s.A;
        d.insns[k].b = s.B;
        d.insns[k].cls = s.Class;
        d.insns[k].column = s.Column;
        d.insns[k].lo = s.ImmLo;
        d.insns[k].hi = s.ImmHi;
    }
    return d;
}

struct DevCols {
    DevCol c[kMaxDevCols];
};

// A value read from a column or computed by a program: exact integer
// (i128), f64, or bool in `i`; `valid` is SQL non-NULL.
struct DVal {
    i128 i;
    double f;
    uint8_t cls;
    uint8_t valid;
};

__device__ __forceinline__ int bit_at(const uint8_t* bitmap, int64_t index, int64_t offset)
{
    int64_t b = index + offset;
    return (bitmap[b >> 3] >> (b & 7)) & 1;
}

__device__ __forceinline__ bool col_valid(const DevCol& c, int64_t row)
{
    return c.mask == nullptr || bit_at(c.mask, row, c.mask_off);
}

__device__ __forceinline__ i128 col_int(const DevCol& c, int64_t row)
{
    const char* d = c.data;
    const bool uns = c.kind == QDB_KIND_UNSIGNED;
    if (c.kind == QDB_KIND_BOOL) return bit_at((const uint8_t*)d, row, c.bit_off);
    switch (c.width) {
    case 1: return uns ? (i128)((const uint8_t*)d)[row] : (i128)((const int8_t*)d)[row];
    case 2: return uns ? (i128)((const uint16_t*)d)[row] : (i128)((const int16_t*)d)[row];
    case 4: return uns ? (i128)((const uint32_t*)d)[row] : (i128)((const int32_t*)d)[row];
    case 8: return uns ? (i128)((const uint64_t*)d)[row] : (i128)((const int64_t*)d)[row];
    default: {
        const uint64_t* p = (const uint64_t*)(d + 16 * row);
        return (i128)(((u128)p[1] << 64) | (u128)p[0]);
    }
    }
}

__device__ __forceinline__ double col_f64(const DevCol& c, int64_t row)
{
    if (c.width == 4) return (double)((const float*)c.data)[row];
    return ((const double*)c.data)[row];
}

__device__ __forceinline__ DVal read_col(const DevCol& c, int64_t row)
{
    DVal v;
    v.valid = col_valid(c, row);
    v.i = 0;
    v.f = 0.0;
    if (c.kind == QDB_KIND_FLOAT) {
        v.cls = QDB_REG_F64;
        if (v.valid) v.f = col_f64(c, row);
    } else {
        v.cls = c.kind == QDB_KIND_BOOL ? QDB_REG_BOOL : QDB_REG_INT;
        if (v.valid) v.i = col_int(c, row);
    }
    return v;
}

// Canonical bits of a float key: +0/-0 merged, every NaN the quiet NaN
// (qdb/kernel/gen.cpp:192-231).
__device__ __forceinline__ uint64_t canonical_f64_bits(double v)
{
    uint64_t b = (uint64_t)__double_as_longlong(v);
    if ((b & 0x7fffffffffffffffull) == 0) return 0;
    if ((b & 0x7ff0000000000000ull) == 0x7ff0000000000000ull && (b & 0x000fffffffffffffull) != 0) {
        return 0x7ff8000000000000ull;
    }
    return b;
}

// rh_hash integer mix (qdb/kernel/gen.cpp:269-277).
__device__ __forceinline__ uint64_t mix_u64(uint64_t h)
{
    h ^= h >> 12;
    h ^= h << 25;
    h ^= h >> 27;
    return h * 2685821657736338717ull;
}

// rh_hash of one valid key field, the zero-extension of its own width.
__device__ __forceinline__ uint64_t leaf_hash(const DevCol& c, int64_t row)
{
    if (c.kind == QDB_KIND_FLOAT) {
        if (c.width == 4) {
            uint32_t b = (uint32_t)__float_as_uint(((const float*)c.data)[row]);
            if ((b & 0x7fffffffu) == 0) b = 0;
            else if ((b & 0x7f800000u) == 0x7f800000u && (b & 0x007fffffu) != 0) b = 0x7fc00000u;
            return mix_u64(b);
        }
        return mix_u64(canonical_f64_bits(((const double*)c.data)[row]));
    }
    if (c.kind == QDB_KIND_BOOL) {
        return (uint64_t)bit_at((const uint8_t*)c.data, row, c.bit_off) * 2685821657736338717ull;
    }
    if (c.width == 16) {
        const uint64_t* p = (const uint64_t*)(c.data + 16 * row);
        uint64_t lo = mix_u64(p[0]);
        uint64_t hi = mix_u64(p[1]);
        return lo ^ (hi + 0x9E3779B97F4A7C15ull + (lo << 6) + (lo >> 2));
    }
    uint64_t bits;
    switch (c.width) {
    case 1: bits = ((const uint8_t*)c.data)[row]; break;
    case 2: bits = ((const uint16_t*)c.data)[row]; break;
    case 4: bits = ((const uint32_t*)c.data)[row]; break;
    default: bits = ((const uint64_t*)c.data)[row]; break;
    }
    return mix_u64(bits);
}

// Boost-style ordered combine of field hashes (qdb/kernel/gen.cpp:344-373).
__device__ __forceinline__ uint64_t hash_combine(uint64_t h, uint64_t fh)
{
    return h ^ (fh + 0x9E3779B97F4A7C15ull + (h << 6) + (h >> 2));
}

// rh_hash of a row's key: the reference always hashes the key STRUCT
// (qdb/kernel/gen.cpp:344-373), whose fields are, per key column, a
// `valid_i` bool when the column is nullable and then `key_i`
// (qdb/kernel/aggregate_key.cpp:246-272). A bool field hashes as
// value * 2685821657736338717 (gen.cpp:319-325) and a NULL key field as 0
// (gen.cpp:357-364); every field hash enters the ordered combine from h = 0,
// so even one non-null BIGINT k hashes to mix(k) + 0x9E3779B97F4A7C15.
template <int N>
__device__ __forceinline__ uint64_t rh_key_hash(const DevCol (&c)[N], int nkeys, int64_t row)
{
    uint64_t h = 0;
#pragma unroll
    for (int k = 0; k < N; ++k) {
        if (k >= nkeys) break;
        bool valid = true;
        if (c[k].nullable) {
            valid = col_valid(c[k], row);
            h = hash_combine(h, valid ? 2685821657736338717ull : 0ull);
        }
        h = hash_combine(h, valid ? leaf_hash(c[k], row) : 0ull);
    }
    return h;
}

// Placement hash of a packed key (not result-visible): a strong 64-bit
// finalizer over the key words.
template <int KW>
__device__ __forceinline__ uint64_t key_place_hash(const uint64_t (&k)[KW])
{
    uint64_t h = 0x243F6A8885A308D3ull;
#pragma unroll
    for (int w = 0; w < KW; ++w) {
        uint64_t z = h ^ k[w];
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        h = z ^ (z >> 31);
    }
    return h;
}

// ---------------------------------------------------------------------------
// Expression interpreter. A program is read through a uniform pointer, so
// every lane runs the same instruction; the registers live in VGPRs.
// ---------------------------------------------------------------------------

__device__ __forceinline__ i128 pow10_i128(int64_t k)
{
    i128 p = 1;
    for (int64_t i = 0; i < k; ++i) p *= 10;
    return p;
}

// |n| / d for u128 operands by shift and subtract (the rare exact path of
// the decimal conversion below; no 128-bit division routine needed).
__device__ __forceinline__ u128 udiv128(u128 n, u128 d, u128* rem)
{
    u128 q = 0, r = 0;
    for (int i = 127; i >= 0; --i) {
        r = (r << 1) | ((n >> i) & 1);
        if (r >= d) {
            r -= d;
            q |= (u128)1 << i;
        }
    }
    *rem = r;
    return q;
}

// v / 10^scale as the nearest double, ties to even: the contract of
// qdb_decimal_to_f64 (qdb/modules/qumirdb.oz:983-1046). Below 2^53 with
// scale <= 22 both operands are exact doubles and one IEEE division rounds
// correctly; otherwise the integer quotient gets fraction bits by long
// division until it holds at least 55 bits, and is rounded to 53 bits with a
// guard bit and a sticky bit (the dropped bits and the remainder).
__device__ __noinline__ double decimal_to_f64_dev(i128 v, int64_t scale)
{
    const bool neg = v < 0;
    const u128 mag = neg ? (u128)0 - (u128)v : (u128)v;
    const u128 p = (u128)pow10_i128(scale);
    double r;
    if (mag < ((u128)1 << 53) && scale <= 22) {
        r = (double)(uint64_t)mag / (double)(uint64_t)p;
    } else if (mag == 0) {
        r = 0.0;
    } else {
        u128 rem;
        u128 q = udiv128(mag, p, &rem);
        int e = 0;
        while (q < ((u128)1 << 54)) {
            // r < p, so compare r with p - r instead of forming 2r.
            q <<= 1;
            if (rem >= p - rem) {
                q |= 1;
                rem -= p - rem;
            } else {
                rem += rem;
            }
            --e;
        }
        int guard = 0, sticky = rem != 0;
        while (q >= ((u128)1 << 53)) {
            sticky |= guard;
            guard = (int)(q & 1);
            q >>= 1;
            ++e;
        }
        if (guard && (sticky || (q & 1))) q += 1;
        r = ldexp((double)(uint64_t)q, e);
    }
    return neg ? -r : r;
}

// One expression instruction on its operand values: the semantics shared
// by the interpreter below and by kernels specialized at run time (which
// call it with a constant instruction, so that the switch folds away).
template <bool kHasCols>
__device__ __forceinline__ DVal exec_insn(const DevInsn& in, const DVal& a, const DVal& b, const DevCols& cols, int64_t row)
{
    DVal o;
    o.i = 0;
    o.f = 0.0;
    o.cls = in.cls;
    o.valid = 0;
    switch (in.op) {
    case QDB_EXPR_COLUMN:
        // The column index is uniform: the descriptor is read from the
        // kernel arguments with scalar loads.
        if (kHasCols) o = read_col(cols.c[in.column], row);
        break;
    case QDB_EXPR_CONST_INT:
        o.valid = 1;
        o.i = (i128)(((u128)(uint64_t)in.hi << 64) | (u128)(uint64_t)in.lo);
        break;
    case QDB_EXPR_CONST_F64:
        o.valid = 1;
        o.f = __longlong_as_double(in.lo);
        break;
    case QDB_EXPR_ADD:
    case QDB_EXPR_SUB:
    case QDB_EXPR_MUL:
        o.valid = a.valid && b.valid;
        if (in.cls == QDB_REG_F64) {
            double x = a.cls == QDB_REG_F64 ? a.f : (double)a.i;
            double y = b.cls == QDB_REG_F64 ? b.f : (double)b.i;
            double z = in.op == QDB_EXPR_ADD ? x + y : in.op == QDB_EXPR_SUB ? x - y : x * y;
            o.f = o.valid ? z : 0.0;
        } else if (o.valid) {
            u128 x = (u128)a.i, y = (u128)b.i;
            if (in.op == QDB_EXPR_MUL) {
                // Operands that fit in 32 bits (decimal prices, rates) have
                // an exact 64-bit product; the full 128-bit multiply only for
                // the others.
                const bool small = a.i >= -(i128)2147483647 && a.i <= (i128)2147483647 && b.i >= -(i128)2147483647 &&
                    b.i <= (i128)2147483647;
                if (small) o.i = (i128)((int64_t)a.i * (int64_t)b.i);
                else o.i = (i128)(x * y);
            } else {
                o.i = (i128)(in.op == QDB_EXPR_ADD ? x + y : x - y);
            }
        }
        break;
    case QDB_EXPR_NEG:
        o.valid = a.valid;
        if (in.cls == QDB_REG_F64) o.f = a.valid ? -a.f : 0.0;
        else if (a.valid) o.i = (i128)((u128)0 - (u128)a.i);
        break;
    case QDB_EXPR_SCALE_UP:
        o.valid = a.valid;
        if (a.valid) o.i = (i128)((u128)a.i * (u128)pow10_i128(in.lo));
        break;
    case QDB_EXPR_CMP: {
        o.cls = QDB_REG_BOOL;
        o.valid = a.valid && b.valid;
        bool c = false;
        if (a.cls == QDB_REG_F64 || b.cls == QDB_REG_F64) {
            double x = a.cls == QDB_REG_F64 ? a.f : (double)a.i;
            double y = b.cls == QDB_REG_F64 ? b.f : (double)b.i;
            switch (in.lo) {
            case QDB_CMP_EQ: c = x == y; break;
            case QDB_CMP_NE: c = x != y; break;
            case QDB_CMP_LT: c = x < y; break;
            case QDB_CMP_LE: c = x <= y; break;
            case QDB_CMP_GT: c = x > y; break;
            default: c = x >= y; break;
            }
        } else {
            i128 x = a.i, y = b.i;
            switch (in.lo) {
            case QDB_CMP_EQ: c = x == y; break;
            case QDB_CMP_NE: c = x != y; break;
            case QDB_CMP_LT: c = x < y; break;
            case QDB_CMP_LE: c = x <= y; break;
            case QDB_CMP_GT: c = x > y; break;
            default: c = x >= y; break;
            }
        }
        o.i = (o.valid && c) ? 1 : 0;
        break;
    }
    case QDB_EXPR_AND: {
        bool af = a.valid && a.i == 0, bf = b.valid && b.i == 0;
        o.cls = QDB_REG_BOOL;
        if (af || bf) { o.valid = 1; o.i = 0; }
        else if (!a.valid || !b.valid) { o.valid = 0; }
        else { o.valid = 1; o.i = 1; }
        break;
    }
    case QDB_EXPR_OR: {
        bool at = a.valid && a.i != 0, bt = b.valid && b.i != 0;
        o.cls = QDB_REG_BOOL;
        if (at || bt) { o.valid = 1; o.i = 1; }
        else if (!a.valid || !b.valid) { o.valid = 0; }
        else { o.valid = 1; o.i = 0; }
        break;
    }
    case QDB_EXPR_NOT:
        o.cls = QDB_REG_BOOL;
        o.valid = a.valid;
        o.i = a.valid ? (a.i == 0) : 0;
        break;
    case QDB_EXPR_IS_NULL:
        o.cls = QDB_REG_BOOL;
        o.valid = 1;
        o.i = a.valid ? 0 : 1;
        break;
    case QDB_EXPR_TO_F64:
        o.cls = QDB_REG_F64;
        o.valid = a.valid;
        if (a.cls == QDB_REG_F64) o.f = a.f;
        else if (a.valid) o.f = decimal_to_f64_dev(a.i, in.lo);
        break;
    case QDB_EXPR_DIV: {
        o.valid = a.valid && b.valid;
        const double x = a.cls == QDB_REG_F64 ? a.f : (double)a.i;
        const double y = b.cls == QDB_REG_F64 ? b.f : (double)b.i;
        o.f = o.valid ? x / y : 0.0;
        break;
    }
    default:
        break;
    }
    return o;
}

template <bool kHasCols>
__device__ __forceinline__ DVal eval_program(const DevProgram* __restrict__ prog, const DevCols& cols, int64_t row)
{
    DVal r[QDB_EXPR_MAX_REGS];
#pragma unroll
    for (int k = 0; k < QDB_EXPR_MAX_REGS; ++k) {
        r[k].i = 0;
        r[k].f = 0.0;
        r[k].cls = 0;
        r[k].valid = 0;
    }
    const int n = prog->n;
    QDB_DESC_LOOP
    for (int k = 0; k < QDB_EXPR_MAX_INSNS; ++k) {
        if (k >= n) break;
        const DevInsn in = prog->insns[k];
        DVal a, b;
        // Registers are indexed by uniform values: select through a switch so
        // that the register file stays in VGPRs.
        a = r[0]; b = r[0];
#pragma unroll
        for (int q = 1; q < QDB_EXPR_MAX_REGS; ++q) {
            if (in.a == q) a = r[q];
            if (in.b == q) b = r[q];
        }
        const DVal o = exec_insn<kHasCols>(in, a, b, cols, row);
#pragma unroll
        for (int q = 0; q < QDB_EXPR_MAX_REGS; ++q) {
            if (in.dst == q) r[q] = o;
        }
    }
    DVal res = r[0];
#pragma unroll
    for (int q = 1; q < QDB_EXPR_MAX_REGS; ++q) {
        if (prog->result == q) res = r[q];
    }
    return res;
}

// ---------------------------------------------------------------------------
// Order-preserving encodings of f64 for integer atomics.
// ---------------------------------------------------------------------------

// Monotone map of doubles to u64 (-inf < ... < -0 < +0 < ... < +inf). For MIN a
// NaN maps just above +inf (ignored unless every value is NaN); for MAX a
// NaN maps to 0, below every number.
__device__ __forceinline__ uint64_t f64_order(double v)
{
    uint64_t b = (uint64_t)__double_as_longlong(v);
    return (b & 0x8000000000000000ull) ? ~b : (b | 0x8000000000000000ull);
}

__device__ __forceinline__ uint64_t f64_min_key(double v)
{
    return v != v ? 0xfff8000000000001ull : f64_order(v);
}

__device__ __forceinline__ uint64_t f64_max_key(double v)
{
    return v != v ? 0ull : f64_order(v);
}

__device__ __forceinline__ double f64_from_order(uint64_t k)
{
    uint64_t b = (k & 0x8000000000000000ull) ? (k & 0x7fffffffffffffffull) : ~k;
    return __longlong_as_double((long long)b);
}

// ---------------------------------------------------------------------------
// Exact f64 summation window: a double whose bits all lie in
// [2^kLimbLo, 2^(kLimbLo + 32 * kLimbs)) is added exactly as 32-bit chunks
// into kLimbs signed 64-bit limbs (no carries; 2^31 additions per limb are
// safe). Anything else (huge, tiny below the window, inf, NaN) is added to
// one f64 side word.
// ---------------------------------------------------------------------------

constexpr int kLimbs = 4;
constexpr int kLimbLo = -64;

// Contribution of v to limb `k` (0..kLimbs-1), or 0; sets *outside when v
// does not fit the window.
__device__ __forceinline__ int64_t f64_limb(double v, int k, bool* outside)
{
    uint64_t b = (uint64_t)__double_as_longlong(v);
    int e = (int)((b >> 52) & 0x7ff);
    uint64_t m = b & 0x000fffffffffffffull;
    if (e == 0x7ff) {
        *outside = true;
        return 0;
    }
    if (e == 0) {
        if (m == 0) return 0;
        e = 1;
    } else {
        m |= 1ull << 52;
    }
    // v = m * 2^(e - 1075). Lowest set bit position relative to 2^kLimbLo:
    int lsb = e - 1075 - kLimbLo;
    int tz = __builtin_ctzll(m);
    if (lsb + tz < 0 || lsb + 53 > 32 * kLimbs - 1) {
        *outside = true;
        return 0;
    }
    // Bits of m land at [lsb, lsb + 53); chunk k covers [32k, 32k + 32).
    int lo = 32 * k - lsb;       // bit of m at the chunk's base
    uint64_t chunk;
    if (lo >= 53) chunk = 0;
    else if (lo >= 0) chunk = (m >> lo) & 0xffffffffull;
    else if (lo > -32) chunk = (m << (-lo)) & 0xffffffffull;
    else chunk = 0;
    int64_t c = (int64_t)chunk;
    return (b >> 63) ? -c : c;
}

// ---------------------------------------------------------------------------
// Wave64 primitives.
// ---------------------------------------------------------------------------

__device__ __forceinline__ int lane_id()
{
    return __lane_id();
}

__device__ __forceinline__ uint64_t shfl_u64(uint64_t v, int src)
{
    return (uint64_t)__shfl((long long)v, src, kWave);
}

__device__ __forceinline__ uint64_t shfl_xor_u64(uint64_t v, int m)
{
    return (uint64_t)__shfl_xor((long long)v, m, kWave);
}

}  // namespace qdbgpu

namespace qdbgpu {

// Exact value of the limb accumulator, sum_k limb[k] * 2^(32k + kLimbLo),
// rounded once to the nearest double (ties to even).
__device__ __forceinline__ double limbs_to_f64(const int64_t (&limb)[kLimbs])
{
    // 256-bit two's complement accumulator, little-endian 64-bit words; bit 0
    // has weight 2^kLimbLo.
    uint64_t acc[4] = {0, 0, 0, 0};
#pragma unroll
    for (int k = 0; k < kLimbs; ++k) {
        // limb << (32k), sign-extended to 256 bits.
        int64_t v = limb[k];
        uint64_t ext = v < 0 ? ~0ull : 0ull;
        uint64_t part[5] = {0, 0, 0, 0, 0};
        int sh = 32 * k;
        int wq = sh / 64, bq = sh % 64;
        uint64_t src[2] = {(uint64_t)v, ext};
        // place (v as 128-bit sign-extended) at bit sh, then extend.
        for (int w = 0; w < 5; ++w) part[w] = ext;
        for (int w = 0; w < wq; ++w) part[w] = 0;
        if (bq == 0) {
            part[wq] = src[0];
            if (wq + 1 < 5) part[wq + 1] = src[1];
        } else {
            part[wq] = src[0] << bq;
            if (wq + 1 < 5) part[wq + 1] = (src[0] >> (64 - bq)) | (src[1] << bq);
        }
        uint64_t carry = 0;
        for (int w = 0; w < 4; ++w) {
            uint64_t s = acc[w] + part[w];
            uint64_t c1 = s < acc[w];
            uint64_t s2 = s + carry;
            uint64_t c2 = s2 < s;
            acc[w] = s2;
            carry = c1 | c2;
        }
    }
    bool neg = (acc[3] >> 63) != 0;
    if (neg) {
        // Two's complement negation: ~x + 1, the carry ripples while zero.
        uint64_t carry = 1;
        for (int w = 0; w < 4; ++w) {
            acc[w] = ~acc[w] + carry;
            carry = (carry && acc[w] == 0) ? 1 : 0;
        }
    }
    int top = -1;
    for (int w = 3; w >= 0 && top < 0; --w) {
        if (acc[w]) top = w * 64 + 63 - __builtin_clzll(acc[w]);
    }
    if (top < 0) return 0.0;
    // The 53 bits from `top` down (fewer when top < 52), then the guard bit
    // below them and the sticky OR of every bit below the guard, by whole
    // words (a word index outside the accumulator reads as zero).
    auto word = [&](int i) -> uint64_t { return i == 0 ? acc[0] : i == 1 ? acc[1] : i == 2 ? acc[2] : i == 3 ? acc[3] : 0ull; };
    const int lsb = top > 52 ? top - 52 : 0;
    const int nbits = top - lsb + 1;
    const int wq = lsb >> 6, bq = lsb & 63;
    uint64_t q = (word(wq) >> bq) | (bq ? word(wq + 1) << (64 - bq) : 0ull);
    q &= (1ull << nbits) - 1ull;
    uint64_t guard = 0, sticky = 0;
    if (lsb >= 1) {
        const int gp = lsb - 1;
        guard = (word(gp >> 6) >> (gp & 63)) & 1ull;
        const int full = gp >> 6, rem = gp & 63;  // whole words, then bits, below the guard
#pragma unroll
        for (int w = 0; w < 4; ++w) {
            if (w < full) sticky |= acc[w];
        }
        if (rem) sticky |= word(full) & ((1ull << rem) - 1ull);
    }
    if (guard && (sticky || (q & 1))) q += 1;
    double r = ldexp((double)q, lsb + kLimbLo);
    return neg ? -r : r;
}

// Output columns of a project launch.
struct ProjectOut {
    char* data[QDB_PROJECT_MAX_OUTPUTS];
    uint8_t* mask[QDB_PROJECT_MAX_OUTPUTS];
    int32_t kind[QDB_PROJECT_MAX_OUTPUTS];
    int32_t width[QDB_PROJECT_MAX_OUTPUTS];
    // DECIMAL(p <= 18) outputs: p, the digits a stored value may have
    // (0: no check). A value with more is a row error: 0 is stored and bit k
    // of *err is set.
    int32_t prec[QDB_PROJECT_MAX_OUTPUTS];
    unsigned int* err;
};

// The row-error check of a narrowing store (qdb_decimal64_narrow_checked,
// qdb/modules/qumirdb.oz:1236-1245): true when |x| >= 10^prec.
__device__ __forceinline__ bool narrow_overflows(i128 x, int prec)
{
    const i128 lim = pow10_i128(prec);
    return x >= lim || x <= -lim;
}

// Writes 64 bits of a bitmap for the 64 rows starting at `base` (a multiple
// of 64); the last word writes only the bytes inside the buffer.
__device__ __forceinline__ void store_bits(uint8_t* bitmap, int64_t base, int64_t n, uint64_t bits, uint64_t covered)
{
    const int64_t first_byte = base >> 3;
    const int64_t rows = n - base < 64 ? n - base : 64;
    const int nbytes = (int)((rows + 7) >> 3);
    for (int b = 0; b < nbytes; ++b) {
        const uint8_t cov = (uint8_t)(covered >> (8 * b));
        if (cov == 0) continue;
        const uint8_t val = (uint8_t)(bits >> (8 * b));
        if (cov == 0xff) bitmap[first_byte + b] = val;
        else bitmap[first_byte + b] = (uint8_t)((bitmap[first_byte + b] & ~cov) | (val & cov));
    }
}

}  // namespace qdbgpu
