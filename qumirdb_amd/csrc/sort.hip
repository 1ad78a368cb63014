// ORDER BY on gfx950: the selected rows of a device batch ordered by key
// columns, as a list of row indexes.
//
// The reference sorts row ids with a least-significant-digit radix sort over
// 8-bit digits of a 64-bit order key per value (qumir_radix_key,
// qdb/kernel/sort/radix.oz:4-60): integers with the sign bit flipped,
// floats in PostgreSQL order (every NaN one key above +inf, -0 the key of
// +0), a descending key by complementing each digit, and the nullable
// variant putting NULL rows in a bucket before or after every digit
// (radix_nullable.oz:5-80; default NULLS LAST ascending, FIRST descending,
// qdb/plan/ops/sort_key.h:21-26). Every pass is stable, so equal keys keep
// the input order, and the keys of an ORDER BY are applied last to first.
//
// Here each key column is one stable device radix sort of (order key, row)
// pairs over the key's bits (a 16-byte decimal: its low word, then its
// high word), and a nullable key one more 1-bit pass that moves the NULL
// rows to the front or the back, stably. The same order as the reference's:
// complementing every digit complements the key, and a NULL bucket before or
// after all digits at every pass is a most-significant null flag.
//
// HBM traffic per key pass: the key column gathered through the current
// order (8 B or less per row), 8 B of order key written, then the radix
// sort's passes over 12 B (key + row index) per row.
#include <hip/hip_runtime.h>

#include <string.h>

#include "device.h"
#include "runtime.h"

namespace qdbgpu {

namespace {

constexpr int kBlock = 256;

__global__ void __launch_bounds__(kBlock) k_sel_flags(const uint8_t* sel, int64_t n, uint32_t* flags)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += stride) flags[i] = !sel || sel[i] ? 1u : 0u;
}

// The selected rows' indexes in input order; count[0] = how many.
__global__ void __launch_bounds__(kBlock) k_sel_compact(const uint32_t* flags, const uint32_t* pos, int64_t n, uint32_t* idx,
    unsigned long long* count)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += stride) {
        if (flags[i]) idx[pos[i]] = (uint32_t)i;
        if (i == n - 1) count[0] = (unsigned long long)pos[i] + flags[i];
    }
}

// qumir_radix_key of a double: PostgreSQL order, every NaN above +inf, -0
// the key of +0.
__device__ __forceinline__ uint64_t order_key_f64(uint64_t bits)
{
    const uint64_t mag = bits & 0x7fffffffffffffffull;
    if (mag > 0x7ff0000000000000ull) return ~0ull;
    if (mag == 0) bits = 0;
    return bits ^ ((bits >> 63) ? ~0ull : 0x8000000000000000ull);
}

__device__ __forceinline__ uint64_t order_key_f32(uint32_t bits)
{
    const uint32_t mag = bits & 0x7fffffffu;
    if (mag > 0x7f800000u) return 0xffffffffull;
    if (mag == 0) bits = 0;
    return (uint64_t)(bits ^ ((bits >> 31) ? 0xffffffffu : 0x80000000u));
}

// The order key of every row idx[i] of a column: part 0 the value (or a
// 16-byte value's low word, unsigned), part 1 a 16-byte value's high word
// (signed). A descending key is complemented within `bits`. NULL rows all
// get key 0, so that the value passes keep their order (the one the less
// significant keys gave them) and the null pass only moves them as a block.
__global__ void __launch_bounds__(kBlock) k_order_keys(DevCol c, int part, int desc, int bits, const uint32_t* idx, int64_t n,
    uint64_t* keys)
{
    const uint64_t mask = bits >= 64 ? ~0ull : ((1ull << bits) - 1ull);
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += stride) {
        const int64_t r = idx[i];
        if (c.mask) {
            const int64_t b = r + c.mask_off;
            if (((c.mask[b >> 3] >> (b & 7)) & 1) == 0) {
                keys[i] = 0;
                continue;
            }
        }
        uint64_t k = 0;
        switch (c.kind) {
        case QDB_KIND_BOOL: k = (c.data[(r + c.bit_off) >> 3] >> ((r + c.bit_off) & 7)) & 1; break;
        case QDB_KIND_FLOAT:
            k = c.width == 4 ? order_key_f32(((const uint32_t*)c.data)[r]) : order_key_f64(((const uint64_t*)c.data)[r]);
            break;
        case QDB_KIND_UNSIGNED:
            switch (c.width) {
            case 1: k = ((const uint8_t*)c.data)[r]; break;
            case 2: k = ((const uint16_t*)c.data)[r]; break;
            case 4: k = ((const uint32_t*)c.data)[r]; break;
            default: k = ((const uint64_t*)c.data)[r]; break;
            }
            break;
        default:  // signed integers, dates, decimals
            switch (c.width) {
            case 1: k = (uint64_t)(uint8_t)(((const uint8_t*)c.data)[r] ^ 0x80u); break;
            case 2: k = (uint64_t)(uint16_t)(((const uint16_t*)c.data)[r] ^ 0x8000u); break;
            case 4: k = (uint64_t)(((const uint32_t*)c.data)[r] ^ 0x80000000u); break;
            case 8: k = ((const uint64_t*)c.data)[r] ^ 0x8000000000000000ull; break;
            default: {
                const uint64_t* w = (const uint64_t*)(c.data + 16 * r);
                k = part == 0 ? w[0] : w[1] ^ 0x8000000000000000ull;
            }
            }
            break;
        }
        keys[i] = (desc ? ~k : k) & mask;
    }
}

// The null pass's key of every row idx[i]: the NULL rows 0 and the others 1
// for NULLS FIRST, the reverse for NULLS LAST.
__global__ void __launch_bounds__(kBlock) k_null_keys(DevCol c, int nulls_first, const uint32_t* idx, int64_t n, uint32_t* keys)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += stride) {
        const int64_t b = (int64_t)idx[i] + c.mask_off;
        const bool valid = ((c.mask[b >> 3] >> (b & 7)) & 1) != 0;
        keys[i] = valid == (nulls_first != 0) ? 1u : 0u;
    }
}

// Bits of a column's order key (per 64-bit part).
int order_bits(const QdbValueType& t)
{
    if (t.Kind == QDB_KIND_BOOL) return 1;
    if (t.Kind == QDB_KIND_FLOAT) return t.Width == 4 ? 32 : 64;
    return t.Width >= 8 ? 64 : 8 * t.Width;
}

}  // namespace

}  // namespace qdbgpu

using namespace qdbgpu;

extern "C" int qdb_gpu_sort_rowset(const TRowSet* rs, const QdbColumnSchema* types, const QdbSortKey* keys, int32_t nkeys,
    uint32_t* row_index, int64_t* count)
{
    clear_error();
    if (!rs || !types || !count || (nkeys > 0 && !keys)) return fail("null argument"), -1;
    *count = 0;
    if (nkeys < 0 || nkeys > QDB_SORT_MAX_KEYS) return fail("sort: %d keys (0..%d)", nkeys, QDB_SORT_MAX_KEYS), -1;
    for (int k = 0; k < nkeys; ++k) {
        const int c = keys[k].Column;
        if (c < 0 || c >= rs->ColumnCount || c >= types->NumColumns) return fail("sort key %d: column %d out of range", k, c), -1;
        const QdbValueType& t = types->Types[c];
        if (t.Kind == QDB_KIND_STRING) return fail("sort key %d: string keys are not supported", k), -1;
        if (t.Kind != QDB_KIND_BOOL && t.Width != 1 && t.Width != 2 && t.Width != 4 && t.Width != 8 && t.Width != 16) {
            return fail("sort key %d: unsupported width %d", k, t.Width), -1;
        }
        if (keys[k].Nulls != QDB_SORT_NULLS_DEFAULT && keys[k].Nulls != QDB_SORT_NULLS_FIRST && keys[k].Nulls != QDB_SORT_NULLS_LAST) {
            return fail("sort key %d: bad NULL placement %d", k, keys[k].Nulls), -1;
        }
    }
    const int64_t n = rs->RowCount;
    if (n <= 0) return 0;
    if (!row_index) return fail("null argument"), -1;
    if (n > (int64_t)0x7fffffff) return fail("sort: batches are limited to 2^31 - 1 rows"), -1;
    if (!ensure_init()) return -1;
    const int grid = grid_for(n, kBlock, 8);
    uint32_t* flags = (uint32_t*)dalloc((size_t)n * 4);
    uint32_t* pos = (uint32_t*)dalloc((size_t)n * 4);
    uint32_t* idx_b = (uint32_t*)dalloc((size_t)n * 4);
    uint64_t* ka = (uint64_t*)dalloc((size_t)n * 8);
    uint64_t* kb = (uint64_t*)dalloc((size_t)n * 8);
    unsigned long long* cnt = (unsigned long long*)dalloc(16);
    auto release = [&]() {
        dfree(flags);
        dfree(pos);
        dfree(idx_b);
        dfree(ka);
        dfree(kb);
        dfree(cnt);
    };
    if (!flags || !pos || !idx_b || !ka || !kb || !cnt) {
        release();
        return -1;
    }
    // The selected rows, in input order, into row_index.
    hipLaunchKernelGGL(k_sel_flags, dim3(grid), dim3(kBlock), 0, stream(), rs->Selection, n, flags);
    bool ok = hip_ok(hipGetLastError(), "sort select") && exclusive_scan_u32(flags, pos, n);
    if (ok) {
        hipLaunchKernelGGL(k_sel_compact, dim3(grid), dim3(kBlock), 0, stream(), flags, pos, n, row_index, cnt);
        ok = hip_ok(hipGetLastError(), "sort compact");
    }
    unsigned long long h_cnt = 0;
    ok = ok && hip_ok(hipMemcpyAsync(&h_cnt, cnt, 8, hipMemcpyDeviceToHost, stream()), "sort count") && sync("sort count");
    const int64_t m = (int64_t)h_cnt;
    // The keys, last to first; every pass stable. The order lives in `cur`
    // and each pass writes the other buffer.
    uint32_t* cur = row_index;
    uint32_t* other = idx_b;
    for (int k = nkeys - 1; ok && k >= 0 && m > 0; --k) {
        const int c = keys[k].Column;
        const QdbValueType& t = types->Types[c];
        const TColumn& src = rs->Columns[c];
        DevCol dc;
        memset(&dc, 0, sizeof(dc));
        dc.data = src.Data;
        dc.mask = src.Mask;
        dc.mask_off = src.MaskBitOffset;
        dc.bit_off = src.DataBitOffset;
        dc.kind = t.Kind;
        dc.width = t.Kind == QDB_KIND_BOOL ? 1 : t.Width;
        const int desc = keys[k].Desc ? 1 : 0;
        const int bits = order_bits(t);
        const int parts = t.Kind != QDB_KIND_BOOL && t.Kind != QDB_KIND_FLOAT && t.Width == 16 ? 2 : 1;
        const int kgrid = grid_for(m, kBlock, 8);
        for (int part = 0; ok && part < parts; ++part) {
            hipLaunchKernelGGL(k_order_keys, dim3(kgrid), dim3(kBlock), 0, stream(), dc, part, desc, bits, cur, m, ka);
            ok = hip_ok(hipGetLastError(), "order keys") && sort_pairs_u64_u32(ka, kb, cur, other, m, 0, bits);
            uint32_t* t2 = cur;
            cur = other;
            other = t2;
        }
        if (ok && src.Mask) {
            const int8_t nulls = keys[k].Nulls;
            const bool first = nulls == QDB_SORT_NULLS_DEFAULT ? desc != 0 : nulls == QDB_SORT_NULLS_FIRST;
            uint32_t* nk = (uint32_t*)ka;
            uint32_t* nk2 = (uint32_t*)kb;
            hipLaunchKernelGGL(k_null_keys, dim3(kgrid), dim3(kBlock), 0, stream(), dc, first ? 1 : 0, cur, m, nk);
            ok = hip_ok(hipGetLastError(), "null keys") && sort_pairs_u32(nk, nk2, cur, other, m, 1);
            uint32_t* t2 = cur;
            cur = other;
            other = t2;
        }
    }
    if (ok && cur != row_index && m > 0) {
        ok = hip_ok(hipMemcpyAsync(row_index, cur, (size_t)m * 4, hipMemcpyDeviceToDevice, stream()), "sort copy");
    }
    ok = ok && sync("sort");
    release();
    if (!ok) return -1;
    *count = m;
    return 0;
}
