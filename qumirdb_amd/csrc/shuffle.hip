// Hash shuffle routing and row gathers on gfx950.
//
// THashShuffleTask (qdb/scheduler/tasks/runtime_adapter.cpp:998-1010) sends
// each selected row to output hash % outputs, or, with a routing table, to
// routes[hash % routes] (a partition routed nowhere is an error). Here the
// routing runs over a whole batch on the device: each selected row gets its
// output as a sort key beside its row number, a stable radix sort of those
// (key, row) pairs groups the rows by output in input order, and a
// histogram gives each output's row range. The hash is the batch's Hash
// column when it has one (the reference's UseInputHash), else the caller's
// (the group-hash entry point computes the reference's rh_hash).
//
// A gather then materializes one output's rows as a new device batch
// (every column, validity bitmap and hash of the rows, in that order),
// kernel-owned like every output of the library.
//
// HBM traffic: routing reads 8 B (hash) + 1 B (selection) per row and writes
// 8 B (key, row); the sort moves 8 B per row per pass (one pass for up to 256
// outputs); a gather reads and writes each gathered row's bytes once.
#include <hip/hip_runtime.h>

#include <string.h>

#include <vector>

#include "device.h"
#include "runtime.h"

namespace qdbgpu {

namespace {

constexpr int kBlock = 256;
constexpr int kMaxOutputs = 4096;

// hash % d, exactly, for 1 <= d < 2^32: q = umulhi(hash, floor(2^64 / d)) is
// floor(hash / d) or one less, so one correction step suffices.
__device__ __forceinline__ uint64_t mod_u64(uint64_t hash, uint64_t d, uint64_t recip)
{
    if (d == 1) return 0;
    const uint64_t q = __umul64hi(hash, recip);
    uint64_t r = hash - q * d;
    if (r >= d) r -= d;
    return r;
}

struct RouteArgs {
    const uint64_t* hashes;
    const uint8_t* sel;
    int64_t n;
    uint32_t outputs;
    uint64_t modulus;      // outputs, or the routing table's size
    uint64_t recip;        // floor(2^64 / modulus)
    const int32_t* routes; // device, or null
    uint32_t* keys;        // output of each row (outputs = not selected)
    uint32_t* rows;        // row numbers
    unsigned int* counts;  // rows per output [outputs]
    unsigned int* bad;     // first partition routed nowhere, + 1
};

__global__ void __launch_bounds__(kBlock) k_route(RouteArgs A)
{
    extern __shared__ unsigned int s_count[];
    const int nout = (int)A.outputs;
    for (int o = threadIdx.x; o < nout; o += blockDim.x) s_count[o] = 0u;
    __syncthreads();
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < A.n; i += stride) {
        uint32_t out = A.outputs;
        if (!A.sel || A.sel[i]) {
            const uint64_t part = mod_u64(A.hashes[i], A.modulus, A.recip);
            if (A.routes) {
                const int32_t r = A.routes[part];
                if (r >= 0 && (uint32_t)r < A.outputs) out = (uint32_t)r;
                else atomicCAS(A.bad, 0u, (unsigned int)part + 1u);
            } else {
                out = (uint32_t)part;
            }
        }
        A.keys[i] = out;
        A.rows[i] = (uint32_t)i;
        if (out < A.outputs) atomicAdd(&s_count[out], 1u);
    }
    __syncthreads();
    for (int o = threadIdx.x; o < nout; o += blockDim.x) {
        if (s_count[o]) atomicAdd(&A.counts[o], s_count[o]);
    }
}

// Gather of one column's fixed-width values.
__global__ void __launch_bounds__(kBlock) k_gather_fixed(const char* src, int width, const uint32_t* idx, int64_t n, char* dst)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const int64_t r = idx[i];
        switch (width) {
        case 1: dst[i] = src[r]; break;
        case 2: ((uint16_t*)dst)[i] = ((const uint16_t*)src)[r]; break;
        case 4: ((uint32_t*)dst)[i] = ((const uint32_t*)src)[r]; break;
        case 8: ((uint64_t*)dst)[i] = ((const uint64_t*)src)[r]; break;
        default: {
            const uint64_t* s = (const uint64_t*)(src + 16 * r);
            uint64_t* d = (uint64_t*)(dst + 16 * i);
            d[0] = s[0];
            d[1] = s[1];
        }
        }
    }
}

// Gather of a bitmap (bool data or validity) into a new bitmap: each wave
// writes whole 64-bit words, so no two lanes share a word.
__global__ void __launch_bounds__(kBlock) k_gather_bits(const uint8_t* src, int64_t src_off, const uint32_t* idx, int64_t n,
    uint64_t* dst)
{
    const int lane = lane_id();
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t base = (int64_t)blockIdx.x * blockDim.x + threadIdx.x - lane; base < n; base += stride) {
        const int64_t i = base + lane;
        bool bit = false;
        if (i < n) {
            const int64_t b = (int64_t)idx[i] + src_off;
            bit = (src[b >> 3] >> (b & 7)) & 1;
        }
        const uint64_t word = __ballot(bit);
        if (lane == 0) dst[base >> 6] = word;
    }
}

__global__ void __launch_bounds__(kBlock) k_gather_sel(const uint8_t* src, const uint32_t* idx, int64_t n, uint8_t* dst)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[idx[i]];
}

}  // namespace

}  // namespace qdbgpu

using namespace qdbgpu;

extern "C" int qdb_gpu_shuffle_partition(const TRowSet* rs, const uint64_t* hashes, int32_t outputs, const int32_t* routes,
    int32_t nroutes, uint32_t* row_index, int64_t* offsets)
{
    clear_error();
    if (!rs || !row_index || !offsets) return fail("null argument"), -1;
    if (outputs < 1 || outputs > kMaxOutputs) return fail("shuffle: %d outputs (1..%d)", outputs, kMaxOutputs), -1;
    if (routes && nroutes < 1) return fail("shuffle: empty routing table"), -1;
    if (!ensure_init()) return -1;
    const int64_t n = rs->RowCount;
    for (int32_t o = 0; o <= outputs; ++o) offsets[o] = 0;
    if (n <= 0) return 0;
    if (n > (int64_t)0x7fffffff) return fail("shuffle: batches are limited to 2^31 - 1 rows"), -1;
    // The batch's own hashes first (computed upstream on the same key).
    const uint64_t* h = rs->Hash ? rs->Hash : hashes;
    if (!h) return fail("shuffle: the batch has no hashes and none were given"), -1;
    RouteArgs A;
    memset(&A, 0, sizeof(A));
    A.hashes = h;
    A.sel = rs->Selection;
    A.n = n;
    A.outputs = (uint32_t)outputs;
    A.modulus = routes ? (uint64_t)nroutes : (uint64_t)outputs;
    A.recip = A.modulus > 1 ? (uint64_t)((((unsigned __int128)1) << 64) / A.modulus) : 0;
    int32_t* d_routes = nullptr;
    uint32_t* keys = (uint32_t*)dalloc((size_t)n * 4);
    uint32_t* keys_sorted = (uint32_t*)dalloc((size_t)n * 4);
    uint32_t* rows = (uint32_t*)dalloc((size_t)n * 4);
    unsigned int* counts = (unsigned int*)dalloc((size_t)(outputs + 1) * 4);
    bool ok = keys && keys_sorted && rows && counts && zero_fill(counts, (size_t)(outputs + 1) * 4);
    if (ok && routes) {
        d_routes = (int32_t*)upload(routes, (size_t)nroutes * 4);
        ok = d_routes != nullptr;
    }
    if (ok) {
        A.routes = d_routes;
        A.keys = keys;
        A.rows = rows;
        A.counts = counts;
        A.bad = counts + outputs;
        hipLaunchKernelGGL(k_route, dim3(grid_for(n, kBlock, 8)), dim3(kBlock), (size_t)outputs * 4, stream(), A);
        ok = hip_ok(hipGetLastError(), "route launch");
    }
    // Stable by output: sort keys of ceil(log2(outputs + 1)) bits.
    int bits = 1;
    while ((1u << bits) <= (uint32_t)outputs) ++bits;
    ok = ok && sort_pairs_u32(keys, keys_sorted, rows, row_index, n, bits);
    std::vector<unsigned int> h_counts((size_t)outputs + 1, 0u);
    ok = ok && hip_ok(hipMemcpyAsync(h_counts.data(), counts, (size_t)(outputs + 1) * 4, hipMemcpyDeviceToHost, stream()), "counts")
        && sync("shuffle");
    dfree(keys);
    dfree(keys_sorted);
    dfree(rows);
    dfree(counts);
    dfree(d_routes);
    if (!ok) return -1;
    if (h_counts[(size_t)outputs]) {
        return fail("hash shuffle got a row of partition %u, which it routes nowhere", h_counts[(size_t)outputs] - 1u), -1;
    }
    int64_t at = 0;
    for (int32_t o = 0; o < outputs; ++o) {
        offsets[o] = at;
        at += h_counts[(size_t)o];
    }
    offsets[outputs] = at;
    return 0;
}

extern "C" int qdb_gpu_gather_rowset(const TRowSet* rs, const QdbColumnSchema* types, const uint32_t* row_index, int64_t count,
    TRowSet* out)
{
    clear_error();
    if (!rs || !types || !out || (count > 0 && !row_index)) return fail("null argument"), -1;
    if (rs->ColumnCount > types->NumColumns) return fail("gather: %lld columns, schema has %d", (long long)rs->ColumnCount, types->NumColumns), -1;
    if (!ensure_init()) return -1;
    const int64_t n = count > 0 ? count : 0;
    OwnedRowSet* o = init_owned_rowset(out, n, (int)rs->ColumnCount);
    const size_t words = (size_t)((n + 63) / 64) + 1;
    const int grid = grid_for(n > 0 ? n : 1, kBlock, 8);
    bool ok = true;
    for (int64_t c = 0; ok && c < rs->ColumnCount; ++c) {
        const TColumn& src = rs->Columns[c];
        const QdbValueType& t = types->Types[c];
        TColumn& dst = o->cols[(size_t)c];
        if (t.Kind == QDB_KIND_STRING) {
            ok = fail("gather: string columns are not supported");
            break;
        }
        const bool bits = t.Kind == QDB_KIND_BOOL;
        const size_t bytes = bits ? words * 8 : (size_t)(n > 0 ? n : 1) * (size_t)t.Width;
        char* data = (char*)dalloc(bytes);
        ok = data != nullptr;
        if (!ok) break;
        o->dev.push_back(data);
        dst.Data = data;
        if (n > 0) {
            if (bits) {
                hipLaunchKernelGGL(k_gather_bits, dim3(grid), dim3(kBlock), 0, stream(), (const uint8_t*)src.Data,
                    (int64_t)src.DataBitOffset, row_index, n, (uint64_t*)data);
            } else {
                hipLaunchKernelGGL(k_gather_fixed, dim3(grid), dim3(kBlock), 0, stream(), src.Data, t.Width, row_index, n, data);
            }
        }
        if (src.Mask) {
            uint8_t* mask = (uint8_t*)dalloc(words * 8);
            ok = mask != nullptr;
            if (!ok) break;
            o->dev.push_back(mask);
            dst.Mask = mask;
            if (n > 0) {
                hipLaunchKernelGGL(k_gather_bits, dim3(grid), dim3(kBlock), 0, stream(), src.Mask, (int64_t)src.MaskBitOffset,
                    row_index, n, (uint64_t*)mask);
            }
        }
        ok = ok && hip_ok(hipGetLastError(), "gather launch");
    }
    if (ok && rs->Hash && n > 0) {
        uint64_t* hh = (uint64_t*)dalloc((size_t)n * 8);
        ok = hh != nullptr;
        if (ok) {
            o->dev.push_back(hh);
            hipLaunchKernelGGL(k_gather_fixed, dim3(grid), dim3(kBlock), 0, stream(), (const char*)rs->Hash, 8, row_index, n, (char*)hh);
            out->Hash = hh;
        }
    }
    // The gathered rows are the selected ones; a selection comes along only
    // when the caller gathered unselected rows too.
    if (ok && rs->Selection && n > 0) {
        uint8_t* sel = (uint8_t*)dalloc((size_t)n);
        ok = sel != nullptr;
        if (ok) {
            o->dev.push_back(sel);
            hipLaunchKernelGGL(k_gather_sel, dim3(grid), dim3(kBlock), 0, stream(), rs->Selection, row_index, n, sel);
            out->Selection = sel;
        }
    }
    ok = ok && hip_ok(hipGetLastError(), "gather launch") && sync("gather");
    if (!ok || !publish_rowset(out)) {
        destroy_kernel_rowset(out);
        memset(out, 0, sizeof(*out));
        return -1;
    }
    out->Destroy = qdb_gpu_destroy_kernel_rowset;
    return 0;
}
