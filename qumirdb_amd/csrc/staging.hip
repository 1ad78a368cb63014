// Decimal storage on the device: the unscaled values of a decoded DECIMAL
// column narrowed or widened to the width the kernels read.
//
// The reference stores DECIMAL(p <= 18) as an int64 and wider ones as an
// int128 (DecimalStorageOf, qdb/plan/types/decimal.h:80-99), and a Parquet
// scan converts Arrow's decimal arrays to that width
// (DecimalStorageValues, qdb/io/parquet/decimal_storage.cpp:36-60): the
// low bytes of the sign-extended little-endian value. Here the decoded
// column is staged into HBM as Arrow gives it (4, 8, 16 or 32 bytes per
// value) and one kernel writes the storage form; per value it reads the
// source width and writes 8 or 16 bytes (HBM-bound).
#include <hip/hip_runtime.h>

#include "device.h"
#include "runtime.h"

namespace qdbgpu {

namespace {

constexpr int kBlock = 256;

template <int From, int To>
__global__ void __launch_bounds__(kBlock) k_decimal_storage(const uint8_t* src, int64_t n, uint8_t* dst)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += stride) {
        // The low 16 bytes of the value, sign-extended from its width.
        uint64_t lo, hi;
        if (From == 4) {
            const int64_t v = (int64_t)((const int32_t*)src)[i];
            lo = (uint64_t)v;
            hi = v < 0 ? ~0ull : 0ull;
        } else if (From == 8) {
            const int64_t v = ((const int64_t*)src)[i];
            lo = (uint64_t)v;
            hi = v < 0 ? ~0ull : 0ull;
        } else {
            const uint64_t* w = (const uint64_t*)(src + (size_t)From * i);
            lo = w[0];
            hi = w[1];
        }
        if (To == 8) {
            ((uint64_t*)dst)[i] = lo;
        } else {
            uint64_t* d = (uint64_t*)(dst + 16 * i);
            d[0] = lo;
            d[1] = hi;
        }
    }
}

template <int From>
void launch_storage(int to, const uint8_t* src, int64_t n, uint8_t* dst, int grid)
{
    if (to == 8) hipLaunchKernelGGL((k_decimal_storage<From, 8>), dim3(grid), dim3(kBlock), 0, stream(), src, n, dst);
    else hipLaunchKernelGGL((k_decimal_storage<From, 16>), dim3(grid), dim3(kBlock), 0, stream(), src, n, dst);
}

}  // namespace

}  // namespace qdbgpu

using namespace qdbgpu;

extern "C" int qdb_gpu_decimal_storage(const void* src, int32_t from_width, int64_t n, int32_t precision, void* dst)
{
    clear_error();
    if (n > 0 && (!src || !dst)) return fail("null argument"), -1;
    if (from_width != 4 && from_width != 8 && from_width != 16 && from_width != 32) {
        return fail("decimal storage: source width %d (4, 8, 16 or 32)", from_width), -1;
    }
    if (precision < 1 || precision > 76) return fail("decimal storage: precision %d", precision), -1;
    if (precision > 38) return fail("decimal storage: DECIMAL(%d) has no storage (at most 38 digits)", precision), -1;
    if (n <= 0) return 0;
    if (!ensure_init()) return -1;
    const int to = precision <= 18 ? 8 : 16;
    const int grid = grid_for(n, kBlock, 8);
    const uint8_t* s = (const uint8_t*)src;
    uint8_t* d = (uint8_t*)dst;
    switch (from_width) {
    case 4: launch_storage<4>(to, s, n, d, grid); break;
    case 8: launch_storage<8>(to, s, n, d, grid); break;
    case 16: launch_storage<16>(to, s, n, d, grid); break;
    default: launch_storage<32>(to, s, n, d, grid); break;
    }
    return hip_ok(hipGetLastError(), "decimal storage launch") ? 0 : -1;
}
