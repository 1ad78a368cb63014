// Device-wide scan and stable radix sort (hipCUB over rocPRIM) used by the
// join build and the bucketed aggregate. Temporary storage comes from the library's caching pool.
#include <hip/hip_runtime.h>

#include <hipcub/device/device_radix_sort.hpp>
#include <hipcub/device/device_scan.hpp>

#include "runtime.h"

namespace qdbgpu {

bool exclusive_scan_u32(const uint32_t* in, uint32_t* out, int64_t n)
{
    if (n <= 0) return true;
    if (n > (int64_t)0x7fffffff) return fail("scan of %lld items exceeds 2^31", (long long)n);
    size_t bytes = 0;
    if (!hip_ok(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, in, out, (int)n, stream()), "scan size")) return false;
    void* tmp = dalloc(bytes ? bytes : 16);
    if (!tmp) return false;
    const bool ok = hip_ok(hipcub::DeviceScan::ExclusiveSum(tmp, bytes, in, out, (int)n, stream()), "scan");
    dfree(tmp);
    return ok;
}

bool sort_pairs_u32(const uint32_t* keys_in, uint32_t* keys_out, const uint32_t* vals_in, uint32_t* vals_out, int64_t n,
    int end_bit)
{
    if (n <= 0) return true;
    if (n > (int64_t)0x7fffffff) return fail("sort of %lld items exceeds 2^31", (long long)n);
    size_t bytes = 0;
    if (!hip_ok(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, keys_in, keys_out, vals_in, vals_out, (int)n, 0, end_bit,
            stream()), "sort size")) {
        return false;
    }
    void* tmp = dalloc(bytes ? bytes : 16);
    if (!tmp) return false;
    const bool ok = hip_ok(hipcub::DeviceRadixSort::SortPairs(tmp, bytes, keys_in, keys_out, vals_in, vals_out, (int)n, 0,
        end_bit, stream()), "sort");
    dfree(tmp);
    return ok;
}

bool sort_keys_u64(const uint64_t* keys_in, uint64_t* keys_out, int64_t n)
{
    if (n <= 0) return true;
    if (n > (int64_t)0x7fffffff) return fail("sort of %lld items exceeds 2^31", (long long)n);
    size_t bytes = 0;
    if (!hip_ok(hipcub::DeviceRadixSort::SortKeys(nullptr, bytes, keys_in, keys_out, (int)n, 0, 64, stream()), "sort size")) {
        return false;
    }
    void* tmp = dalloc(bytes ? bytes : 16);
    if (!tmp) return false;
    const bool ok = hip_ok(hipcub::DeviceRadixSort::SortKeys(tmp, bytes, keys_in, keys_out, (int)n, 0, 64, stream()), "sort");
    dfree(tmp);
    return ok;
}

bool sort_pairs_u64_u32(const uint64_t* keys_in, uint64_t* keys_out, const uint32_t* vals_in, uint32_t* vals_out, int64_t n,
    int begin_bit, int end_bit)
{
    if (n <= 0) return true;
    if (n > (int64_t)0x7fffffff) return fail("sort of %lld items exceeds 2^31", (long long)n);
    size_t bytes = 0;
    if (!hip_ok(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, keys_in, keys_out, vals_in, vals_out, (int)n, begin_bit,
            end_bit, stream()), "sort size")) {
        return false;
    }
    void* tmp = dalloc(bytes ? bytes : 16);
    if (!tmp) return false;
    const bool ok = hip_ok(hipcub::DeviceRadixSort::SortPairs(tmp, bytes, keys_in, keys_out, vals_in, vals_out, (int)n,
        begin_bit, end_bit, stream()), "sort");
    dfree(tmp);
    return ok;
}

}  // namespace qdbgpu
