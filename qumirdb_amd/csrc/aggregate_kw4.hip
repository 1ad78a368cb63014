// Aggregate kernels instantiated for packed keys of 4 64-bit word(s); one
// translation unit per key width so that they compile in parallel.
#include "aggregate_kernels.h"

namespace qdbgpu {
namespace agg {

void launch_update_kw4(int S, const AggArgs& A, int grid, size_t lds)
{
    switch (S) {
    case 8: hipLaunchKernelGGL((k_agg_update<4, 8>), dim3(grid), dim3(kBlock), lds, stream(), A); break;
    case 16: hipLaunchKernelGGL((k_agg_update<4, 16>), dim3(grid), dim3(kBlock), lds, stream(), A); break;
    default: hipLaunchKernelGGL((k_agg_update<4, 32>), dim3(grid), dim3(kBlock), lds, stream(), A); break;
    }
}

void launch_batch_kw4(int S, const AggArgs* insts, const BatchItem* items, int nitems, unsigned int* head, int grid, size_t lds)
{
    switch (S) {
    case 8: hipLaunchKernelGGL((k_agg_batch<4, 8>), dim3(grid), dim3(kBlock), lds, stream(), insts, items, nitems, head); break;
    case 16: hipLaunchKernelGGL((k_agg_batch<4, 16>), dim3(grid), dim3(kBlock), lds, stream(), insts, items, nitems, head); break;
    default: hipLaunchKernelGGL((k_agg_batch<4, 32>), dim3(grid), dim3(kBlock), lds, stream(), insts, items, nitems, head); break;
    }
}

void launch_records_kw4(int S, const AggArgs& A, const uint64_t* recs, unsigned int n)
{
    const int grid = grid_for(n, kBlock, 8);
    switch (S) {
    case 8: hipLaunchKernelGGL((k_agg_apply_records<4, 8>), dim3(grid), dim3(kBlock), 0, stream(), A, recs, n); break;
    case 16: hipLaunchKernelGGL((k_agg_apply_records<4, 16>), dim3(grid), dim3(kBlock), 0, stream(), A, recs, n); break;
    default: hipLaunchKernelGGL((k_agg_apply_records<4, 32>), dim3(grid), dim3(kBlock), 0, stream(), A, recs, n); break;
    }
}

void launch_rehash_kw4(const AggArgs& A, const uint64_t* old_recs, const unsigned long long* old_ctrl, uint64_t old_slots)
{
    const int grid = grid_for((int64_t)old_slots, kBlock, 8);
    hipLaunchKernelGGL((k_agg_rehash<4>), dim3(grid), dim3(kBlock), 0, stream(), A, old_recs, old_ctrl, old_slots);
}

}  // namespace agg
}  // namespace qdbgpu
