// Host-side runtime of the library: device selection, the stream every
// entry point runs on, a caching device-memory pool, pinned staging buffers,
// error text, and HIP-event timing of kernel families.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <shared_mutex>
#include <string>
#include <vector>

#include "../../include/qdb_gpu.h"

namespace qdbgpu {

// Records an error for qdb_gpu_last_error (thread-local) and returns false.
bool fail(const char* fmt, ...) __attribute__((format(printf, 1, 2)));
void clear_error();
// The calling thread's current error text (empty when none).
const std::string& last_error_text();

// Lazily selects the device and creates the library stream; false on error.
bool ensure_init();
// Streams the library issues work on: [0] the library stream, [1..3] more
// (at the default priority), one per thread that selects it.
constexpr int kMaxStreams = 4;
// The calling thread's stream (the library stream unless the thread chose
// another with set_thread_stream).
hipStream_t stream();
bool set_thread_stream(int index);
int thread_stream_index();
int num_cus();
// Largest LDS allocation of one workgroup on the device, in bytes.
size_t max_lds_per_block();

// Checks a HIP status; on failure records "<what>: <hip error>".
bool hip_ok(hipError_t e, const char* what);

// Caching device allocator: freed blocks are kept per size class and
// reused, so repeated queries do not call hipMalloc.
void* dalloc(size_t bytes);
void dfree(void* p);
size_t dalloc_size(const void* p);
// Returns every cached device block to the HIP runtime (hipFree).
void trim_device_cache();
// Pinned host staging memory (hipHostMalloc), also cached.
void* halloc_pinned(size_t bytes);
void hfree_pinned(void* p);

// Synchronises the library stream; false (with the error text) on failure.
bool sync(const char* what);

// Fills `bytes` of device memory with zero through a kernel (no memset
// node, so the call is safe under stream capture).
bool zero_fill(void* p, size_t bytes);

// Times one kernel family with HIP events on the library stream when timing
// is enabled (qdb_gpu_timing_enable).
class TimedRegion {
public:
    explicit TimedRegion(const char* family);
    ~TimedRegion();
    TimedRegion(const TimedRegion&) = delete;
    TimedRegion& operator=(const TimedRegion&) = delete;

private:
    const char* family_;
    hipEvent_t start_ = nullptr;
};

// Grid size for a grid-stride kernel over `items` with `block` threads: at
// most `per_cu` blocks per CU and no more than the items need.
int grid_for(int64_t items, int block, int per_cu);
// Held exclusively while a thread captures a graph, and shared around the
// device pool's hipMalloc and hipFree: on ROCm an allocation or a free by
// another thread can invalidate a capture in progress.
std::shared_mutex& capture_mutex();

// Exclusive prefix sum of n u32 values (out may alias nothing in in), and a
// stable LSD radix sort of (key, value) u32 pairs over key bits [0, end_bit).
// Both run on the library stream.
bool exclusive_scan_u32(const uint32_t* in, uint32_t* out, int64_t n);
bool sort_pairs_u32(const uint32_t* keys_in, uint32_t* keys_out, const uint32_t* vals_in, uint32_t* vals_out, int64_t n,
    int end_bit);
// Radix sort of u64 keys (all 64 bits), and a stable radix sort of
// (u64 key, u32 value) pairs over key bits [begin_bit, end_bit).
bool sort_keys_u64(const uint64_t* keys_in, uint64_t* keys_out, int64_t n);
bool sort_pairs_u64_u32(const uint64_t* keys_in, uint64_t* keys_out, const uint32_t* vals_in, uint32_t* vals_out, int64_t n,
    int begin_bit, int end_bit);

// Device-resident copy of a host struct (uploaded once, freed with dfree).
void* upload(const void* host, size_t bytes);

// The host heap output rowsets are allocated from (qdb_gpu_set_host_heap:
// the reference's qdb_alloc / qdb_free; malloc / free by default).
void* heap_alloc(size_t bytes);
void heap_free(void* p);
void heap_detach(void* p);

// The buffers of an output rowset while an entry point builds it: its
// TColumns, device blocks (dalloc) and host blocks (heap_alloc).
struct OwnedRowSet {
    std::vector<TColumn> cols;
    std::vector<void*> dev;
    std::vector<void*> host;
};

// Starts `out` as a rowset under construction of `ncols` zeroed TColumns over
// `rows` rows; returns the builder to which buffers are added. Until it is
// published, out->Destroy frees everything the builder holds.
OwnedRowSet* init_owned_rowset(TRowSet* out, int64_t rows, int ncols);
// Moves the device output columns of a rowset under construction to the
// host heap (for OutputOnHost specs).
bool move_rowset_to_host(TRowSet* rs, const QdbValueType* types);
// Turns a rowset under construction into the kernel-owned form of the
// reference (qdb/exec/kernel_rowset.h:14-38): Columns from the host heap,
// Private the owners list, Destroy NULL. False (everything freed) on failure.
bool publish_rowset(TRowSet* rs);
// Frees a kernel-owned rowset (or one under construction).
void destroy_kernel_rowset(TRowSet* rs);

}  // namespace qdbgpu
