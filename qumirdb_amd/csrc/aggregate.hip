// Hash aggregation on gfx950: agg_dispatch / agg_finish_rowset.
//
// Reference: agg_dispatch (qdb/kernel/gen.cpp:2075-2299), reducers
// (gen.cpp:3301-3592), finalize (gen.cpp:2836-3132), op-codes and the
// HashTable blob (qdb/kernel/compiler.h:208-233,603-616,
// qdb/exec/aggregate_exec.cpp:18-39).
//
// Device design. A group's state is a record of 64-bit words, each with a
// combine operation (add, min, max, f64 add), so that a row's contribution,
// a wave's partial, a workgroup's partial and the global state all combine
// the same way:
//   - integer SUM: one wrapping i64 add word (the reference's i64 state);
//   - DECIMAL SUM (exact i128): three add words holding bits [0,32), [32,64)
//     and [64,128) of every value, recombined modulo 2^128 at the end;
//   - f64 SUM: four 32-bit-chunk limbs (exact over a 128-bit window) plus an
//     f64 side word for anything outside it, rounded once at the end;
//   - MIN/MAX: i64 or order-encoded f64 words with atomic min/max;
//   - COUNT and the group's row count: add words.
// Rows are pre-aggregated per wave (ballot/match on equal keys), then in a
// workgroup-private open-addressing table in LDS; keys that miss a full LDS
// table, and every LDS group at the end, go to the HBM table (64-bit CAS on a
// control word per slot, dense group records). A launch that fills the HBM
// table records what it could not insert; the host grows the table and
// replays exactly those rows and records, so every row is applied once.
#include <hip/hip_runtime.h>

#include <stdlib.h>
#include <string.h>

#include <memory>
#include <mutex>
#include <utility>
#include <algorithm>
#include <atomic>
#include <chrono>
#include <thread>
#include <vector>

#include "aggregate_kernels.h"
#include "autotune.h"
#include "jit.h"
#include "slots.h"

namespace qdbgpu {

namespace {

using namespace agg;

typedef long long ll2 __attribute__((ext_vector_type(2)));

constexpr int kMaxStageWords = 8;     // staged row words the bucket sort moves
// Rows of a batch from which the bucketed path pays (below, its fixed
// costs -- four launches, three host synchronisations -- exceed the atomic
// path's time), and table capacity
// (slots, from Init / ExpectRowGroups or growth) that selects it.
constexpr int64_t kSortedMinRows = (int64_t)1 << 22;
constexpr uint64_t kSortedMinSlots = (uint64_t)1 << 20;
constexpr int64_t kSortedItemRows = (int64_t)1 << 16;
constexpr int kSortedBlock = 1024;    // threads of a bucket-aggregation workgroup
// Rows of a bucketed batch from which its hot keys are sampled
// (QDB_GPU_HOT_MIN_ROWS moves it, down to the sample's size).
constexpr int64_t kHotMinRows = (int64_t)1 << 22;
bool hot_keys_enabled();
// Partial records of an update (AggArgs::part_recs): at most kPartCap
// groups per workgroup slot, kPartsPerReducer slots per workgroup of the
// combining launch; used for tables of at most kPartMaxSlots slots that
// hold at most kPartCap groups.
constexpr int kPartCap = 64;
constexpr int kPartsPerReducer = 32;
constexpr uint64_t kPartMaxSlots = 4096;
static_assert(kPartsPerReducer <= agg::kMaxPartsPerReducer, "combining workgroups take at most kMaxPartsPerReducer slots");

struct StageLayout {
    int words = 0;                    // 0: the plan does not stage
    int valid_word = -1;
    int off[kMaxSrc] = {};
    int cls[kMaxSrc] = {};
    bool two[kMaxSrc] = {};
};

__device__ __forceinline__ void store_width(char* data, int width, unsigned int g, u128 v)
{
    switch (width) {
    case 1: ((uint8_t*)data)[g] = (uint8_t)v; break;
    case 2: ((uint16_t*)data)[g] = (uint16_t)v; break;
    case 4: ((uint32_t*)data)[g] = (uint32_t)v; break;
    case 8: ((uint64_t*)data)[g] = (uint64_t)v; break;
    default: {
        uint64_t* p = (uint64_t*)(data + 16 * (size_t)g);
        p[0] = (uint64_t)v;
        p[1] = (uint64_t)(v >> 64);
    }
    }
}

// Output groups [base, base + 64) of F, one per lane (base a multiple of
// 64): the wave reads their records together (word k of 64 records per
// load) and writes each validity bitmap's 64 bits, and a bool key's data
// bits, as one word from its ballot. `index`: the slot of output group
// g - pend_n (null: the group's own number).
__device__ __forceinline__ void finalize_groups_wave(const FinArgs& F, const uint32_t* index, uint64_t base)
{
    const int lane = lane_id();
    {
        const unsigned int g = (unsigned int)(base + lane);
        const bool live = g < F.groups;
        const uint64_t* r = F.recs;
        // The run holding the wave's first pending group (one binary search
        // per wave: the last k with pend_prefix[k] <= base); runs hold
        // hundreds of groups, so a lane walks at most a step or two from it.
        int run0 = 0;
        if (base < F.pend_n) {
            int lo = 0, hi = F.pend_runs - 1;
            const unsigned int b0 = (unsigned int)base;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (F.pend_prefix[mid] <= b0) lo = mid;
                else hi = mid - 1;
            }
            run0 = lo;
        }
        if (live) {
            if (g < F.pend_n) {
                int lo = run0;
                while (lo + 1 < F.pend_runs && F.pend_prefix[lo + 1] <= g) ++lo;
                r = F.pend_recs + ((uint64_t)F.pend_items[lo].row0 + (g - F.pend_prefix[lo])) * (uint64_t)F.R;
            } else {
                const unsigned int gg = g - F.pend_n;
                r = F.recs + (size_t)(index ? index[gg] : gg) * (size_t)F.stride;
            }
        }
        const uint64_t word_at = base >> 6;  // the wave's 64-bit bitmap word
        for (int f = 0; f < F.nkeys; ++f) {
            const KeyFieldDesc kf = F.keys[f];
            bool is_null = true;
            u128 v = 0;
            if (live) {
                is_null = kf.nullable && ((r[kf.null_word] >> kf.null_bit) & 1ull);
                if (!is_null) {
                    if (kf.width == 16) {
                        v = ((u128)r[kf.word + 1] << 64) | (u128)r[kf.word];
                    } else {
                        uint64_t m = kf.width == 8 ? ~0ull : ((1ull << (8 * kf.width)) - 1ull);
                        v = (r[kf.word] >> kf.shift) & m;
                    }
                }
            }
            if (kf.kind == QDB_KIND_BOOL) {
                // bool keys are emitted as an LSB bitmap in the data buffer.
                const uint64_t bits = __ballot(live && v != 0);
                if (lane == 0) ((uint64_t*)F.out_data[f])[word_at] = bits;
            } else if (live) {
                store_width(F.out_data[f], F.key_out_width[f], g, v);
            }
            const uint64_t valid_bits = __ballot(live && !is_null);
            if (F.out_mask[f] && lane == 0) ((uint64_t*)F.out_mask[f])[word_at] = valid_bits;
        }
        const uint64_t* s = r + F.KW;
        const int64_t rows = live ? (int64_t)s[F.rows_word] : 0;
        for (int a = 0; a < F.nred; ++a) {
            const RedOut ro = F.red[a];
            const int col = F.nkeys + a;
            bool valid = false;
            u128 v = 0;
            if (live) {
                const int64_t contributions = ro.wvalid >= 0 ? (int64_t)s[ro.wvalid] : rows;
                valid = true;
                switch (ro.kind) {
                case RK_COUNT_ROWS: v = (u128)(uint64_t)rows; break;
                case RK_COUNT_VALID: v = (u128)s[ro.w0]; break;
                case RK_SUM64:
                    valid = contributions > 0;
                    v = (u128)s[ro.w0];
                    break;
                case RK_SUM128:
                    valid = contributions > 0;
                    v = (u128)s[ro.w0] + ((u128)s[ro.w0 + 1] << 32) + ((u128)s[ro.w0 + 2] << 64);
                    break;
                case RK_SUMF64: {
                    valid = contributions > 0;
                    int64_t limb[kLimbs];
                    for (int k = 0; k < kLimbs; ++k) limb[k] = (int64_t)s[ro.w0 + k];
                    double exact = limbs_to_f64(limb);
                    double side = __longlong_as_double((long long)s[ro.w0 + kLimbs]);
                    double total = side != 0.0 || side != side ? exact + side : exact;
                    v = (u128)(uint64_t)__double_as_longlong(total);
                    break;
                }
                case RK_MINMAX_I64: {
                    valid = contributions > 0;
                    int64_t x = (int64_t)s[ro.w0];
                    v = (u128)(i128)x;
                    break;
                }
                case RK_MINMAX_I128:
                    valid = contributions > 0;
                    v = ((u128)s[ro.w0 + 1] << 64) | (u128)s[ro.w0];
                    break;
                case RK_MINMAX_F64: {
                    valid = contributions > 0;
                    uint64_t k = s[ro.w0];
                    double d;
                    if (ro.is_max) d = k == 0ull ? __longlong_as_double(0x7ff8000000000000ll) : f64_from_order(k);
                    else d = k >= 0xfff8000000000001ull ? __longlong_as_double(0x7ff8000000000000ll) : f64_from_order(k);
                    v = (u128)(uint64_t)__double_as_longlong(d);
                    break;
                }
                default: break;
                }
                if (!valid) v = 0;
                store_width(F.out_data[col], ro.out_width, g, v);
            }
            const uint64_t valid_bits = __ballot(valid);
            if (F.out_mask[col] && lane == 0) ((uint64_t*)F.out_mask[col])[word_at] = valid_bits;
        }
    }
}

__global__ void __launch_bounds__(kBlock) k_agg_finalize(FinArgs F)
{
    const int lane = lane_id();
    const uint64_t step = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t base = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x - lane; base < F.groups; base += step) {
        finalize_groups_wave(F, F.index, base);
    }
}

// The finalize of many small tables of one aggregate in one launch (a
// batched finish): workgroup b takes table b whole. It lists the slots
// that hold groups, in slot order, in LDS (the occupancy words split into
// one contiguous range per thread, counted, scanned, then each thread
// writes its range's slots; the extra record's slot last), then its waves
// write the output groups as k_agg_finalize does. One-word keys only, no
// pending runs, at most kSmallFinMaxSlots slots and kSmallFinMaxGroups
// groups.
constexpr int kSmallFinBlock = 1024;
constexpr unsigned int kSmallFinMaxGroups = 16384;
constexpr uint64_t kSmallFinMaxSlots = 65536;

struct SmallFin {
    const uint32_t* occ;   // one bit per slot
    uint64_t slots;        // the slot table's capacity (the extra record is slot `slots`)
    int extra_used;
    int pad;
};

__global__ void __launch_bounds__(kSmallFinBlock) k_agg_finalize_small(const FinArgs* tabs, const SmallFin* info)
{
    __shared__ uint32_t s_index[kSmallFinMaxGroups];
    __shared__ unsigned int s_wave[kSmallFinBlock / kWave];
    const FinArgs& F = tabs[blockIdx.x];
    const SmallFin I = info[blockIdx.x];
    const int lane = lane_id(), wave = threadIdx.x / kWave;
    const uint32_t nw = (uint32_t)((I.slots + 31) / 32);
    const uint32_t per = (nw + blockDim.x - 1) / blockDim.x;
    const uint32_t w0 = threadIdx.x * per, w1 = w0 + per < nw ? w0 + per : nw;
    auto word_at = [&](uint32_t w) -> uint32_t {
        uint32_t v = I.occ[w];
        const uint64_t first = (uint64_t)w * 32;
        if (first + 32 > I.slots) v &= (uint32_t)((1ull << (I.slots - first)) - 1ull);
        return v;
    };
    unsigned int cnt = 0;
    for (uint32_t w = w0; w < w1; ++w) cnt += (unsigned int)__popc(word_at(w));
    // Exclusive scan of the threads' counts: within the wave, then the
    // waves before it.
    unsigned int incl = cnt;
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        const unsigned int y = (unsigned int)__shfl_up((int)incl, o, kWave);
        if (lane >= o) incl += y;
    }
    if (lane == kWave - 1) s_wave[wave] = incl;
    __syncthreads();
    unsigned int before = incl - cnt;
    unsigned int all = 0;
    for (int v = 0; v < (int)(blockDim.x / kWave); ++v) {
        before += v < wave ? s_wave[v] : 0u;
        all += s_wave[v];
    }
    unsigned int at = before;
    for (uint32_t w = w0; w < w1; ++w) {
        uint32_t bits = word_at(w);
        while (bits) {
            const int b = __ffs(bits) - 1;
            bits &= bits - 1u;
            if (at < kSmallFinMaxGroups) s_index[at] = w * 32u + (uint32_t)b;
            ++at;
        }
    }
    if (threadIdx.x == 0 && I.extra_used && all < kSmallFinMaxGroups) s_index[all] = (uint32_t)I.slots;
    __syncthreads();
    const uint64_t step = (uint64_t)blockDim.x;
    for (uint64_t base = (uint64_t)wave * kWave; base < F.groups; base += step) finalize_groups_wave(F, s_index, base);
}

// Fills `n` records of `stride` words with the empty-table pattern (the
// empty key marker, then every state word's identity), 16 bytes per store.
struct RecordPattern {
    uint64_t w[2 * kMaxWords];
};

__global__ void __launch_bounds__(kBlock) k_fill_records(uint64_t* recs, uint64_t words, int stride, RecordPattern pat)
{
    const uint64_t pairs = words / 2;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < pairs; i += (uint64_t)gridDim.x * blockDim.x) {
        const int w = (int)((2 * i) % (uint64_t)stride);
        const ll2 v = {(long long)pat.w[w], (long long)pat.w[w + 1]};
        *(ll2*)(recs + 2 * i) = v;
    }
}

constexpr int kOccChunk = 4096;   // slots per workgroup in the occupancy passes

// One-word keys: the table's occupancy bitmap (4 bytes per 32 slots
// rather than one record line per slot); wider keys: their control words.
template <int KW>
__device__ __forceinline__ bool occupied(const uint64_t* recs, const unsigned long long* ctrl, const uint32_t* occ, int stride,
    uint64_t slots, int extra_used, uint64_t s)
{
    if (s >= slots) return s == slots && extra_used;
    if (KW == 1) return ((occ[s >> 5] >> (s & 31u)) & 1u) != 0u;
    return slot_used<KW>(recs, ctrl, stride, s);
}

// Groups per chunk of kOccChunk slots (slots [0, slots] with the extra
// record last).
template <int KW>
__global__ void __launch_bounds__(kBlock) k_occ_counts(const uint64_t* recs, const unsigned long long* ctrl, const uint32_t* occ,
    int stride, uint64_t slots, int extra_used, uint32_t* counts)
{
    __shared__ unsigned int total;
    if (threadIdx.x == 0) total = 0u;
    __syncthreads();
    const uint64_t base = (uint64_t)blockIdx.x * kOccChunk;
    unsigned int n = 0;
    for (int k = threadIdx.x; k < kOccChunk; k += blockDim.x) {
        n += occupied<KW>(recs, ctrl, occ, stride, slots, extra_used, base + (uint64_t)k) ? 1u : 0u;
    }
    atomicAdd(&total, n);
    __syncthreads();
    if (threadIdx.x == 0) counts[blockIdx.x] = total;
}

// The slot of every group, in slot order: chunk offset + rank in the chunk.
template <int KW>
__global__ void __launch_bounds__(kBlock) k_occ_index(const uint64_t* recs, const unsigned long long* ctrl, const uint32_t* occ,
    int stride, uint64_t slots, int extra_used, const uint32_t* offsets, uint32_t* index)
{
    __shared__ unsigned int wave_total[kBlock / kWave];
    const uint64_t base = (uint64_t)blockIdx.x * kOccChunk;
    const int lane = lane_id(), wave = threadIdx.x / kWave;
    unsigned int out = offsets[blockIdx.x];
    for (int k0 = 0; k0 < kOccChunk; k0 += kBlock) {
        const uint64_t s = base + (uint64_t)(k0 + threadIdx.x);
        const bool used = occupied<KW>(recs, ctrl, occ, stride, slots, extra_used, s);
        const uint64_t m = __ballot(used);
        if (lane == 0) wave_total[wave] = (unsigned int)__popcll(m);
        __syncthreads();
        unsigned int before = 0, all = 0;
        for (int v = 0; v < kBlock / kWave; ++v) {
            before += v < wave ? wave_total[v] : 0u;
            all += wave_total[v];
        }
        if (used) index[out + before + (unsigned int)__popcll(m & ((1ull << lane) - 1ull))] = (uint32_t)s;
        out += all;
        __syncthreads();
    }
}

// starts[b] = offs[b * nblk]: where bucket b's run of the first row range
// starts in the partitioned rows.
__global__ void __launch_bounds__(kBlock) k_gather_starts(const unsigned int* offs, unsigned int nb, unsigned int nblk, uint32_t* starts)
{
    for (unsigned int b = blockIdx.x * blockDim.x + threadIdx.x; b < nb; b += gridDim.x * blockDim.x) {
        starts[b] = offs[(uint64_t)b * nblk];
    }
}


// ---------------------------------------------------------------------------
// Host: plans, tables, entry points.
// ---------------------------------------------------------------------------

// A compiled aggregate spec.
struct AggPlan {
    QdbAggregateSpec spec;
    int KW = 1;
    int S = 8;                 // word-count bucket (template argument)
    AggArgs proto;             // launch arguments without per-batch fields
    int slot_col[kMaxDevCols]; // device slot -> input column
    int nslots = 0;
    DevProgram* d_pred = nullptr;
    DevProgram* d_progs = nullptr;
    FinArgs fin;               // finalize arguments without buffers
    int nout = 0;
    QdbValueType out_types[QDB_AGG_MAX_KEYS + QDB_AGG_MAX_AGGS];
    bool has_pairs = false;
    uint64_t word_init_host[kMaxWords];
    // Host copies of the programs and the value type of every device slot,
    // for the kernel generated for this aggregate.
    bool has_pred = false;
    DevProgram h_pred;
    std::vector<DevProgram> h_progs;
    QdbValueType slot_type[kMaxDevCols];
    std::string jit_src;
    bool jit_tried = false;
    // The partial-state projection (qdb_gpu_agg_partial_states): compiled at
    // its first use, its project slot released with the plan.
    std::mutex partial_mu;
    std::mutex jit_mu;      // jit_tried / jit_src, set by the first caller
    bool partial_tried = false;
    QdbProjectFn partial_fn = nullptr;
    int partial_slot = -1;

    ~AggPlan()
    {
        dfree(d_pred);
        dfree(d_progs);
        if (partial_slot >= 0) qdb_gpu_release_slot(QDB_SLOT_PROJECT, partial_slot);
    }
};

// The device table behind a HashTable blob.
struct AggTable {
    AggPlan* plan = nullptr;
    unsigned long long* ctrl = nullptr; // wider keys only
    uint32_t* occ = nullptr;            // one-word keys: a bit per slot that holds a group
    uint64_t capacity = 0;              // slots (a power of two); the extra record follows
    uint64_t* recs = nullptr;
    int rstride = 1;                    // words per record
    unsigned int extra_used = 0;        // the extra record holds a group
    unsigned int* counters = nullptr;   // device [4]
    unsigned int groups = 0;            // groups in the slot table
    int64_t expect_groups = 0;          // ExpectRowGroups hint
    int match_rounds = 0;
    // A table made with many slots gets its buffers at first need: a
    // bucketed update into it while it is empty keeps the groups of the
    // buckets that one workgroup held whole as record runs (pend_*), which
    // finish reads directly and any other use first merges into the slot
    // table. Their keys occur in no other run and not in the slot table.
    bool allocated = true;
    uint64_t cap_hint = 0;              // the slots asked for at Init (at least 1024)
    int64_t init_slots = 0;             // the Init argument itself
    uint64_t* pend_recs = nullptr;      // records of R words
    SortedItem* pend_items = nullptr;   // device: the runs (row0 = first record)
    uint32_t* pend_counts = nullptr;    // device: records per run
    uint32_t* pend_prefix = nullptr;    // device: exclusive prefix of the counts (runs + 1)
    int pend_runs = 0;
    unsigned int pend_groups = 0;
};

// Groups of a table: the slot table's and the pending runs'.
inline unsigned int table_size(const AggTable* T)
{
    return T->groups + T->pend_groups;
}

void table_drop_pending(AggTable* T)
{
    dfree(T->pend_recs);
    dfree(T->pend_items);
    dfree(T->pend_counts);
    dfree(T->pend_prefix);
    T->pend_recs = nullptr;
    T->pend_items = nullptr;
    T->pend_counts = nullptr;
    T->pend_prefix = nullptr;
    T->pend_runs = 0;
    T->pend_groups = 0;
}

// The HashTable blob (qdb/exec/aggregate_exec.cpp:18-39). The processor reads
// Size; the first word holds the device table.
struct HtBlob {
    AggTable* table;
    void* unused_ptrs[7];
    int64_t Capacity;
    int64_t Size;
    int64_t NumAggs;
    int64_t NumKeys;
    int64_t KeySize;
    uint8_t RowLayout[QDB_ROW_ARENA_SIZE + QDB_ROW_TABLE_SIZE + QDB_ROW_SCRATCH_SIZE];
    int64_t RowSwitchGroups;
    int64_t RowExpectedGroups;
    int64_t RowRebuilds;
    int64_t RowMovedFromSlots;
};
static_assert(sizeof(HtBlob) == QDB_HASH_TABLE_SIZE, "HashTable blob size");

SlotTable<AggPlan>& agg_slots()
{
    static SlotTable<AggPlan>* t = new SlotTable<AggPlan>();
    return *t;
}

uint64_t word_init_host(const WordDesc& d)
{
    if (d.kind == WK_PAIR_LO) return d.part ? 0ull : ~0ull;
    if (d.kind == WK_PAIR_HI) return d.part ? 0x8000000000000000ull : 0x7fffffffffffffffull;
    switch (d.op) {
    case OP_MIN_I64: return 0x7fffffffffffffffull;
    case OP_MAX_I64: return 0x8000000000000000ull;
    case OP_MIN_U64: return ~0ull;
    default: return 0ull;
    }
}

int key_bytes_of(const QdbValueType& t)
{
    if (t.Kind == QDB_KIND_BOOL) return 1;
    if (t.Kind == QDB_KIND_FLOAT) return t.Width == 4 ? 4 : 8;
    return t.Width;
}

// Result types, AggregateResultType (qdb/plan/ops/aggregate.cpp:21-61).
void output_types(const QdbAggregateSpec& spec, const QdbValueType* col_types, QdbValueType* out)
{
    int n = 0;
    for (int k = 0; k < spec.NumKeys; ++k) out[n++] = col_types[spec.KeyColumns[k]];
    for (int a = 0; a < spec.NumAggs; ++a) {
        QdbValueType t;
        memset(&t, 0, sizeof(t));
        const QdbAggregateFn& fn = spec.Aggs[a];
        if (fn.Func == QDB_AGG_COUNT_STAR || fn.Func == QDB_AGG_COUNT) {
            t.Kind = QDB_KIND_SIGNED;
            t.Width = 8;
        } else {
            const QdbValueType& in = col_types[fn.Column];
            t.Nullable = in.Nullable || spec.NumEmptyInputGroups > 0;
            if (in.Kind == QDB_KIND_FLOAT) {
                t.Kind = QDB_KIND_FLOAT;
                t.Width = 8;
            } else if (in.Kind == QDB_KIND_DECIMAL) {
                t.Kind = QDB_KIND_DECIMAL;
                t.Scale = in.Scale;
                t.Width = fn.Func == QDB_AGG_SUM ? 16 : in.Width;
            } else if (in.Kind == QDB_KIND_DATE) {
                t.Kind = QDB_KIND_DATE;
                t.Width = 8;
            } else {
                t.Kind = QDB_KIND_SIGNED;
                t.Width = 8;
            }
        }
        out[n++] = t;
    }
}

bool compile_plan(const QdbAggregateSpec& spec, AggPlan* P, bool on_device = true)
{
    P->spec = spec;
    AggArgs& A = P->proto;
    memset(&A, 0, sizeof(A));
    const int nin = spec.Input.NumColumns;
    // Types of input and computed columns.
    QdbValueType types[64 + QDB_PROJECT_MAX_OUTPUTS];
    for (int c = 0; c < nin; ++c) types[c] = spec.Input.Types[c];
    for (int k = 0; k < spec.NumComputed; ++k) types[nin + k] = spec.ComputedTypes[k];
    const int ncols = nin + spec.NumComputed;
    auto bad_col = [&](int c) { return c < 0 || c >= ncols; };

    int col_slot[64];
    for (int c = 0; c < 64; ++c) col_slot[c] = -1;
    auto slot_of = [&](int c) -> int {
        if (c < 0 || c >= nin) return -1;
        if (col_slot[c] < 0) {
            if (P->nslots >= kMaxDevCols) return -2;
            col_slot[c] = P->nslots;
            P->slot_type[P->nslots] = spec.Input.Types[c];
            P->slot_col[P->nslots++] = c;
        }
        return col_slot[c];
    };
    // Programs: remap input columns to device slots.
    auto remap = [&](const QdbExprProgram& in, QdbExprProgram* out) -> bool {
        *out = in;
        if (in.NumInsns < 0 || in.NumInsns > QDB_EXPR_MAX_INSNS) return fail("program has %d instructions", in.NumInsns);
        for (int k = 0; k < in.NumInsns; ++k) {
            QdbExprInsn& ins = out->Insns[k];
            if (ins.Dst >= QDB_EXPR_MAX_REGS || ins.A >= QDB_EXPR_MAX_REGS || ins.B >= QDB_EXPR_MAX_REGS) {
                return fail("program register out of range");
            }
            if (ins.Op == QDB_EXPR_DIV && ins.Class != QDB_REG_F64) return fail("DIV is an f64 division (register class F64)");
            if (ins.Op == QDB_EXPR_COLUMN) {
                if (ins.Column < 0 || ins.Column >= nin) return fail("program column %d is not an input column", ins.Column);
                int s = slot_of(ins.Column);
                if (s < 0) return fail("too many columns in one aggregate (max %d)", kMaxDevCols);
                ins.Column = s;
            }
        }
        return true;
    };
    if (spec.NumKeys < 0 || spec.NumKeys > QDB_AGG_MAX_KEYS) return fail("bad key count %d", spec.NumKeys);
    if (spec.NumAggs < 0 || spec.NumAggs > QDB_AGG_MAX_AGGS) return fail("bad aggregate count %d", spec.NumAggs);
    if (spec.NumEmptyInputGroups < 0 || spec.NumEmptyInputGroups > QDB_AGG_MAX_EMPTY_INPUT_GROUPS) {
        return fail("bad empty-input group count %d", spec.NumEmptyInputGroups);
    }
    if (spec.NumEmptyInputGroups > 0) {
        // gen.cpp:2226-2232: the seeded groups need a leading integer key.
        const int c0 = spec.NumKeys > 0 ? spec.KeyColumns[0] : -1;
        const QdbValueType* t0 = c0 >= 0 && c0 < nin ? &spec.Input.Types[c0] : nullptr;
        if (!t0 || t0->Nullable || t0->Width > 8
            || (t0->Kind != QDB_KIND_SIGNED && t0->Kind != QDB_KIND_UNSIGNED && t0->Kind != QDB_KIND_DATE)) {
            return fail("an aggregate with empty-input groups requires a leading integer key");
        }
    }
    if (spec.NumComputed < 0 || spec.NumComputed > QDB_PROJECT_MAX_OUTPUTS) return fail("bad computed count");
    for (int k = 0; k < spec.NumComputed; ++k) {
        // A computed DECIMAL(p <= 18) argument would be stored by the
        // project operator with a row-error check (qdb_decimal64_narrow_
        // checked); the fused update has no row-error channel, so such an
        // argument runs through the project kernel instead.
        const QdbValueType& t = spec.ComputedTypes[k];
        if (t.Kind == QDB_KIND_DECIMAL && t.Width == 8) {
            return fail("computed column %d is DECIMAL(p <= 18): run it through the project kernel", k);
        }
    }

    // Keys: the null bits of nullable fields first, then every field at its
    // natural alignment (a 16-byte field on a word boundary).
    int nnull = 0;
    for (int k = 0; k < spec.NumKeys; ++k) {
        const int c = spec.KeyColumns[k];
        if (c >= 0 && c < nin && types[c].Nullable) ++nnull;
    }
    int byte = (nnull + 7) / 8;
    int nullbit = 0;
    for (int k = 0; k < spec.NumKeys; ++k) {
        const int c = spec.KeyColumns[k];
        if (bad_col(c) || c >= nin) return fail("key column %d must be an input column", c);
        const QdbValueType& t = types[c];
        if (t.Kind == QDB_KIND_STRING) return fail("string keys are not supported on this device");
        const int w = key_bytes_of(t);
        if (w != 1 && w != 2 && w != 4 && w != 8 && w != 16) return fail("key width %d unsupported", w);
        const int align = w > 8 ? 8 : w;
        byte = (byte + align - 1) / align * align;
        KeyFieldDesc& kf = A.keys[k];
        int s = slot_of(c);
        if (s < 0) return fail("too many columns");
        kf.slot = s;
        kf.kind = t.Kind;
        kf.width = w;
        kf.nullable = t.Nullable;
        kf.word = byte / 8;
        kf.shift = (byte % 8) * 8;
        if (t.Nullable) {
            kf.null_word = nullbit / 64;
            kf.null_bit = nullbit % 64;
            ++nullbit;
        }
        byte += w;
    }
    P->KW = (byte + 7) / 8;
    if (P->KW < 1) P->KW = 1;
    if (P->KW > 4) return fail("a key of %d bytes exceeds the 32 bytes this device packs", byte);
    A.nkeys = spec.NumKeys;

    // Sources and state words.
    int src_col[kMaxSrc];
    A.nsrc = 0;
    auto src_of = [&](int c) -> int {
        for (int j = 0; j < A.nsrc; ++j) {
            if (src_col[j] == c) return j;
        }
        if (A.nsrc >= kMaxSrc) return -1;
        src_col[A.nsrc] = c;
        SrcDesc& sd = A.src[A.nsrc];
        if (c >= nin) {
            sd.slot = -1;
            sd.prog = c - nin;
        } else {
            int s = slot_of(c);
            if (s < 0) return -1;
            sd.slot = s;
            sd.prog = -1;
        }
        return A.nsrc++;
    };
    int nw = 0;
    auto word = [&](uint8_t kind, uint8_t op, int src, int part) -> int {
        if (nw >= kMaxWords) return -1;
        WordDesc& d = A.words[nw];
        d.kind = kind;
        d.op = op;
        d.src = src;
        d.part = part;
        return nw++;
    };
    const int rows_word = word(WK_ROWS, OP_ADD, 0, 0);
    FinArgs& F = P->fin;
    memset(&F, 0, sizeof(F));
    F.nred = spec.NumAggs;
    F.rows_word = rows_word;
    for (int a = 0; a < spec.NumAggs; ++a) {
        const QdbAggregateFn& fn = spec.Aggs[a];
        RedOut& ro = F.red[a];
        ro.wvalid = -1;
        if (fn.Func == QDB_AGG_COUNT_STAR) {
            ro.kind = RK_COUNT_ROWS;
            ro.out_width = 8;
            continue;
        }
        if (bad_col(fn.Column)) return fail("aggregate %d: bad column %d", a, fn.Column);
        const QdbValueType& t = types[fn.Column];
        if (t.Kind == QDB_KIND_STRING) return fail("string aggregates are not supported on this device");
        // An identical earlier aggregate (AVG lowers to SUM + COUNT of the same
        // argument) shares its state words; COUNT of a non-nullable argument is
        // the group's row count.
        bool shared = false;
        for (int b = 0; b < a && !shared; ++b) {
            if (spec.Aggs[b].Func == fn.Func && spec.Aggs[b].Column == fn.Column) {
                ro = F.red[b];
                shared = true;
            }
        }
        if (shared) continue;
        if (fn.Func == QDB_AGG_COUNT && !t.Nullable) {
            ro.kind = RK_COUNT_ROWS;
            ro.out_width = 8;
            continue;
        }
        const int src = src_of(fn.Column);
        if (src < 0) return fail("too many aggregate arguments (max %d)", kMaxSrc);
        if (fn.Func == QDB_AGG_COUNT) {
            ro.kind = RK_COUNT_VALID;
            ro.out_width = 8;
            ro.w0 = word(WK_VALID, OP_ADD, src, 0);
            continue;
        }
        const bool is_max = fn.Func == QDB_AGG_MAX;
        if (fn.Func == QDB_AGG_SUM) {
            if (t.Kind == QDB_KIND_DATE) return fail("function sum(DATE) does not exist");
            if (t.Kind == QDB_KIND_FLOAT) {
                ro.kind = RK_SUMF64;
                ro.out_width = 8;
                ro.w0 = word(WK_LIMB, OP_ADD, src, 0);
                for (int k = 1; k < kLimbs; ++k) word(WK_LIMB, OP_ADD, src, k);
                word(WK_OUTSIDE, OP_FADD, src, 0);
            } else if (t.Kind == QDB_KIND_DECIMAL) {
                ro.kind = RK_SUM128;
                ro.out_width = 16;
                ro.w0 = word(WK_I128_P0, OP_ADD, src, 0);
                word(WK_I128_P1, OP_ADD, src, 0);
                word(WK_I128_P2, OP_ADD, src, 0);
            } else {
                ro.kind = RK_SUM64;
                ro.out_width = 8;
                ro.w0 = word(WK_SUM64, OP_ADD, src, 0);
            }
        } else if (fn.Func == QDB_AGG_MIN || fn.Func == QDB_AGG_MAX) {
            if (t.Kind == QDB_KIND_FLOAT) {
                ro.kind = RK_MINMAX_F64;
                ro.out_width = 8;
                ro.is_max = is_max;
                ro.w0 = word(is_max ? WK_MAX_F64 : WK_MIN_F64, is_max ? OP_MAX_U64 : OP_MIN_U64, src, 0);
            } else if (t.Kind == QDB_KIND_DECIMAL && t.Width == 16) {
                ro.kind = RK_MINMAX_I128;
                ro.out_width = 16;
                ro.is_max = is_max;
                ro.w0 = word(WK_PAIR_LO, is_max ? OP_PAIR_MAX : OP_PAIR_MIN, src, is_max);
                word(WK_PAIR_HI, OP_PAIR_HI, src, is_max);
                word(WK_LOCK, OP_LOCK, src, 0);
                P->has_pairs = true;
                A.has_pairs = 1;
            } else {
                ro.kind = RK_MINMAX_I64;
                ro.out_width = 8;
                ro.is_max = is_max;
                ro.w0 = word(is_max ? WK_MAX_I64 : WK_MIN_I64, is_max ? OP_MAX_I64 : OP_MIN_I64, src, 0);
            }
        } else {
            return fail("aggregate %d: unknown function %d", a, fn.Func);
        }
        if (t.Nullable) ro.wvalid = word(WK_VALID, OP_ADD, src, 0);
        if (ro.w0 < 0 || (t.Nullable && ro.wvalid < 0)) return fail("too many state words (max %d)", kMaxWords);
    }
    A.nwords = nw;
    P->S = nw <= 8 ? 8 : nw <= 16 ? 16 : 32;
    for (int w = 0; w < nw; ++w) P->word_init_host[w] = word_init_host(A.words[w]);

    // Programs.
    if (spec.HasPredicate) {
        QdbExprProgram pred;
        if (!remap(spec.Predicate, &pred)) return false;
        const DevProgram dp = to_dev_program(pred);
        P->has_pred = true;
        P->h_pred = dp;
        if (on_device) {
            P->d_pred = (DevProgram*)upload(&dp, sizeof(dp));
            if (!P->d_pred) return false;
        }
    }
    if (spec.NumComputed > 0) {
        std::vector<DevProgram> progs((size_t)spec.NumComputed);
        for (int k = 0; k < spec.NumComputed; ++k) {
            QdbExprProgram prog;
            if (!remap(spec.Computed[k], &prog)) return false;
            progs[(size_t)k] = to_dev_program(prog);
        }
        P->h_progs = progs;
        if (on_device) {
            P->d_progs = (DevProgram*)upload(progs.data(), progs.size() * sizeof(DevProgram));
            if (!P->d_progs) return false;
        }
    }

    // Finalize layout.
    F.KW = P->KW;
    F.R = P->KW + nw;
    F.nkeys = spec.NumKeys;
    for (int k = 0; k < spec.NumKeys; ++k) {
        F.keys[k] = A.keys[k];
        F.key_out_width[k] = key_bytes_of(types[spec.KeyColumns[k]]);
    }
    output_types(spec, types, P->out_types);
    P->nout = spec.NumKeys + spec.NumAggs;
    for (int a = 0; a < spec.NumAggs; ++a) F.red[a].nullable_out = P->out_types[spec.NumKeys + a].Nullable;
    return true;
}

// Fills the per-batch fields of the launch arguments.
bool bind_batch(const AggPlan* P, const TRowSet* batch, AggArgs* A)
{
    *A = P->proto;
    for (int s = 0; s < P->nslots; ++s) {
        const int c = P->slot_col[s];
        if (c >= batch->ColumnCount) return fail("batch has %lld columns, the spec reads column %d", (long long)batch->ColumnCount, c);
        const TColumn& col = batch->Columns[c];
        const QdbValueType& t = P->spec.Input.Types[c];
        DevCol& d = A->cols.c[s];
        d.data = col.Data;
        d.mask = col.Mask;
        d.mask_off = col.MaskBitOffset;
        d.bit_off = col.DataBitOffset;
        d.kind = t.Kind;
        d.width = t.Width;
        d.nullable = t.Nullable;
    }
    A->nrows = batch->RowCount;
    A->sel = batch->Selection;
    A->pred = P->d_pred;
    A->progs = P->d_progs;
    return true;
}


// Words per record: a power of two up to 16 words (a record is then one
// 128-byte line, aligned), else a multiple of 16.
int record_stride(int R)
{
    if (R > 16) return (R + 15) / 16 * 16;
    int s = 1;
    while (s < R) s *= 2;
    return s;
}

bool fill_records(const AggPlan* P, uint64_t* recs, uint64_t nrecs, int stride)
{
    RecordPattern pat;
    memset(&pat, 0, sizeof(pat));
    const int R = P->KW + P->proto.nwords;
    for (int w = 0; w < 2 * kMaxWords; ++w) {
        const int x = w % stride;
        if (x >= R) pat.w[w] = 0;
        else if (x < P->KW) pat.w[w] = (P->KW == 1 && x == 0) ? kEmptyKey : 0ull;
        else pat.w[w] = P->word_init_host[x - P->KW];
    }
    const uint64_t words = nrecs * (uint64_t)stride;
    const int grid = grid_for((int64_t)(words / 2), kBlock, 8);
    hipLaunchKernelGGL(k_fill_records, dim3(grid), dim3(kBlock), 0, stream(), recs, words, stride, pat);
    return hip_ok(hipGetLastError(), "record fill");
}

// The buffers of a table of `capacity` slots plus the extra record, every
// slot empty. Nothing of T changes; on failure whatever was allocated here is
// freed again.
struct TableBufs {
    unsigned long long* ctrl = nullptr;
    uint32_t* occ = nullptr;
    uint64_t* recs = nullptr;
    int rstride = 1;
    void release()
    {
        dfree(ctrl);
        dfree(occ);
        dfree(recs);
        ctrl = nullptr;
        occ = nullptr;
        recs = nullptr;
    }
};

bool table_bufs_alloc(const AggPlan* P, uint64_t capacity, TableBufs* B)
{
    *B = TableBufs();
    B->rstride = record_stride(P->KW + P->proto.nwords);
    bool ok = true;
    if (P->KW > 1) {
        B->ctrl = (unsigned long long*)dalloc(capacity * 8);
        ok = B->ctrl && zero_fill(B->ctrl, capacity * 8);
    } else {
        const size_t ob = (size_t)((capacity + 31) / 32) * 4;
        B->occ = (uint32_t*)dalloc(ob);
        ok = B->occ && zero_fill(B->occ, ob);
    }
    if (ok) {
        B->recs = (uint64_t*)dalloc((size_t)(capacity + 1) * B->rstride * 8);
        ok = B->recs && fill_records(P, B->recs, capacity + 1, B->rstride);
    }
    if (!ok) B->release();
    return ok;
}

// T gets a new empty table of `capacity` slots (T's buffers must be free).
bool table_alloc(AggTable* T, uint64_t capacity)
{
    TableBufs B;
    if (!table_bufs_alloc(T->plan, capacity, &B)) return false;
    T->capacity = capacity;
    T->rstride = B.rstride;
    T->ctrl = B.ctrl;
    T->occ = B.occ;
    T->recs = B.recs;
    return true;
}

// ---------------------------------------------------------------------------
// The update kernel generated for one aggregate: the state-word layout as
// constants and the row's contribution (selection, predicate, arguments,
// key) as straight-line code over the shared helpers, so that the compiler
// folds every descriptor switch; the tables and their protocol are the
// shared agg_update_body.
// ---------------------------------------------------------------------------

std::string i64_lit(int64_t v)
{
    char b[40];
    snprintf(b, sizeof(b), "(int64_t)0x%llxull", (unsigned long long)v);
    return b;
}

std::string u64_lit(uint64_t v)
{
    char b[40];
    snprintf(b, sizeof(b), "0x%llxull", (unsigned long long)v);
    return b;
}

// Emits the program as one DVal per register and returns the name holding
// its result.
// Straight-line code for a program over columns of the given slot types;
// returns the name of its result register.
std::string emit_program_typed(std::string& o, const std::string& name, const DevProgram& prog, const QdbValueType* slot_type)
{
    for (int q = 0; q < QDB_EXPR_MAX_REGS; ++q) o += "    DVal " + name + "_r" + std::to_string(q) + " = dval_zero();\n";
    auto reg = [&](int q) { return name + "_r" + std::to_string(q >= 0 && q < QDB_EXPR_MAX_REGS ? q : 0); };
    for (int k = 0; k < prog.n && k < QDB_EXPR_MAX_INSNS; ++k) {
        const DevInsn& in = prog.insns[k];
        if (in.dst < 0 || in.dst >= QDB_EXPR_MAX_REGS) continue;
        if (in.op == QDB_EXPR_COLUMN) {
            const QdbValueType& t = slot_type[in.column];
            o += "    " + reg(in.dst) + " = read_col_t<" + std::to_string(t.Kind) + ", " + std::to_string(t.Width) + ">(C.c[" +
                std::to_string(in.column) + "], row);\n";
        } else {
            char b[400];
            snprintf(b, sizeof(b), "    %s = exec_insn<false>(DevInsn{%d, %d, %d, %d, %d, %d, %s, %s}, %s, %s, C, row);\n",
                reg(in.dst).c_str(), in.op, in.dst, in.a, in.b, in.cls, in.column, i64_lit(in.lo).c_str(), i64_lit(in.hi).c_str(),
                reg(in.a).c_str(), reg(in.b).c_str());
            o += b;
        }
    }
    return reg(prog.result);
}

std::string emit_program(std::string& o, const std::string& name, const DevProgram& prog, const AggPlan* P)
{
    return emit_program_typed(o, name, prog, P->slot_type);
}

// Class (QDB_REG_*) of a plan's argument j as a kernel computes it.
int source_class(const AggPlan* P, int j)
{
    const SrcDesc& sd = P->proto.src[j];
    auto col_cls = [](const QdbValueType& t) {
        return t.Kind == QDB_KIND_FLOAT ? QDB_REG_F64 : t.Kind == QDB_KIND_BOOL ? QDB_REG_BOOL : QDB_REG_INT;
    };
    if (sd.prog < 0) return col_cls(P->slot_type[sd.slot]);
    const DevProgram& prog = P->h_progs[(size_t)sd.prog];
    int cls[QDB_EXPR_MAX_REGS] = {0};
    for (int k = 0; k < prog.n && k < QDB_EXPR_MAX_INSNS; ++k) {
        const DevInsn& in = prog.insns[k];
        if (in.dst < 0 || in.dst >= QDB_EXPR_MAX_REGS) continue;
        if (in.op == QDB_EXPR_COLUMN) cls[in.dst] = col_cls(P->slot_type[in.column]);
        else if (in.op == QDB_EXPR_CMP || in.op == QDB_EXPR_AND || in.op == QDB_EXPR_OR || in.op == QDB_EXPR_NOT || in.op == QDB_EXPR_IS_NULL)
            cls[in.dst] = QDB_REG_BOOL;
        else cls[in.dst] = in.cls;
    }
    return cls[prog.result >= 0 && prog.result < QDB_EXPR_MAX_REGS ? prog.result : 0];
}

// Where a row's key and arguments sit in its staged words (the bucketed
// high-cardinality path): key words, then per argument one word (f64 bits,
// or an integer that fits 64 bits) or two (an i128), then a word of
// argument validity bits when an argument can be NULL. words = 0: the plan
// does not stage (i128 MIN/MAX, or more than kMaxStageWords words).
StageLayout stage_layout(const AggPlan* P)
{
    StageLayout L;
    const AggArgs& A = P->proto;
    int w = P->KW;
    bool nullable = false;
    for (int j = 0; j < A.nsrc; ++j) {
        const SrcDesc& sd = A.src[j];
        L.cls[j] = source_class(P, j);
        L.off[j] = w;
        bool two = false;
        if (sd.prog >= 0) {
            two = L.cls[j] == QDB_REG_INT;
            nullable = true;
        } else {
            two = P->slot_type[sd.slot].Width == 16;
            nullable |= P->slot_type[sd.slot].Nullable != 0;
        }
        L.two[j] = two;
        w += two ? 2 : 1;
    }
    L.valid_word = nullable ? w++ : -1;
    // The sort moves 2..6 or 8 words: pad to the next such count.
    if (w < 2) w = 2;
    if (w == 7) w = 8;
    L.words = (P->has_pairs || w > kMaxStageWords) ? 0 : w;
    return L;
}

// `const DVal v = <argument j of the row>;`
void emit_source(std::string& o, const AggPlan* P, int j, const std::string& v)
{
    const SrcDesc& sd = P->proto.src[j];
    if (sd.prog >= 0) {
        const std::string r = emit_program(o, v + "p", P->h_progs[(size_t)sd.prog], P);
        o += "        const DVal " + v + " = " + r + ";\n";
    } else {
        const QdbValueType& t = P->slot_type[sd.slot];
        o += "        const DVal " + v + " = read_col_t<" + std::to_string(t.Kind) + ", " + std::to_string(t.Width) + ">(C.c[" +
            std::to_string(sd.slot) + "], row);\n";
    }
}

// The state words argument j contributes to, from its value v.
void emit_contribs(std::string& o, const AggPlan* P, int j, const std::string& v)
{
    const AggArgs& A = P->proto;
    o += "        if (" + v + ".valid) {\n";
    o += "            int64_t limb[kLimbs] = {};\n            bool outside = false;\n";
    bool limbs = false;
    for (int w = 0; w < A.nwords; ++w) limbs |= A.words[w].src == j && (A.words[w].kind == WK_LIMB || A.words[w].kind == WK_OUTSIDE);
    if (limbs) o += "            if (" + v + ".cls == QDB_REG_F64) f64_limbs(" + v + ".f, limb, &outside);\n";
    for (int w = 0; w < A.nwords; ++w) {
        const WordDesc& d = A.words[w];
        if (d.src != j || d.kind == WK_ROWS || d.kind == WK_LOCK) continue;
        char b[200];
        snprintf(b, sizeof(b), "            c[%d] = word_contrib(WordDesc{%d, %d, %d, %d}, %s, limb, outside);\n", w, d.kind, d.op, d.src, d.part,
            v.c_str());
        o += b;
    }
    o += "        }\n";
}

// keep_: whether the row takes part (its selection byte and the predicate).
// The generated functions decide at their end, not by an early return, so
// that the loads of every column the row needs issue together instead of
// waiting for the predicate's columns first; the code after it only loads
// in-bounds values of the same row and computes, without side effects.
void emit_keep(std::string& o, const AggPlan* P)
{
    o += "        bool keep_ = !(A.sel && !A.sel[row]);\n";
    if (P->has_pred) {
        o += "        {\n";
        const std::string r = emit_program(o, "p", P->h_pred, P);
        o += "    keep_ = keep_ && " + r + ".valid && " + r + ".i != 0;\n        }\n";
    }
}

// stage(): the row's staged words (false: the row does not take part);
// unstage(): the key and contributions back from them.
void emit_stage(std::string& o, const AggPlan* P)
{
    const StageLayout L = stage_layout(P);
    const AggArgs& A = P->proto;
    const int nw = A.nwords;
    o += "    static constexpr int kStage = " + std::to_string(L.words > 0 ? L.words : 1) + ";\n";
    if (L.words == 0) return;
    o += "    template <int KW>\n";
    o += "    __device__ static __forceinline__ bool stage(const AggArgs& A, int64_t row, uint64_t (&st)[kStage])\n    {\n";
    o += "        const DevCols& C = A.cols;\n";
    o += "#pragma unroll\n        for (int w = 0; w < kStage; ++w) st[w] = 0;\n";
    emit_keep(o, P);
    o += "        uint64_t key[KW];\n#pragma unroll\n        for (int w = 0; w < KW; ++w) key[w] = 0;\n";
    for (int f = 0; f < A.nkeys; ++f) {
        const KeyFieldDesc& k = A.keys[f];
        char b[300];
        snprintf(b, sizeof(b), "        pack_key_field<KW>(KeyFieldDesc{%d, %d, %d, %d, %d, %d, %d, %d}, C.c[%d], row, key);\n", k.slot, k.kind,
            k.width, k.nullable, k.word, k.shift, k.null_word, k.null_bit, k.slot);
        o += b;
    }
    o += "#pragma unroll\n        for (int w = 0; w < KW; ++w) st[w] = key[w];\n";
    o += "        uint64_t vb = 0;\n";
    for (int j = 0; j < A.nsrc; ++j) {
        const std::string v = "s" + std::to_string(j), off = std::to_string(L.off[j]);
        o += "        {\n";
        emit_source(o, P, j, v);
        o += "        if (" + v + ".valid) vb |= 1ull << " + std::to_string(j) + ";\n";
        if (L.cls[j] == QDB_REG_F64) {
            o += "        st[" + off + "] = (uint64_t)__double_as_longlong(" + v + ".f);\n";
        } else if (L.two[j]) {
            o += "        st[" + off + "] = (uint64_t)(u128)" + v + ".i;\n";
            o += "        st[" + std::to_string(L.off[j] + 1) + "] = (uint64_t)((u128)" + v + ".i >> 64);\n";
        } else {
            o += "        st[" + off + "] = (uint64_t)(int64_t)" + v + ".i;\n";
        }
        o += "        }\n";
    }
    if (L.valid_word >= 0) o += "        st[" + std::to_string(L.valid_word) + "] = vb;\n";
    o += "        (void)vb;\n        return keep_;\n    }\n";
    // stage_key(): the row's key alone (the partition's counting pass reads
    // only the key and predicate columns).
    o += "    template <int KW>\n";
    o += "    __device__ static __forceinline__ bool stage_key(const AggArgs& A, int64_t row, uint64_t (&key)[KW])\n    {\n";
    o += "        const DevCols& C = A.cols;\n";
    o += "#pragma unroll\n        for (int w = 0; w < KW; ++w) key[w] = 0;\n";
    emit_keep(o, P);
    for (int f = 0; f < A.nkeys; ++f) {
        const KeyFieldDesc& k = A.keys[f];
        char b[300];
        snprintf(b, sizeof(b), "        pack_key_field<KW>(KeyFieldDesc{%d, %d, %d, %d, %d, %d, %d, %d}, C.c[%d], row, key);\n", k.slot, k.kind,
            k.width, k.nullable, k.word, k.shift, k.null_word, k.null_bit, k.slot);
        o += b;
    }
    o += "        return keep_;\n    }\n";
    o += "    template <int KW, int S>\n";
    o += "    __device__ static __forceinline__ void unstage(const AggArgs& A, const uint64_t (&st)[kStage], uint64_t (&key)[KW], uint64_t (&c)[S])\n    {\n";
    o += "#pragma unroll\n        for (int w = 0; w < KW; ++w) key[w] = st[w];\n";
    o += "#pragma unroll\n        for (int w = 0; w < S; ++w) c[w] = 0;\n";
    for (int w = 0; w < nw; ++w) {
        const WordDesc& d = A.words[w];
        o += "        c[" + std::to_string(w) + "] = " + (d.kind == WK_ROWS ? std::string("1ull") : u64_lit(P->word_init_host[w])) + ";\n";
    }
    for (int j = 0; j < A.nsrc; ++j) {
        const std::string v = "s" + std::to_string(j), off = std::to_string(L.off[j]);
        o += "        {\n        DVal " + v + " = dval_zero();\n";
        o += "        " + v + ".cls = " + std::to_string(L.cls[j]) + ";\n";
        o += "        " + v + ".valid = " +
            (L.valid_word >= 0 ? "(uint8_t)((st[" + std::to_string(L.valid_word) + "] >> " + std::to_string(j) + ") & 1ull)" : std::string("1")) +
            ";\n";
        if (L.cls[j] == QDB_REG_F64) {
            o += "        " + v + ".f = __longlong_as_double((long long)st[" + off + "]);\n";
        } else if (L.two[j]) {
            o += "        " + v + ".i = (i128)(((u128)st[" + std::to_string(L.off[j] + 1) + "] << 64) | (u128)st[" + off + "]);\n";
        } else {
            o += "        " + v + ".i = (i128)(int64_t)st[" + off + "];\n";
        }
        emit_contribs(o, P, j, v);
        o += "        }\n";
    }
    o += "    }\n";
}

std::string agg_jit_source(const AggPlan* P)
{
    const AggArgs& A = P->proto;
    const int nw = A.nwords;
    std::string o;
    o += "#define QDB_SPECIALIZED 1\n#include \"aggregate_kernels.h\"\n";
    o += "namespace qdbgpu {\nnamespace agg {\n";
    o += "__device__ __forceinline__ DVal dval_zero()\n{\n    DVal v;\n    v.i = 0;\n    v.f = 0.0;\n    v.cls = 0;\n    v.valid = 0;\n    return v;\n}\n";
    o += "struct JitPolicy {\n";
    o += "    static constexpr int kNW = " + std::to_string(nw) + ";\n";
    o += "    __device__ static __forceinline__ int nwords(const AggArgs&) { return kNW; }\n";
    o += "    __device__ static __forceinline__ uint8_t op(const AggArgs&, int w)\n    {\n        constexpr uint8_t ops[] = {";
    for (int w = 0; w < nw; ++w) o += std::to_string(A.words[w].op) + ", ";
    o += "0};\n        return w < kNW ? ops[w] : 0;\n    }\n";
    o += "    __device__ static __forceinline__ uint64_t init(const AggArgs&, int w)\n    {\n        constexpr uint64_t v[] = {";
    for (int w = 0; w < nw; ++w) o += u64_lit(P->word_init_host[w]) + ", ";
    o += "0};\n        return w < kNW ? v[w] : 0;\n    }\n";
    o += std::string("    __device__ static __forceinline__ bool has_pairs(const AggArgs&) { return ") + (P->has_pairs ? "true" : "false") + "; }\n";
    o += "    __device__ static __forceinline__ int match_rounds(const AggArgs& A) { return A.match_rounds; }\n";
    o += "    template <int KW, int S>\n";
    o += "    __device__ static __forceinline__ bool contrib(const AggArgs& A, int64_t row, uint64_t (&key)[KW], uint64_t (&c)[S])\n    {\n";
    o += "        const DevCols& C = A.cols;\n";
    emit_keep(o, P);
    o += "#pragma unroll\n        for (int w = 0; w < S; ++w) c[w] = 0;\n";
    for (int w = 0; w < nw; ++w) {
        const WordDesc& d = A.words[w];
        o += "        c[" + std::to_string(w) + "] = " + (d.kind == WK_ROWS ? std::string("1ull") : u64_lit(P->word_init_host[w])) + ";\n";
    }
    for (int j = 0; j < A.nsrc; ++j) {
        const std::string v = "s" + std::to_string(j);
        o += "        {\n";
        emit_source(o, P, j, v);
        emit_contribs(o, P, j, v);
        o += "        }\n";
    }
    o += "#pragma unroll\n        for (int w = 0; w < KW; ++w) key[w] = 0;\n";
    for (int f = 0; f < A.nkeys; ++f) {
        const KeyFieldDesc& k = A.keys[f];
        char b[300];
        snprintf(b, sizeof(b), "        pack_key_field<KW>(KeyFieldDesc{%d, %d, %d, %d, %d, %d, %d, %d}, C.c[%d], row, key);\n", k.slot, k.kind,
            k.width, k.nullable, k.word, k.shift, k.null_word, k.null_bit, k.slot);
        o += b;
    }
    o += "        return keep_;\n    }\n";
    emit_stage(o, P);
    o += "};\n}  // namespace agg\n}  // namespace qdbgpu\n";
    o += "extern \"C\" __global__ void __launch_bounds__(" + std::to_string(kBlock) + ") qdb_agg_update(qdbgpu::agg::AggArgs A)\n{\n";
    o += "    qdbgpu::agg::agg_update_body<" + std::to_string(P->KW) + ", " + std::to_string(P->S) + ", qdbgpu::agg::JitPolicy>(A);\n}\n";
    o += "extern \"C\" __global__ void __launch_bounds__(" + std::to_string(kBlock) +
        ") qdb_agg_batch(const qdbgpu::agg::AggArgs* insts, const qdbgpu::agg::BatchItem* items, int nitems, unsigned int* head)\n{\n";
    o += "    qdbgpu::agg::agg_batch_body<" + std::to_string(P->KW) + ", " + std::to_string(P->S) +
        ", qdbgpu::agg::JitPolicy>(insts, items, nitems, head);\n}\n";
    o += "extern \"C\" __global__ void __launch_bounds__(" + std::to_string(kBlock) +
        ") qdb_agg_reduce(qdbgpu::agg::AggArgs A, const unsigned long long* parts, const unsigned int* part_n, int nparts, int per)\n{\n";
    o += "    qdbgpu::agg::agg_reduce_parts_body<" + std::to_string(P->KW) + ", " + std::to_string(P->S) +
        ", qdbgpu::agg::JitPolicy>(A, (const uint64_t*)parts, part_n, nparts, per);\n}\n";
    const StageLayout L = stage_layout(P);
    if (L.words > 0) {
        const std::string kw = std::to_string(P->KW), sw = std::to_string(P->S), w = std::to_string(L.words);
        o += "extern \"C\" __global__ void __launch_bounds__(" + std::to_string(kSortedBlock) +
            ") qdb_agg_count(qdbgpu::agg::AggArgs A, unsigned int* hist, int shift, unsigned int nb, long long chunk, "
            "const unsigned long long* hot_keys, const unsigned int* hot_n)\n{\n";
        o += "    qdbgpu::agg::agg_count_body<" + kw + ", " + w + ", qdbgpu::agg::JitPolicy>(A, hist, shift, nb, chunk, hot_keys, hot_n);\n}\n";
        o += "extern \"C\" __global__ void __launch_bounds__(" + std::to_string(kSortedBlock) +
            ") qdb_agg_scatter(qdbgpu::agg::AggArgs A, const unsigned int* offs, unsigned long long* staged, int shift, unsigned int nb, "
            "long long chunk, const unsigned long long* hot_keys, const unsigned int* hot_n, unsigned long long* hot_parts)\n{\n";
        o += "    qdbgpu::agg::agg_scatter_body<" + kw + ", " + sw + ", " + w +
            ", qdbgpu::agg::JitPolicy>(A, offs, staged, shift, nb, chunk, hot_keys, hot_n, hot_parts);\n}\n";
        if (P->KW == 1) {
            o += "extern \"C\" __global__ void __launch_bounds__(" + std::to_string(agg::kHotSampleBlock) +
                ") qdb_agg_hot_sample(qdbgpu::agg::AggArgs A, long long n, unsigned long long* hot_keys, unsigned int* hot_n)\n{\n";
            o += "    qdbgpu::agg::agg_hot_sample_body<1, qdbgpu::agg::JitPolicy>(A, n, hot_keys, hot_n);\n}\n";
            o += "extern \"C\" __global__ void __launch_bounds__(" + std::to_string(kBlock) +
                ") qdb_agg_hot_reduce(qdbgpu::agg::AggArgs A, const unsigned long long* parts, int nblk, const unsigned int* hot_n, "
                "unsigned long long* out)\n{\n";
            o += "    qdbgpu::agg::agg_hot_reduce_body<1, qdbgpu::agg::JitPolicy>(A, parts, nblk, hot_n, out);\n}\n";
        }
        o += "extern \"C\" __global__ void __launch_bounds__(" + std::to_string(kSortedBlock) +
            ") qdb_agg_sorted(qdbgpu::agg::AggArgs A, const unsigned long long* staged, const qdbgpu::agg::SortedItem* items, int nitems, "
            "unsigned int* head, unsigned long long* recs, unsigned int* counts)\n{\n";
        o += "    qdbgpu::agg::agg_sorted_body<" + kw + ", " + sw + ", " + w +
            ", qdbgpu::agg::JitPolicy>(A, staged, items, nitems, head, recs, counts);\n}\n";
        o += "extern \"C\" __global__ void __launch_bounds__(" + std::to_string(kBlock) +
            ") qdb_agg_merge(qdbgpu::agg::AggArgs A, const unsigned long long* recs, const qdbgpu::agg::SortedItem* items, "
            "const unsigned int* counts, int nitems, unsigned int* head)\n{\n";
        o += "    qdbgpu::agg::agg_merge_body<" + kw + ", " + sw + ", qdbgpu::agg::JitPolicy>(A, recs, items, counts, nitems, head);\n}\n";
    }
    return o;
}

// A kernel of the plan's generated source ("qdb_agg_update" or
// "qdb_agg_batch"), compiled on first use; nullptr when run-time kernels are
// off or the compile failed (the generic kernels run then).
hipFunction_t plan_jit_function(AggPlan* P, const char* name)
{
    if (!jit_enabled()) return nullptr;
    {
        std::lock_guard<std::mutex> lock(P->jit_mu);
        if (!P->jit_tried) {
            P->jit_tried = true;
            P->jit_src = agg_jit_source(P);
        }
    }
    std::string err;
    return jit_function(P->jit_src, name, &err);
}

AggTable* table_new(AggPlan* P, int64_t slots_hint)
{
    AggTable* T = new AggTable();
    T->plan = P;
    uint64_t cap = 1024;
    while (cap < (uint64_t)slots_hint) cap *= 2;
    const int64_t groups_hint = slots_hint / 2;
    T->counters = (unsigned int*)dalloc(16);
    T->cap_hint = cap;
    T->init_slots = slots_hint;
    // Tables large enough for the bucketed path get their buffers at first
    // need (see AggTable).
    const bool lazy = cap >= kSortedMinSlots;
    if (lazy) {
        T->capacity = cap;
        T->rstride = record_stride(P->KW + P->proto.nwords);
        T->allocated = false;
    }
    if (!T->counters || !zero_fill(T->counters, 16) || (!lazy && !table_alloc(T, cap))) {
        dfree(T->counters);
        dfree(T->ctrl);
        dfree(T->recs);
        delete T;
        return nullptr;
    }
    // Few expected groups: wave-level matching pays; many: it does not.
    // Wave-level matching of equal keys before the table updates: measured
    // slower than letting every lane apply its own row for both few groups
    // (the per-round reduction of every state word costs more than the LDS
    // atomics it saves) and many (most keys differ within a wave), so it is
    // off; the generic kernels keep the machinery.
    T->match_rounds = 0;
    (void)groups_hint;
    return T;
}

void table_free(AggTable* T)
{
    if (!T) return;
    dfree(T->ctrl);
    dfree(T->occ);
    dfree(T->recs);
    dfree(T->counters);
    table_drop_pending(T);
    delete T;
}

// Doubles the HBM table until `need` groups fill at most half of it; moves
// the groups over.
bool table_grow(AggTable* T, uint64_t need)
{
    const AggPlan* P = T->plan;
    uint64_t cap = T->capacity * 2;
    while (cap / 2 < need) cap *= 2;
    // The new buffers are made first; the table keeps its own until every
    // allocation has succeeded.
    TableBufs B;
    if (!table_bufs_alloc(P, cap, &B)) return false;
    unsigned long long* old_ctrl = T->ctrl;
    uint32_t* old_occ = T->occ;
    uint64_t* old_recs = T->recs;
    const uint64_t old_cap = T->capacity;
    T->capacity = cap;
    T->rstride = B.rstride;
    T->ctrl = B.ctrl;
    T->occ = B.occ;
    T->recs = B.recs;
    AggArgs A = P->proto;
    A.ctrl = T->ctrl;
    A.occ = T->occ;
    A.cmask = T->capacity - 1;
    A.recs = T->recs;
    A.rstride = T->rstride;
    agg::launch_rehash(P->KW, A, old_recs, old_ctrl, old_cap);
    if (!hip_ok(hipGetLastError(), "rehash launch")) return false;
    // The extra record keeps its contents at its new place.
    const size_t rb = (size_t)T->rstride * 8;
    if (!hip_ok(hipMemcpyAsync(T->recs + cap * (uint64_t)T->rstride, old_recs + old_cap * (uint64_t)T->rstride, rb,
            hipMemcpyDeviceToDevice, stream()), "grow copy")) {
        return false;
    }
    if (!sync("grow")) return false;
    dfree(old_ctrl);
    dfree(old_occ);
    dfree(old_recs);
    return true;
}

void table_args(const AggTable* T, AggArgs* A)
{
    A->ctrl = T->ctrl;
    A->occ = T->occ;
    A->cmask = T->capacity - 1;
    A->recs = T->recs;
    A->rstride = T->rstride;
    A->counters = T->counters;
    A->match_rounds = T->match_rounds;
}

bool read_counters(AggTable* T, unsigned int out[4])
{
    if (!hip_ok(hipMemcpyAsync(out, T->counters, 16, hipMemcpyDeviceToHost, stream()), "counters")) return false;
    if (!sync("aggregate update")) return false;
    T->groups = out[0];
    T->extra_used = out[3];
    return true;
}

__global__ void k_set_counters(unsigned int* c, unsigned int groups, unsigned int extra_used)
{
    if (threadIdx.x == 0) {
        c[0] = groups;
        c[1] = 0u;
        c[2] = 0u;
        c[3] = extra_used;
    }
}

// The groups an empty input has (QdbAggregateSpec.EmptyInputGroups).
struct SeedGroups {
    int n;
    int64_t id[QDB_AGG_MAX_EMPTY_INPUT_GROUPS];
};

// Inserts the seeded groups into a fresh HBM table, as the reference's Init
// upserts its seed keys (qdb/kernel/gen.cpp:2233-2266): the leading key field
// holds the id (the zero-extension of its width, as pack_key_field packs an
// integer key), every other field is NULL when nullable and 0 otherwise; the
// records keep their identities. One thread does it all: a handful of keys,
// no contention, no wait on another lane.
template <int KW>
__global__ void k_seed_groups(AggArgs A, SeedGroups G)
{
    __shared__ unsigned int s_inserted;
    if (threadIdx.x != 0) return;
    s_inserted = 0u;
    agg::LdsTable T;
    T.ctrl = nullptr;
    T.count = nullptr;
    T.recs = nullptr;
    T.inserted = &s_inserted;
    T.fresh = 0;
    for (int i = 0; i < G.n; ++i) {
        uint64_t key[KW];
        for (int w = 0; w < KW; ++w) key[w] = 0;
        for (int f = 0; f < A.nkeys; ++f) {
            const agg::KeyFieldDesc kf = A.keys[f];
            if (f == 0) {
                const uint64_t id = (uint64_t)G.id[i];
                const uint64_t lo = kf.width >= 8 ? id : (id & ((1ull << (8 * kf.width)) - 1));
                for (int w = 0; w < KW; ++w) {
                    if (w == kf.word) key[w] |= lo << kf.shift;
                }
            } else if (kf.nullable) {
                for (int w = 0; w < KW; ++w) {
                    if (w == kf.null_word) key[w] |= 1ull << kf.null_bit;
                }
            }
        }
        if (agg::global_find_or_insert<KW, 1>(A, T, key) < 0) atomicAdd(&A.counters[1], 1u);
    }
    atomicAdd(&A.counters[0], s_inserted);
}

// Seeds the empty-input groups of a new table; T->groups counts them.
bool table_ready(AggTable* T);

bool table_seed(AggTable* T)
{
    const AggPlan* P = T->plan;
    SeedGroups G;
    memset(&G, 0, sizeof(G));
    G.n = P->spec.NumEmptyInputGroups;
    for (int i = 0; i < G.n; ++i) G.id[i] = P->spec.EmptyInputGroups[i];
    if (G.n == 0) return true;
    if (!table_ready(T)) return false;
    AggArgs A = P->proto;
    table_args(T, &A);
    switch (P->KW) {
    case 1: hipLaunchKernelGGL(k_seed_groups<1>, dim3(1), dim3(64), 0, stream(), A, G); break;
    case 2: hipLaunchKernelGGL(k_seed_groups<2>, dim3(1), dim3(64), 0, stream(), A, G); break;
    case 3: hipLaunchKernelGGL(k_seed_groups<3>, dim3(1), dim3(64), 0, stream(), A, G); break;
    default: hipLaunchKernelGGL(k_seed_groups<4>, dim3(1), dim3(64), 0, stream(), A, G); break;
    }
    if (!hip_ok(hipGetLastError(), "seed launch")) return false;
    unsigned int cnt[4];
    if (!read_counters(T, cnt)) return false;
    if (cnt[1] != 0u) return fail("seeding the empty-input groups found no room");
    return true;
}

// The table's counters before a launch (a kernel, not an upload from
// pageable memory, which would wait for the copy).
bool set_counters(AggTable* T)
{
    hipLaunchKernelGGL(k_set_counters, dim3(1), dim3(64), 0, stream(), T->counters, T->groups, T->extra_used);
    return hip_ok(hipGetLastError(), "counters");
}

// LDS table of an update launch: as many groups as fit in lds_kib KiB, a
// power of two (none for i128 MIN/MAX, which live in HBM only).
void lds_layout(const AggPlan* P, AggArgs* A, size_t* lds_bytes, int lds_kib = 64, int copies = 1)
{
    const int R = P->KW + A->nwords;
    const size_t budget = (size_t)lds_kib * 1024;
    if (copies < 1 || P->has_pairs) copies = 1;
    // Records, hash slots (two per group), the counter, then the copies of
    // the state words: an odd number of words per group and per copy, so
    // that a word's copies, and one word of different groups, fall in
    // different LDS banks.
    const int G = A->nwords | 1;  // words per group in a copy (odd)
    auto bytes_of = [&](int g) {
        return (size_t)g * R * 8 + (size_t)g * 2 * 4 + 16 + (size_t)(copies - 1) * ((size_t)g * G + 1) * 8;
    };
    int lds_groups = 4096;
    while (lds_groups > 16 && bytes_of(lds_groups) > budget) lds_groups /= 2;
    if (P->has_pairs) lds_groups = 0;
    A->lds_groups = lds_groups;
    A->lds_slots = lds_groups * 2;
    A->lds_copies = copies;
    A->lds_copy_stride = lds_groups * G + 1;
    *lds_bytes = lds_groups ? bytes_of(lds_groups) : 0;
}

// One update launch over rows [base, base + n) of a batch into table T: its
// arguments and the buffers for what it cannot insert (every row of the
// launch, and `recs_cap` group records). The second pair of buffers, for a
// replay that spills again, is made only when a replay runs.
struct UpdateLaunch {
    AggTable* T = nullptr;
    AggArgs A;
    int64_t n = 0;
    size_t recs_cap = 0;
    uint32_t* ov_rows[2] = {nullptr, nullptr};
    uint64_t* ov_recs[2] = {nullptr, nullptr};

    void release()
    {
        for (int k = 0; k < 2; ++k) {
            dfree(ov_rows[k]);
            dfree(ov_recs[k]);
            ov_rows[k] = nullptr;
            ov_recs[k] = nullptr;
        }
    }
};

bool prepare_launch(AggTable* T, const AggArgs& bound, int64_t base, int64_t n, size_t recs_cap, UpdateLaunch* L)
{
    if (!table_ready(T)) return false;
    const AggPlan* P = T->plan;
    const int R = P->KW + bound.nwords;
    L->T = T;
    L->A = bound;
    L->n = n;
    L->recs_cap = recs_cap;
    L->ov_rows[0] = (uint32_t*)dalloc((size_t)n * 4 + 16);
    L->ov_recs[0] = (uint64_t*)dalloc(recs_cap * R * 8);
    if (!L->ov_rows[0] || !L->ov_recs[0]) {
        L->release();
        return false;
    }
    table_args(T, &L->A);
    L->A.row_base = base;
    L->A.nrows = n;
    L->A.row_index = nullptr;
    L->A.ov_rows = L->ov_rows[0];
    L->A.ov_recs = L->ov_recs[0];
    L->A.ov_rec_cap = (unsigned int)recs_cap;
    return true;
}

// After the launch: while rows or records were left over, grows the table
// and replays exactly those (generic kernels), so every row is applied
// once; then keeps the table at most half full for the next batch.
bool settle_launch(UpdateLaunch* L, unsigned int cnt[4])
{
    AggTable* T = L->T;
    const AggPlan* P = T->plan;
    const int KW = P->KW, S = P->S;
    const int R = KW + L->A.nwords;
    AggArgs A = L->A;
    int cur = 0;
    while (cnt[1] || cnt[2]) {
        if (cnt[2] > L->recs_cap || cnt[1] > (uint64_t)L->n) {
            return fail("aggregate spill overflow (%u rows, %u records)", cnt[1], cnt[2]);
        }
        const unsigned int spilled_rows = cnt[1];
        const unsigned int spilled_recs = cnt[2];
        const int nxt = cur ^ 1;
        if (!L->ov_rows[nxt]) {
            L->ov_rows[nxt] = (uint32_t*)dalloc((size_t)L->n * 4 + 16);
            L->ov_recs[nxt] = (uint64_t*)dalloc(L->recs_cap * R * 8);
            if (!L->ov_rows[nxt] || !L->ov_recs[nxt]) return false;
        }
        if (!table_grow(T, (uint64_t)T->groups + spilled_rows + spilled_recs + 1024)) return false;
        table_args(T, &A);
        A.ov_rows = L->ov_rows[nxt];
        A.ov_recs = L->ov_recs[nxt];
        if (!set_counters(T)) return false;
        if (spilled_rows) {
            AggArgs B = A;
            B.row_index = L->ov_rows[cur];
            B.nrows = spilled_rows;
            B.row_base = L->A.row_base;
            B.lds_slots = 0;
            B.lds_groups = 0;
            const int g2 = grid_for(spilled_rows, kBlock, 2);
            agg::launch_update(KW, S, B, g2, 0);
        }
        if (spilled_recs) agg::launch_records(KW, S, A, L->ov_recs[cur], spilled_recs);
        if (!hip_ok(hipGetLastError(), "agg replay launch")) return false;
        if (!read_counters(T, cnt)) return false;
        cur = nxt;
    }
    // Keep the load at most one half for the next batch.
    if ((uint64_t)T->groups > T->capacity / 2 && !table_grow(T, (uint64_t)T->groups)) return false;
    return true;
}

// Merges the pending runs into the slot table: each of their keys occurs
// once and not in the table, so the merge stores the words of a key it
// inserts as they are.
bool table_merge_pending(AggTable* T)
{
    if (T->pend_runs == 0) return true;
    AggPlan* P = T->plan;
    hipFunction_t f_merge = plan_jit_function(P, "qdb_agg_merge");
    if (!f_merge) return fail("bucketed aggregate kernels unavailable");
    // Taken off the table first: the launch below sees a table without them.
    uint64_t* recs = T->pend_recs;
    SortedItem* items = T->pend_items;
    uint32_t* counts = T->pend_counts;
    uint32_t* prefix = T->pend_prefix;
    int nitems = T->pend_runs;
    const unsigned int total = T->pend_groups;
    T->pend_recs = nullptr;
    T->pend_items = nullptr;
    T->pend_counts = nullptr;
    T->pend_prefix = nullptr;
    T->pend_runs = 0;
    T->pend_groups = 0;
    unsigned int* head = (unsigned int*)dalloc(16);
    bool ok = head != nullptr;
    if (ok && (uint64_t)T->groups + total > T->capacity / 2) ok = table_grow(T, (uint64_t)T->groups + total);
    UpdateLaunch L;
    ok = ok && prepare_launch(T, P->proto, 0, 1, (size_t)total + 1024, &L) && set_counters(T) && zero_fill(head, 16);
    if (ok) {
        const int grid = nitems < num_cus() * 4 ? nitems : num_cus() * 4;
        void* args[] = {&L.A, &recs, &items, &counts, &nitems, &head};
        ok = hip_ok(hipModuleLaunchKernel(f_merge, (unsigned)grid, 1, 1, kBlock, 1, 1, 0, stream(), args, nullptr), "merge launch");
    }
    unsigned int cnt[4];
    ok = ok && read_counters(T, cnt) && settle_launch(&L, cnt);
    L.release();
    dfree(head);
    dfree(recs);
    dfree(items);
    dfree(counts);
    dfree(prefix);
    return ok;
}

// The slot table made (at its capacity) and every pending group in it.
bool table_ready(AggTable* T)
{
    if (!T->allocated) {
        if (!table_alloc(T, T->capacity)) return false;
        T->allocated = true;
    }
    return table_merge_pending(T);
}


bool sorted_path_wanted(AggPlan* P, const AggTable* T, int64_t rows)
{
    // QDB_GPU_BUCKETED_MIN_ROWS moves the row threshold (0: never).
    int64_t min_rows = kSortedMinRows;
    if (const char* e = getenv("QDB_GPU_BUCKETED_MIN_ROWS")) {
        min_rows = atoll(e);
        if (min_rows <= 0) return false;
    }
    const uint64_t slots = T->capacity > T->cap_hint ? T->capacity : T->cap_hint;
    return rows >= min_rows && slots >= kSortedMinSlots && stage_layout(P).words > 0 &&
        plan_jit_function(P, "qdb_agg_sorted") != nullptr;
}

// The bucketed update of rows [0, n) of a bound batch (n <= 2^30): stage,
// sort by bucket, aggregate every bucket in LDS into records, grow the
// table for them, merge them into the HBM table.
bool table_update_sorted(AggTable* T, const AggArgs& bound, int64_t n, int fanout_shift = 0, int part_blocks = 0)
{
    AggPlan* P = T->plan;
    // An empty table whose buffers are not made yet keeps the groups of the
    // buckets that one workgroup held whole as pending runs; otherwise the
    // slot table takes every group.
    const bool lazy = !T->allocated && T->groups == 0 && T->pend_runs == 0;
    if (!lazy && !table_ready(T)) return false;
    const StageLayout SL = stage_layout(P);
    const int W = SL.words, KW = P->KW, R = KW + bound.nwords;
    // LDS table of the bucket workgroups: the most groups that fit next to
    // the kernel's own LDS with the hash slots (a power of two) at most 45%
    // full, and buckets sized to fill about 60% of it (fewer buckets: longer
    // runs in the partition's writes).
    AggArgs A = bound;
    table_args(T, &A);
    A.row_base = 0;
    A.nrows = n;
    A.row_index = nullptr;
    const size_t budget = max_lds_per_block() - 4096;
    int g = 0, slots = 0;
    for (int s = 16384; s >= 128; s /= 2) {
        if ((size_t)s * 4 + 16 >= budget) continue;
        const size_t fit = (budget - (size_t)s * 4 - 16) / ((size_t)R * 8);
        const int cap = (int)((size_t)s * 45 / 100 < fit ? (size_t)s * 45 / 100 : fit);
        if (cap > g) {
            g = cap;
            slots = s;
        }
    }
    A.lds_groups = g;
    A.lds_slots = slots;
    A.lds_copies = 1;
    const size_t lds = (size_t)g * R * 8 + (size_t)A.lds_slots * 4 + 16;
    const uint64_t expect = (uint64_t)n < T->capacity / 2 ? (uint64_t)n : T->capacity / 2;
    int bits = 8;
    while (bits < 15 && ((uint64_t)1 << bits) * (uint64_t)(g * 6 / 10) < expect) ++bits;
    bits += fanout_shift;
    bits = bits < 6 ? 6 : bits > 15 ? 15 : bits;
    const uint32_t B = 1u << bits;
    hipFunction_t f_sorted = plan_jit_function(P, "qdb_agg_sorted");
    hipFunction_t f_merge = plan_jit_function(P, "qdb_agg_merge");
    if (!f_sorted || !f_merge) return fail("bucketed aggregate kernels unavailable");

    // Partition: per-workgroup bucket counts over contiguous row ranges, a
    // scan, then every row's staged words at its bucket run's next slot.
    const unsigned int nb = B + 1;  // the last bucket takes rows that do not take part
    // Partition workgroups: one per 16K rows, at most 2048 (or the tuned count).
    const int64_t max_blk = part_blocks > 0 ? part_blocks : 2048;
    int64_t nblk = (n + 16383) / 16384;
    if (nblk > max_blk) nblk = max_blk;
    long long chunk = (long long)((n + nblk - 1) / nblk);
    nblk = (n + chunk - 1) / chunk;
    const size_t nhist = (size_t)nb * (size_t)nblk;
    unsigned int* hist = (unsigned int*)dalloc(nhist * 4 + 16);
    unsigned int* offs = (unsigned int*)dalloc(nhist * 4 + 16);
    unsigned long long* staged = (unsigned long long*)dalloc((size_t)n * W * 8 + 16);
    uint32_t* starts = (uint32_t*)dalloc(((size_t)nb + 1) * 4);
    unsigned int* head = (unsigned int*)dalloc(16);
    unsigned long long* recs = nullptr;
    SortedItem* d_items = nullptr;
    unsigned int* counts = nullptr;
    // Hot keys (one-word keys, batches of at least kHotMinRows rows;
    // QDB_GPU_HOT_KEYS=0 turns them off): a sample names them, the count
    // leaves their rows out and the scatter aggregates them in LDS.
    const size_t hot_base = ((size_t)nb * 4 + 7) & ~(size_t)7;
    const size_t scatter_hot_lds = hot_base + agg::kHotSetBytes + (size_t)agg::kHotMax * (size_t)R * 8;
    hipFunction_t f_hot_sample = nullptr, f_hot_reduce = nullptr;
    int64_t hot_min_rows = kHotMinRows;
    if (const char* e = getenv("QDB_GPU_HOT_MIN_ROWS")) hot_min_rows = atoll(e) > agg::kHotSamples ? atoll(e) : agg::kHotSamples;
    bool hot = KW == 1 && n >= hot_min_rows && hot_keys_enabled() && scatter_hot_lds + 1024 <= max_lds_per_block();
    if (hot) {
        f_hot_sample = plan_jit_function(P, "qdb_agg_hot_sample");
        f_hot_reduce = plan_jit_function(P, "qdb_agg_hot_reduce");
        hot = f_hot_sample && f_hot_reduce;
    }
    unsigned long long* hot_keys = hot ? (unsigned long long*)dalloc((size_t)agg::kHotMax * 8) : nullptr;
    unsigned int* hot_n = hot ? (unsigned int*)dalloc(16) : nullptr;
    unsigned long long* hot_parts = hot ? (unsigned long long*)dalloc((size_t)nblk * agg::kHotMax * (size_t)R * 8) : nullptr;
    auto release = [&]() {
        dfree(hist);
        dfree(offs);
        dfree(staged);
        dfree(starts);
        dfree(head);
        dfree(recs);
        dfree(d_items);
        dfree(counts);
        dfree(hot_keys);
        dfree(hot_n);
        dfree(hot_parts);
    };
    if (!hist || !offs || !staged || !starts || !head || (hot && (!hot_keys || !hot_n || !hot_parts))) {
        release();
        return false;
    }
    hipFunction_t f_count = plan_jit_function(P, "qdb_agg_count");
    hipFunction_t f_scatter = plan_jit_function(P, "qdb_agg_scatter");
    if (!f_count || !f_scatter) {
        release();
        return fail("bucketed aggregate kernels unavailable");
    }
    TimedRegion timer("agg_update");
    int shift = 64 - bits;
    const unsigned int hist_lds = nb * 4;
    if (hot) {
        long long nn = (long long)n;
        void* args[] = {&A, &nn, &hot_keys, &hot_n};
        if (!hip_ok(hipModuleLaunchKernel(f_hot_sample, 1, 1, 1, agg::kHotSampleBlock, 1, 1, 0, stream(), args, nullptr),
                "hot sample launch")) {
            release();
            return false;
        }
    }
    {
        unsigned int nbv = nb;
        const unsigned long long* hk = hot_keys;
        const unsigned int* hn = hot_n;
        void* args[] = {&A, &hist, &shift, &nbv, &chunk, &hk, &hn};
        const unsigned int lds = hot ? (unsigned int)(hot_base + agg::kHotSetBytes) : hist_lds;
        if (!hip_ok(hipModuleLaunchKernel(f_count, (unsigned)nblk, 1, 1, kSortedBlock, 1, 1, lds, stream(), args, nullptr),
                "count launch")) {
            release();
            return false;
        }
    }
    if (!exclusive_scan_u32(hist, offs, (int64_t)nhist)) {
        release();
        return false;
    }
    {
        unsigned int nbv = nb;
        const unsigned long long* hk = hot_keys;
        const unsigned int* hn = hot_n;
        void* args[] = {&A, &offs, &staged, &shift, &nbv, &chunk, &hk, &hn, &hot_parts};
        const unsigned int lds = hot ? (unsigned int)scatter_hot_lds : hist_lds;
        if (!hip_ok(hipModuleLaunchKernel(f_scatter, (unsigned)nblk, 1, 1, kSortedBlock, 1, 1, lds, stream(), args, nullptr),
                "scatter launch")) {
            release();
            return false;
        }
    }
    // Bucket b starts where the first range's run of it starts.
    hipLaunchKernelGGL(k_gather_starts, dim3(grid_for(nb, kBlock, 1)), dim3(kBlock), 0, stream(), offs, nb, (unsigned int)nblk, starts);
    std::vector<uint32_t> h_starts((size_t)nb + 1);
    unsigned int nhot = 0;
    if (!hip_ok(hipGetLastError(), "bucket launch") ||
        !hip_ok(hipMemcpyAsync(h_starts.data(), starts, (size_t)nb * 4, hipMemcpyDeviceToHost, stream()), "starts") ||
        (hot && !hip_ok(hipMemcpyAsync(&nhot, hot_n, 4, hipMemcpyDeviceToHost, stream()), "hot count")) || !sync("starts")) {
        release();
        return false;
    }
    h_starts[nb] = (uint32_t)n;
    std::vector<SortedItem> items;
    for (uint32_t b = 0; b < B; ++b) {
        const int64_t r0 = h_starts[b], r1 = h_starts[b + 1];
        if (r1 <= r0) continue;
        const bool owned = r1 - r0 <= kSortedItemRows;
        for (int64_t r = r0; r < r1; r += kSortedItemRows) {
            SortedItem it;
            it.row0 = r;
            it.row1 = r + kSortedItemRows < r1 ? r + kSortedItemRows : r1;
            it.owned = owned ? 1 : 0;
            it.pad = 0;
            items.push_back(it);
        }
    }
    // The bucket aggregate takes items [0, nsorted); the hot keys' records,
    // when there are any, are one more item after them (records n, n + 1,
    // ...), owned: no bucket holds a hot key.
    int nsorted = (int)items.size();
    if (nhot > 0) {
        SortedItem it;
        it.row0 = n;
        it.row1 = n + (int64_t)nhot;
        it.owned = 1;
        it.pad = 0;
        items.push_back(it);
    }
    int nitems = (int)items.size();
    if (nitems == 0) {
        release();
        return true;
    }
    recs = (unsigned long long*)dalloc((size_t)(n + (hot ? agg::kHotMax : 0)) * R * 8 + 16);
    d_items = (SortedItem*)upload(items.data(), items.size() * sizeof(SortedItem));
    counts = (unsigned int*)dalloc((size_t)nitems * 4 + 16);
    if (!recs || !d_items || !counts || !zero_fill(head, 16)) {
        release();
        return false;
    }
    if (nhot > 0) {
        unsigned long long* hot_out = recs + (size_t)n * R;
        int nblk_i = (int)nblk;
        void* args[] = {&A, &hot_parts, &nblk_i, &hot_n, &hot_out};
        const unsigned int grid = (nhot * kWave + kBlock - 1) / kBlock;
        if (!hip_ok(hipModuleLaunchKernel(f_hot_reduce, grid, 1, 1, kBlock, 1, 1, 0, stream(), args, nullptr), "hot reduce launch") ||
            !hip_ok(hipMemcpyAsync(counts + nsorted, &nhot, 4, hipMemcpyHostToDevice, stream()), "hot item count")) {
            release();
            return false;
        }
    }
    if (nsorted > 0) {
        const int grid = nsorted < num_cus() ? nsorted : num_cus();
        void* args[] = {&A, &staged, &d_items, &nsorted, &head, &recs, &counts};
        if (!hip_ok(hipModuleLaunchKernel(f_sorted, (unsigned)grid, 1, 1, kSortedBlock, 1, 1, (unsigned)lds, stream(), args, nullptr),
                "bucket aggregate launch")) {
            release();
            return false;
        }
    }
    std::vector<uint32_t> h_counts((size_t)nitems);
    if (!hip_ok(hipMemcpyAsync(h_counts.data(), counts, h_counts.size() * 4, hipMemcpyDeviceToHost, stream()), "counts") || !sync("bucket aggregate")) {
        release();
        return false;
    }
    uint64_t total = 0;
    for (uint32_t c : h_counts) total += c & ~kSharedItem;
    if (lazy) {
        // Owned runs stay pending; the records of split buckets (and of
        // items whose rows overflowed the LDS table) may repeat keys and go
        // to a slot table sized for them.
        std::vector<SortedItem> own, shared;
        std::vector<uint32_t> own_n, shared_n, prefix;
        uint64_t own_total = 0, shared_total = 0;
        for (int i = 0; i < nitems; ++i) {
            const uint32_t c = h_counts[(size_t)i];
            const uint32_t k = c & ~kSharedItem;
            if (k == 0) continue;
            if (items[(size_t)i].owned && !(c & kSharedItem)) {
                prefix.push_back((uint32_t)own_total);
                own.push_back(items[(size_t)i]);
                own_n.push_back(k);
                own_total += k;
            } else {
                shared.push_back(items[(size_t)i]);
                shared_n.push_back(c);
                shared_total += k;
            }
        }
        prefix.push_back((uint32_t)own_total);
        bool ok = true;
        if (!shared.empty()) {
            uint64_t cap = 1024;
            while (cap / 2 < shared_total + 1024) cap *= 2;
            ok = table_alloc(T, cap);
            if (ok) T->allocated = true;
            SortedItem* d_shared = ok ? (SortedItem*)upload(shared.data(), shared.size() * sizeof(SortedItem)) : nullptr;
            unsigned int* d_shared_n = ok ? (unsigned int*)upload(shared_n.data(), shared_n.size() * 4) : nullptr;
            int nshared = (int)shared.size();
            UpdateLaunch L;
            ok = ok && d_shared && d_shared_n && prepare_launch(T, bound, 0, 1, (size_t)shared_total + 1024, &L) && set_counters(T) &&
                zero_fill(head, 16);
            if (ok) {
                const int grid = nshared < num_cus() * 4 ? nshared : num_cus() * 4;
                void* args[] = {&L.A, &recs, &d_shared, &d_shared_n, &nshared, &head};
                ok = hip_ok(hipModuleLaunchKernel(f_merge, (unsigned)grid, 1, 1, kBlock, 1, 1, 0, stream(), args, nullptr), "merge launch");
            }
            unsigned int cnt[4];
            ok = ok && read_counters(T, cnt) && settle_launch(&L, cnt);
            L.release();
            dfree(d_shared);
            dfree(d_shared_n);
        }
        if (ok && !own.empty()) {
            T->pend_items = (SortedItem*)upload(own.data(), own.size() * sizeof(SortedItem));
            T->pend_counts = (uint32_t*)upload(own_n.data(), own_n.size() * 4);
            T->pend_prefix = (uint32_t*)upload(prefix.data(), prefix.size() * 4);
            ok = T->pend_items && T->pend_counts && T->pend_prefix;
            if (ok) {
                T->pend_recs = (uint64_t*)recs;
                recs = nullptr;  // kept by the table
                T->pend_runs = (int)own.size();
                T->pend_groups = (unsigned int)own_total;
            } else {
                table_drop_pending(T);
            }
        }
        release();
        return ok;
    }
    // Room for every record as a new group keeps the merge free of spills.
    if ((uint64_t)T->groups + total > T->capacity / 2 && !table_grow(T, (uint64_t)T->groups + total)) {
        release();
        return false;
    }
    UpdateLaunch L;
    if (!prepare_launch(T, bound, 0, 1, (size_t)total + 1024, &L) || !set_counters(T) || !zero_fill(head, 16)) {
        L.release();
        release();
        return false;
    }
    {
        const int grid = nitems < num_cus() * 4 ? nitems : num_cus() * 4;
        void* args[] = {&L.A, &recs, &d_items, &counts, &nitems, &head};
        if (!hip_ok(hipModuleLaunchKernel(f_merge, (unsigned)grid, 1, 1, kBlock, 1, 1, 0, stream(), args, nullptr), "merge launch")) {
            L.release();
            release();
            return false;
        }
    }
    unsigned int cnt[4];
    const bool ok = read_counters(T, cnt) && settle_launch(&L, cnt);
    L.release();
    release();
    return ok;
}

// Whether bucketed updates look for hot keys (QDB_GPU_HOT_KEYS=0: no).
bool hot_keys_enabled()
{
    const char* e = getenv("QDB_GPU_HOT_KEYS");
    return !(e && e[0] == '0');
}

// Whether the bucketed path can take a batch of `rows` rows of this plan.
bool bucketed_possible(AggPlan* P, int64_t rows)
{
    return rows <= ((int64_t)1 << 30) && stage_layout(P).words > 0 && plan_jit_function(P, "qdb_agg_sorted") != nullptr;
}

// The configuration of an update when nothing is tuned: the bucketed path
// for large batches into large tables, else LDS tables of 64 KiB at two
// workgroups per CU.
TuneConfig default_config(AggPlan* P, const AggTable* T, int64_t rows)
{
    TuneConfig c = {0, 64, 2, 0, 0, 1};
    if (rows <= ((int64_t)1 << 30) && sorted_path_wanted(P, T, rows)) c.path = 1;
    // A table made for a handful of groups: four workgroups per CU with
    // small LDS tables (their groups leave through partial records, so the
    // larger grid adds no same-line atomics on the table).
    if (T->init_slots <= 64) {
        c.lds_kib = 32;
        c.grid_per_cu = 4;
    }
    return c;
}

TuneConfig choose_config(AggTable* T, const TRowSet* batch);

// agg_dispatch(Update) in configuration `cfg`.
bool table_update_cfg(AggTable* T, const TRowSet* batch, const TuneConfig& cfg)
{
    AggPlan* P = T->plan;
    if (batch->RowCount <= 0) return true;
    if (cfg.path == 1 && bucketed_possible(P, batch->RowCount)) {
        AggArgs bound;
        if (!bind_batch(P, batch, &bound)) return false;
        return table_update_sorted(T, bound, batch->RowCount, cfg.fanout_shift, cfg.part_blocks);
    }
    AggArgs bound;
    if (!bind_batch(P, batch, &bound)) return false;
    const int KW = P->KW, S = P->S;
    size_t lds = 0;
    lds_layout(P, &bound, &lds, cfg.lds_kib, cfg.lds_copies);
    const int per_cu = cfg.grid_per_cu >= 1 ? cfg.grid_per_cu : 2;
    const int64_t chunk = (int64_t)1 << 30;
    for (int64_t base = 0; base < batch->RowCount; base += chunk) {
        const int64_t n = batch->RowCount - base < chunk ? batch->RowCount - base : chunk;
        const int grid = grid_for(n, kBlock, per_cu);
        UpdateLaunch L;
        if (!prepare_launch(T, bound, base, n, (size_t)grid * (bound.lds_groups ? bound.lds_groups : 1) + 1024, &L)) return false;
        if (!set_counters(T)) {
            L.release();
            return false;
        }
        hipError_t launched = hipSuccess;
        hipFunction_t fn = plan_jit_function(P, "qdb_agg_update");
        // A table of few groups updated by a large grid: the workgroups
        // write their groups as partial records to slots of their own, and
        // a second launch (one workgroup per kPartsPerReducer slots)
        // combines them in LDS before its atomics on the table.
        uint64_t* parts = nullptr;
        unsigned int* part_n = nullptr;
        hipFunction_t f_reduce = nullptr;
        const int part_cap = bound.lds_groups < kPartCap ? bound.lds_groups : kPartCap;
        if (fn && grid >= 4 * kPartsPerReducer && part_cap > 0 && T->groups <= (uint64_t)part_cap && T->capacity <= kPartMaxSlots &&
            (f_reduce = plan_jit_function(P, "qdb_agg_reduce")) != nullptr) {
            parts = (uint64_t*)dalloc((size_t)grid * (size_t)part_cap * (size_t)(KW + bound.nwords) * 8);
            part_n = (unsigned int*)dalloc((size_t)grid * 4);
            if (!parts || !part_n) {
                dfree(parts);
                dfree(part_n);
                L.release();
                return false;
            }
        }
        {
            TimedRegion timer("agg_update");
            if (fn) {
                AggArgs A = L.A;
                if (parts) {
                    A.part_recs = parts;
                    A.part_n = part_n;
                    A.part_cap = part_cap;
                }
                void* args[] = {&A};
                launched = hipModuleLaunchKernel(fn, (unsigned)grid, 1, 1, kBlock, 1, 1, (unsigned)lds, stream(), args, nullptr);
                if (parts && launched == hipSuccess) {
                    AggArgs B = L.A;
                    B.part_cap = part_cap;
                    int nparts = grid, per = kPartsPerReducer;
                    const int nred = (grid + per - 1) / per;
                    void* rargs[] = {&B, &parts, &part_n, &nparts, &per};
                    launched = hipModuleLaunchKernel(f_reduce, (unsigned)nred, 1, 1, kBlock, 1, 1, (unsigned)lds, stream(), rargs, nullptr);
                }
            } else {
                agg::launch_update(KW, S, L.A, grid, lds);
            }
        }
        unsigned int cnt[4];
        const bool ok = hip_ok(launched, "agg update launch") && hip_ok(hipGetLastError(), "agg update launch") &&
            read_counters(T, cnt) && settle_launch(&L, cnt);
        dfree(parts);
        dfree(part_n);
        L.release();
        if (!ok) return false;
    }
    return true;
}

// agg_dispatch(Update): every selected row of the batch, exactly once, in
// the configuration chosen for its shape.
bool table_update(AggTable* T, const TRowSet* batch)
{
    if (batch->RowCount <= 0) return true;
    return table_update_cfg(T, batch, choose_config(T, batch));
}

// Rows per work item of a batched update.
constexpr int64_t kBatchItemRows = 32768;

// The batched update's instances that run on their own (the bucketed path,
// batches above 2^30 rows), largest first. Called from the library stream,
// they are spread over that stream and up to two more, each driven by a
// thread of its own (QDB_GPU_BATCH_STREAMS = 1: one after another), so that
// one instance's host round trips overlap another's kernels; tables are
// independent, and the pool reuses a block across streams only after its
// free has completed.
bool update_own(AggTable* const* tables, const TRowSet* const* batches, const std::vector<TuneConfig>& cfgs,
    std::vector<int64_t> own)
{
    if (own.empty()) return true;
    std::sort(own.begin(), own.end(), [&](int64_t a, int64_t b) { return batches[a]->RowCount > batches[b]->RowCount; });
    int helpers = 2;
    if (const char* e = getenv("QDB_GPU_BATCH_STREAMS")) helpers = atoi(e) - 1;
    if (helpers > kMaxStreams - 2) helpers = kMaxStreams - 2;
    if (thread_stream_index() != 0 || own.size() < 2 || helpers < 1) {
        for (int64_t i : own) {
            if (!table_update_cfg(tables[i], batches[i], cfgs[(size_t)i])) return false;
        }
        return true;
    }
    if ((size_t)helpers > own.size() - 1) helpers = (int)own.size() - 1;
    std::atomic<size_t> next(0);
    std::atomic<bool> failed(false);
    std::vector<std::string> errors((size_t)helpers + 1);
    auto work = [&](int slot) {
        for (;;) {
            const size_t k = next.fetch_add(1);
            if (k >= own.size() || failed.load()) break;
            const int64_t i = own[k];
            if (!table_update_cfg(tables[i], batches[i], cfgs[(size_t)i])) {
                errors[(size_t)slot] = last_error_text().empty() ? std::string("aggregate update failed") : last_error_text();
                failed.store(true);
                break;
            }
        }
        if (slot > 0 && !sync("batched update helper") && errors[(size_t)slot].empty()) {
            errors[(size_t)slot] = last_error_text();
            failed.store(true);
        }
    };
    std::vector<std::thread> threads;
    for (int h = 1; h <= helpers; ++h) {
        threads.emplace_back([&, h]() {
            if (!set_thread_stream(h)) {
                errors[(size_t)h] = last_error_text();
                failed.store(true);
                return;
            }
            work(h);
        });
    }
    work(0);
    for (auto& t : threads) t.join();
    if (failed.load()) {
        for (const std::string& e : errors) {
            if (!e.empty()) return fail("%s", e.c_str());
        }
        return fail("batched update failed");
    }
    return true;
}

// Updates tables[i] with batches[i] for every i in ONE persistent launch:
// workgroups take (instance, row range) items from a device queue. Leftover
// rows and records are replayed per table afterwards, as after a plain
// update. Batches above 2^30 rows take the plain path.
bool batched_update(AggPlan* P, AggTable* const* tables, const TRowSet* const* batches, int64_t n)
{
    const int KW = P->KW, S = P->S;
    std::vector<UpdateLaunch> launches;
    std::vector<AggArgs> insts;
    std::vector<BatchItem> items;
    std::vector<int64_t> which;
    launches.reserve((size_t)n);
    size_t lds = 0;
    bool ok = true;
    // One instance (after those updated on their own) gains nothing from
    // the queue: it takes the plain update.
    // Each instance's configuration (tuned per shape when tuning is on):
    // the bucketed ones, and batches above 2^30 rows, update on their own.
    std::vector<TuneConfig> cfgs((size_t)n);
    int64_t queued = 0;
    for (int64_t i = 0; i < n; ++i) {
        const int64_t rows = batches[i]->RowCount;
        if (rows <= 0) continue;
        cfgs[(size_t)i] = choose_config(tables[i], batches[i]);
        queued += rows <= ((int64_t)1 << 30) && cfgs[(size_t)i].path == 0;
    }
    if (queued <= 1) {
        for (int64_t i = 0; i < n && ok; ++i) {
            if (batches[i]->RowCount > 0) ok = table_update_cfg(tables[i], batches[i], cfgs[(size_t)i]);
        }
        return ok;
    }
    auto release_all = [&]() {
        for (auto& L : launches) L.release();
    };
    // Very large batches, and those the bucketed path takes, are updated on
    // their own (after the persistent launch is queued, see below).
    std::vector<int64_t> own;
    for (int64_t i = 0; i < n && ok; ++i) {
        AggTable* T = tables[i];
        const TRowSet* b = batches[i];
        if (b->RowCount <= 0) continue;
        if (b->RowCount > ((int64_t)1 << 30) || cfgs[(size_t)i].path != 0) {
            own.push_back(i);
            continue;
        }
        AggArgs bound;
        if (!bind_batch(P, b, &bound)) {
            ok = false;
            break;
        }
        lds_layout(P, &bound, &lds);
        const int64_t nitems = (b->RowCount + kBatchItemRows - 1) / kBatchItemRows;
        launches.emplace_back();
        if (!prepare_launch(T, bound, 0, b->RowCount, (size_t)nitems * (bound.lds_groups ? bound.lds_groups : 1) + 1024,
                &launches.back()) ||
            !set_counters(T)) {
            launches.pop_back();
            ok = false;
            break;
        }
        const int inst = (int)insts.size();
        insts.push_back(launches.back().A);
        which.push_back(i);
        for (int64_t r = 0; r < b->RowCount; r += kBatchItemRows) {
            BatchItem it;
            it.inst = inst;
            it.pad = 0;
            it.row0 = r;
            it.row1 = r + kBatchItemRows < b->RowCount ? r + kBatchItemRows : b->RowCount;
            items.push_back(it);
        }
    }
    if (!ok || insts.empty()) {
        release_all();
        return ok && update_own(tables, batches, cfgs, own);
    }
    AggArgs* d_insts = (AggArgs*)upload(insts.data(), insts.size() * sizeof(AggArgs));
    BatchItem* d_items = (BatchItem*)upload(items.data(), items.size() * sizeof(BatchItem));
    unsigned int* d_head = (unsigned int*)dalloc(16);
    ok = d_insts && d_items && d_head && zero_fill(d_head, 16);
    if (ok) {
        const int nitems = (int)items.size();
        const int grid = nitems < num_cus() * 2 ? nitems : num_cus() * 2;
        hipFunction_t fn = plan_jit_function(P, "qdb_agg_batch");
        hipError_t launched = hipSuccess;
        {
            TimedRegion timer("agg_batch");
            if (fn) {
                void* args[] = {&d_insts, &d_items, (void*)&nitems, &d_head};
                launched = hipModuleLaunchKernel(fn, (unsigned)grid, 1, 1, kBlock, 1, 1, (unsigned)lds, stream(), args, nullptr);
            } else {
                agg::launch_batch(KW, S, d_insts, d_items, nitems, d_head, grid, lds);
            }
        }
        ok = hip_ok(launched, "agg batch launch") && hip_ok(hipGetLastError(), "agg batch launch");
    }
    // The instances on their own run while the persistent launch does.
    ok = ok && update_own(tables, batches, cfgs, own);
    // Every table's counters in one synchronisation.
    std::vector<unsigned int> cnt(launches.size() * 4, 0u);
    for (size_t k = 0; ok && k < launches.size(); ++k) {
        ok = hip_ok(hipMemcpyAsync(&cnt[k * 4], launches[k].T->counters, 16, hipMemcpyDeviceToHost, stream()), "counters");
    }
    ok = ok && sync("agg batch");
    for (size_t k = 0; ok && k < launches.size(); ++k) {
        AggTable* T = launches[k].T;
        T->groups = cnt[k * 4];
        T->extra_used = cnt[k * 4 + 3];
        ok = settle_launch(&launches[k], &cnt[k * 4]);
    }
    release_all();
    dfree(d_insts);
    dfree(d_items);
    dfree(d_head);
    return ok;
}

// The finalize's device temporaries, released once its kernels are done.
struct FinishTemps {
    uint32_t* index = nullptr;
    uint32_t* counts = nullptr;
    uint32_t* offsets = nullptr;
    void release()
    {
        dfree(index);
        dfree(counts);
        dfree(offsets);
        index = counts = offsets = nullptr;
    }
};

// agg_finish_rowset, first half: groups -> columnar TRowSet, every kernel
// launched, nothing waited for; returns the row count or -1.
// The output rowset of T (every column's data and validity bitmap in one
// block of 256-byte aligned parts: one pool call per output instead of one
// per buffer) and the finalize arguments that write it, the index aside;
// returns the group count or -1.
int64_t table_finish_outputs(AggTable* T, TRowSet* out, FinArgs* F)
{
    AggPlan* P = T->plan;
    *F = P->fin;
    F->recs = T->recs;
    F->stride = T->rstride;
    F->index = nullptr;
    // The pending runs first, then the slot table's groups.
    F->pend_recs = T->pend_recs;
    F->pend_items = T->pend_items;
    F->pend_prefix = T->pend_prefix;
    F->pend_runs = T->pend_runs;
    F->pend_n = T->pend_groups;
    const unsigned int n = table_size(T);
    OwnedRowSet* o = init_owned_rowset(out, n, P->nout);
    F->groups = n;
    // Every column's data and validity bitmap in one block, 256-byte aligned
    // parts (one pool call per output instead of one per buffer).
    const size_t mask_bytes = (size_t)((n + 7) / 8) + 8;
    auto aligned = [](size_t b) { return (b + 255) & ~(size_t)255; };
    size_t total = 0;
    for (int c = 0; c < P->nout; ++c) {
        const QdbValueType& t = P->out_types[c];
        total += aligned(t.Kind == QDB_KIND_BOOL ? mask_bytes : (size_t)n * (size_t)t.Width + 8);
        if (t.Nullable) total += aligned(mask_bytes);
    }
    char* block = (char*)dalloc(total ? total : 256);
    if (!block) {
        return -1;
    }
    o->dev.push_back(block);
    size_t at = 0;
    for (int c = 0; c < P->nout; ++c) {
        const QdbValueType& t = P->out_types[c];
        const size_t bytes = t.Kind == QDB_KIND_BOOL ? mask_bytes : (size_t)n * (size_t)t.Width + 8;
        char* data = block + at;
        at += aligned(bytes);
        if (t.Kind == QDB_KIND_BOOL && !zero_fill(data, bytes)) {
            return -1;
        }
        F->out_data[c] = data;
        o->cols[(size_t)c].Data = data;
        uint8_t* mask = nullptr;
        if (t.Nullable) {
            mask = (uint8_t*)(block + at);
            at += aligned(mask_bytes);
        }
        F->out_mask[c] = mask;
        o->cols[(size_t)c].Mask = mask;
    }
    return (int64_t)n;
}

int64_t table_finish_launch(AggTable* T, TRowSet* out, FinishTemps* tmp)
{
    AggPlan* P = T->plan;
    FinArgs F;
    uint32_t*& index = tmp->index;
    uint32_t*& counts = tmp->counts;
    uint32_t*& offsets = tmp->offsets;
    auto release = [&]() { tmp->release(); };
    if (T->groups > 0) {
        // The slots that hold groups, in slot order.
        const uint64_t slots = T->capacity + 1;
        const int64_t chunks = (int64_t)((slots + kOccChunk - 1) / kOccChunk);
        index = (uint32_t*)dalloc((size_t)T->groups * 4 + 16);
        counts = (uint32_t*)dalloc((size_t)chunks * 4 + 16);
        offsets = (uint32_t*)dalloc((size_t)chunks * 4 + 16);
        if (!index || !counts || !offsets) {
            release();
            return -1;
        }
        const int extra = (int)T->extra_used;
        TimedRegion timer("agg_finalize");
        switch (P->KW) {
        case 1: hipLaunchKernelGGL(k_occ_counts<1>, dim3(chunks), dim3(kBlock), 0, stream(), T->recs, T->ctrl, T->occ, T->rstride, T->capacity, extra, counts); break;
        case 2: hipLaunchKernelGGL(k_occ_counts<2>, dim3(chunks), dim3(kBlock), 0, stream(), T->recs, T->ctrl, T->occ, T->rstride, T->capacity, extra, counts); break;
        case 3: hipLaunchKernelGGL(k_occ_counts<3>, dim3(chunks), dim3(kBlock), 0, stream(), T->recs, T->ctrl, T->occ, T->rstride, T->capacity, extra, counts); break;
        default: hipLaunchKernelGGL(k_occ_counts<4>, dim3(chunks), dim3(kBlock), 0, stream(), T->recs, T->ctrl, T->occ, T->rstride, T->capacity, extra, counts); break;
        }
        if (!hip_ok(hipGetLastError(), "occupancy launch") || !exclusive_scan_u32(counts, offsets, chunks)) {
            release();
            return -1;
        }
        switch (P->KW) {
        case 1: hipLaunchKernelGGL(k_occ_index<1>, dim3(chunks), dim3(kBlock), 0, stream(), T->recs, T->ctrl, T->occ, T->rstride, T->capacity, extra, offsets, index); break;
        case 2: hipLaunchKernelGGL(k_occ_index<2>, dim3(chunks), dim3(kBlock), 0, stream(), T->recs, T->ctrl, T->occ, T->rstride, T->capacity, extra, offsets, index); break;
        case 3: hipLaunchKernelGGL(k_occ_index<3>, dim3(chunks), dim3(kBlock), 0, stream(), T->recs, T->ctrl, T->occ, T->rstride, T->capacity, extra, offsets, index); break;
        default: hipLaunchKernelGGL(k_occ_index<4>, dim3(chunks), dim3(kBlock), 0, stream(), T->recs, T->ctrl, T->occ, T->rstride, T->capacity, extra, offsets, index); break;
        }
        if (!hip_ok(hipGetLastError(), "index launch")) {
            release();
            return -1;
        }
    }
    const int64_t n = table_finish_outputs(T, out, &F);
    if (n < 0) {
        release();
        return -1;
    }
    F.index = T->groups > 0 ? index : nullptr;
    if (n > 0) {
        TimedRegion timer("agg_finalize");
        const int grid = grid_for(n, kBlock, 8);
        hipLaunchKernelGGL(k_agg_finalize, dim3(grid), dim3(kBlock), 0, stream(), F);
        if (!hip_ok(hipGetLastError(), "agg finalize launch")) {
            release();
            return -1;
        }
    }
    return (int64_t)n;
}

// Second half, once the stream has been synchronized: the temporaries
// released, the output moved to the host when the spec asks.
bool table_finish_done(AggTable* T, TRowSet* out, FinishTemps* tmp)
{
    tmp->release();
    if (T->plan->spec.OutputOnHost && !move_rowset_to_host(out, T->plan->out_types)) return false;
    return publish_rowset(out);
}

// Host milliseconds of one query in configuration `cfg` on a scratch table
// of T's size: set-up, the update of `batch` and the finish into HBM (the
// paths leave different work to the finish); < 0 when it fails.
double time_config(AggTable* T, const TRowSet* batch, const TuneConfig& cfg)
{
    AggPlan* P = T->plan;
    if (!sync("tune")) return -1.0;
    const auto t0 = std::chrono::steady_clock::now();
    AggTable* X = table_new(P, (int64_t)(T->cap_hint > T->capacity ? T->cap_hint : T->capacity));
    if (!X) return -1.0;
    bool ok = table_seed(X) && table_update_cfg(X, batch, cfg);
    TRowSet out;
    memset(&out, 0, sizeof(out));
    FinishTemps tmp;
    if (ok) ok = table_finish_launch(X, &out, &tmp) >= 0;
    ok = sync("tune") && ok;
    tmp.release();
    destroy_kernel_rowset(&out);
    table_free(X);
    const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return ok ? ms : -1.0;
}

// The autotuner: the configuration of an update of `batch` into T. A forced
// configuration (QDB_GPU_TUNE_FORCE) wins; else the one cached for the
// shape bucket (key width, table slots, rows); else, with tuning on, every
// candidate is timed twice on a scratch table (the second time counts: the
// first may include one-time costs) and the fastest is cached; else the
// sizing rule's default.
TuneConfig choose_config(AggTable* T, const TRowSet* batch)
{
    AggPlan* P = T->plan;
    const int64_t rows = batch->RowCount;
    const bool bucketed_ok = bucketed_possible(P, rows);
    TuneConfig cfg;
    if (tune_forced(&cfg)) {
        if (cfg.path == 1 && !bucketed_ok) cfg = TuneConfig{0, 64, 2, 0, 0, 1};
        return cfg;
    }
    const uint64_t slots = T->cap_hint > T->capacity ? T->cap_hint : T->capacity;
    const TuneKey key = tune_key(P->KW, slots, rows);
    if (tune_lookup(key, &cfg)) {
        if (cfg.path == 1 && !bucketed_ok) cfg = TuneConfig{0, 64, 2, 0, 0, 1};
        return cfg;
    }
    if (!tune_enabled() || rows <= 0) return default_config(P, T, rows);
    TuneConfig best = default_config(P, T, rows);
    double best_ms = -1.0;
    for (const TuneConfig& c : tune_candidates(bucketed_ok)) {
        double ms = time_config(T, batch, c);
        if (ms >= 0) ms = time_config(T, batch, c);
        if (ms < 0) {
            clear_error();
            continue;
        }
        if (best_ms < 0 || ms < best_ms) {
            best_ms = ms;
            best = c;
        }
    }
    if (best_ms >= 0) tune_store(key, best, best_ms);
    return best;
}

// agg_finish_rowset: groups -> columnar TRowSet; destroys the table.
int64_t table_finish(AggTable* T, TRowSet* out)
{
    FinishTemps tmp;
    const int64_t n = table_finish_launch(T, out, &tmp);
    if (n < 0) {
        tmp.release();
        return -1;
    }
    const bool synced = sync("agg finalize");
    if (!table_finish_done(T, out, &tmp) || !synced) return -1;
    return n;
}

int64_t agg_dispatch_impl(AggPlan* P, void* ht, TRowSet* batch, int64_t arg, int64_t op)
{
    clear_error();
    HtBlob* B = (HtBlob*)ht;
    if (!P) return op == QDB_AGG_OP_UPDATE ? -1 : 0;
    switch (op) {
    case QDB_AGG_OP_INIT:
    case QDB_AGG_OP_INIT_ROWS: {
        if (op == QDB_AGG_OP_INIT && (arg < 8 || (arg & (arg - 1)) != 0)) {
            fail("Init needs a power-of-two capacity of at least 8, got %lld", (long long)arg);
            return 0;
        }
        if (B->table) table_free(B->table);
        memset(B, 0, sizeof(*B));
        // Init's arg is the reference's swiss slot count; InitRows' arg is
        // groups, which want twice as many slots here (load at most 1/2).
        const int64_t slots = op == QDB_AGG_OP_INIT ? arg : 2 * arg;
        AggTable* T = table_new(P, slots);
        if (!T) return 0;
        if (!table_seed(T)) {
            table_free(T);
            return 0;
        }
        B->table = T;
        B->Capacity = (int64_t)T->capacity;
        B->Size = T->groups;
        B->NumAggs = P->spec.NumAggs;
        B->NumKeys = P->spec.NumKeys;
        B->KeySize = P->KW * 8;
        return 1;
    }
    case QDB_AGG_OP_SWITCH_TO_ROWS_AT:
        // One layout serves every cardinality here; the hint is recorded.
        B->RowSwitchGroups = arg;
        return 1;
    case QDB_AGG_OP_EXPECT_ROW_GROUPS:
        B->RowExpectedGroups = arg;
        if (B->table) {
            B->table->expect_groups = arg;
            (void)arg;
        }
        return 1;
    case QDB_AGG_OP_UPDATE: {
        AggTable* T = B->table;
        if (!T) {
            fail("aggregate update before Init");
            return -1;
        }
        if (!batch) return (int64_t)table_size(T);
        if (!table_update(T, batch)) return -1;
        B->Size = table_size(T);
        B->Capacity = (int64_t)T->capacity;
        return (int64_t)table_size(T);
    }
    default:
        return 0;
    }
}

int64_t agg_finish_impl(AggPlan* P, void* ht, TRowSet* out)
{
    clear_error();
    HtBlob* B = (HtBlob*)ht;
    AggTable* T = B->table;
    if (!P || !T) {
        fail("aggregate finish without a table");
        return -1;
    }
    int64_t n = table_finish(T, out);
    table_free(T);
    memset(B, 0, sizeof(*B));
    return n;
}

template <int N>
struct AggDispatchTramp {
    static int64_t call(void* ht, TRowSet* batch, int64_t arg, int64_t op)
    {
        return agg_dispatch_impl(agg_slots().get(N), ht, batch, arg, op);
    }
};

template <int N>
struct AggFinishTramp {
    static int64_t call(void* ht, TRowSet* out)
    {
        return agg_finish_impl(agg_slots().get(N), ht, out);
    }
};

const auto kAggDispatchFns = make_tramp_table<QdbAggDispatchFn, AggDispatchTramp>(std::make_integer_sequence<int, kSlots>{});
const auto kAggFinishFns = make_tramp_table<QdbAggFinishFn, AggFinishTramp>(std::make_integer_sequence<int, kSlots>{});

}  // namespace
}  // namespace qdbgpu

using namespace qdbgpu;

extern "C" int qdb_gpu_compile_aggregate(const QdbAggregateSpec* spec, QdbAggregateKernels* out)
{
    clear_error();
    if (!spec || !out) return fail("null argument"), -1;
    if (!ensure_init()) return -1;
    AggPlan* P = new AggPlan();
    if (!compile_plan(*spec, P)) {
        delete P;
        return -1;
    }
    const int slot = agg_slots().add(P);
    if (slot < 0) {
        delete P;
        return fail("all %d aggregate slots are in use", kSlots), -1;
    }
    memset(out, 0, sizeof(*out));
    out->Dispatch = kAggDispatchFns.fns[slot];
    out->FinishRowSet = kAggFinishFns.fns[slot];
    out->NumAggs = spec->NumAggs;
    out->StateWords = P->proto.nwords;
    out->RowBytes = (int64_t)(P->KW + P->proto.nwords) * 8;
    out->Slot = slot;
    out->NumOutputColumns = P->nout;
    for (int c = 0; c < P->nout; ++c) out->OutputTypes[c] = P->out_types[c];
    return 0;
}

extern "C" int qdb_gpu_agg_update_batched(const QdbAggregateKernels* kernels, void* const* hts, const TRowSet* const* batches,
    int64_t n)
{
    clear_error();
    if (!kernels || (n > 0 && (!hts || !batches))) return fail("null argument"), -1;
    if (kernels->Slot < 0 || kernels->Slot >= kSlots) return fail("bad aggregate slot %lld", (long long)kernels->Slot), -1;
    AggPlan* P = agg_slots().get((int)kernels->Slot);
    if (!P) return fail("aggregate kernels are not compiled"), -1;
    std::vector<AggTable*> tables((size_t)n);
    for (int64_t i = 0; i < n; ++i) {
        HtBlob* B = (HtBlob*)hts[i];
        if (!B || !B->table) return fail("batched update of table %lld before Init", (long long)i), -1;
        if (B->table->plan != P) return fail("table %lld belongs to another aggregate", (long long)i), -1;
        if (!batches[i]) return fail("null batch %lld", (long long)i), -1;
        tables[(size_t)i] = B->table;
    }
    if (!batched_update(P, tables.data(), batches, n)) return -1;
    for (int64_t i = 0; i < n; ++i) {
        HtBlob* B = (HtBlob*)hts[i];
        B->Size = table_size(B->table);
        B->Capacity = (int64_t)B->table->capacity;
    }
    return 0;
}

// The finalize kernels of every table of a batched finish, launched without
// waiting. Called from the library stream with at least kSpreadMinTables
// tables, the tables (largest first) are spread over that stream and up to
// two more, each driven by a thread of its own that synchronizes its stream
// at the end (QDB_GPU_BATCH_STREAMS = 1: all on the calling stream): one
// table's launches are a handful of small kernels, which a single stream
// runs one after another. rows[i] >= 0 marks a table whose kernels were
// launched; false after the first failure (the error is kept).
constexpr int64_t kSpreadMinTables = 8;

bool finish_launch_spread(void* const* hts, TRowSet* outs, int64_t* rows, std::vector<FinishTemps>& tmp, std::vector<int64_t> order)
{
    int helpers = 2;
    if (const char* e = getenv("QDB_GPU_BATCH_STREAMS")) helpers = atoi(e) - 1;
    if (helpers > kMaxStreams - 2) helpers = kMaxStreams - 2;
    if (thread_stream_index() != 0 || (int64_t)order.size() < kSpreadMinTables || helpers < 1) {
        for (int64_t i : order) {
            rows[i] = table_finish_launch(((HtBlob*)hts[i])->table, &outs[i], &tmp[(size_t)i]);
            if (rows[i] < 0) return false;
        }
        return true;
    }
    std::sort(order.begin(), order.end(),
        [&](int64_t a, int64_t b) { return table_size(((HtBlob*)hts[a])->table) > table_size(((HtBlob*)hts[b])->table); });
    std::atomic<size_t> next(0);
    std::atomic<bool> failed(false);
    std::vector<std::string> errors((size_t)helpers + 1);
    auto work = [&](int slot) {
        for (;;) {
            const size_t k = next.fetch_add(1);
            if (k >= order.size() || failed.load()) break;
            const int64_t i = order[k];
            rows[i] = table_finish_launch(((HtBlob*)hts[i])->table, &outs[i], &tmp[(size_t)i]);
            if (rows[i] < 0) {
                errors[(size_t)slot] = last_error_text().empty() ? std::string("aggregate finalize failed") : last_error_text();
                failed.store(true);
                break;
            }
        }
        if (slot > 0 && !sync("batched finalize helper") && errors[(size_t)slot].empty()) {
            errors[(size_t)slot] = last_error_text();
            failed.store(true);
        }
    };
    std::vector<std::thread> threads;
    for (int h = 1; h <= helpers; ++h) {
        threads.emplace_back([&, h]() {
            if (!set_thread_stream(h)) {
                errors[(size_t)h] = last_error_text();
                failed.store(true);
                return;
            }
            work(h);
        });
    }
    work(0);
    for (auto& t : threads) t.join();
    if (failed.load()) {
        for (const std::string& e : errors) {
            if (!e.empty()) return fail("%s", e.c_str());
        }
        return fail("batched finalize failed");
    }
    return true;
}

// Whether table T's finalize can take the one-workgroup form
// (k_agg_finalize_small).
bool small_finish_ok(const AggTable* T)
{
    return T->plan->KW == 1 && T->pend_runs == 0 && T->groups > 0 && T->occ != nullptr && T->capacity <= kSmallFinMaxSlots &&
        table_size(T) <= kSmallFinMaxGroups;
}

// The outputs of tables hts[i], i in `which` (every one small_finish_ok),
// and one launch of k_agg_finalize_small for all of them, not waited for.
bool small_finish_launch(void* const* hts, TRowSet* outs, int64_t* rows, const std::vector<int64_t>& which)
{
    if (which.empty()) return true;
    std::vector<FinArgs> fa(which.size());
    std::vector<SmallFin> sf(which.size());
    for (size_t k = 0; k < which.size(); ++k) {
        const int64_t i = which[k];
        AggTable* T = ((HtBlob*)hts[i])->table;
        rows[i] = table_finish_outputs(T, &outs[i], &fa[k]);
        if (rows[i] < 0) return false;
        sf[k].occ = T->occ;
        sf[k].slots = T->capacity;
        sf[k].extra_used = T->extra_used ? 1 : 0;
        sf[k].pad = 0;
    }
    FinArgs* d_fa = (FinArgs*)upload(fa.data(), fa.size() * sizeof(FinArgs));
    SmallFin* d_sf = (SmallFin*)upload(sf.data(), sf.size() * sizeof(SmallFin));
    bool ok = d_fa && d_sf;
    if (ok) {
        TimedRegion timer("agg_finalize");
        hipLaunchKernelGGL(k_agg_finalize_small, dim3((unsigned)which.size()), dim3(kSmallFinBlock), 0, stream(), d_fa, d_sf);
        ok = hip_ok(hipGetLastError(), "small finalize launch");
    }
    dfree(d_fa);  // stream-ordered: reused only by work queued after the launch
    dfree(d_sf);
    return ok;
}

// agg_finish_rowset of hts[i] into outs[i] for every i: every table's
// finalize kernels are launched before one synchronization; each table is
// destroyed as by agg_finish_rowset. rows[i] gets the row count. Returns 0,
// or -1 (then every output that was made is released).
extern "C" int qdb_gpu_agg_finish_batched(const QdbAggregateKernels* kernels, void* const* hts, TRowSet* outs, int64_t* rows,
    int64_t n)
{
    clear_error();
    if (!kernels || (n > 0 && (!hts || !outs || !rows))) return fail("null argument"), -1;
    if (kernels->Slot < 0 || kernels->Slot >= kSlots) return fail("bad aggregate slot %lld", (long long)kernels->Slot), -1;
    AggPlan* P = agg_slots().get((int)kernels->Slot);
    if (!P) return fail("aggregate kernels are not compiled"), -1;
    for (int64_t i = 0; i < n; ++i) {
        HtBlob* B = (HtBlob*)hts[i];
        if (!B || !B->table) return fail("batched finish of table %lld before Init", (long long)i), -1;
        if (B->table->plan != P) return fail("table %lld belongs to another aggregate", (long long)i), -1;
    }
    std::vector<FinishTemps> tmp((size_t)n);
    for (int64_t i = 0; i < n; ++i) {
        memset(&outs[i], 0, sizeof(TRowSet));
        rows[i] = -1;
    }
    // Small tables: their outputs made here and their finalize in one
    // launch; the others on up to three streams.
    std::vector<int64_t> small, rest;
    for (int64_t i = 0; i < n; ++i) (small_finish_ok(((HtBlob*)hts[i])->table) ? small : rest).push_back(i);
    if (small.size() < 2) {
        rest.insert(rest.end(), small.begin(), small.end());
        small.clear();
    }
    bool ok = small_finish_launch(hts, outs, rows, small) && finish_launch_spread(hts, outs, rows, tmp, rest);
    ok = sync("agg finalize") && ok;
    for (int64_t i = 0; i < n; ++i) {
        HtBlob* B = (HtBlob*)hts[i];
        if (ok && rows[i] >= 0 && !table_finish_done(B->table, &outs[i], &tmp[(size_t)i])) ok = false;
        tmp[(size_t)i].release();
    }
    if (!ok) {
        for (int64_t i = 0; i < n; ++i) {
            destroy_kernel_rowset(&outs[i]);
            memset(&outs[i], 0, sizeof(TRowSet));
        }
        return -1;
    }
    for (int64_t i = 0; i < n; ++i) {
        HtBlob* B = (HtBlob*)hts[i];
        table_free(B->table);
        memset(B, 0, sizeof(*B));
    }
    return 0;
}

namespace {

// The projection that turns one input row into the partial states of its
// group (the pass-through of the reference's adaptive partial aggregate,
// docs/arch/aggregation.md:376-389): the keys as they are, 1 for COUNT(*),
// IS NOT NULL for COUNT(x), and the argument of SUM/MIN/MAX stored in the
// state's type. False (with the error text) when the aggregate has fused
// stages, which a projection of its input cannot reproduce.
bool partial_spec_of(const AggPlan* P, QdbProjectSpec* ps)
{
    const QdbAggregateSpec& a = P->spec;
    if (a.HasPredicate || a.NumComputed > 0) return fail("partial states: the aggregate has fused stages");
    if (P->nout > QDB_PROJECT_MAX_OUTPUTS) {
        return fail("partial states: %d output columns exceed the projection's %d", P->nout, QDB_PROJECT_MAX_OUTPUTS);
    }
    memset(ps, 0, sizeof(*ps));
    ps->Input = a.Input;
    ps->NumOutputs = P->nout;
    auto col = [](QdbExprProgram& p, int c, int cls) {
        QdbExprInsn& in = p.Insns[p.NumInsns++];
        in.Op = QDB_EXPR_COLUMN;
        in.Dst = 0;
        in.Class = (uint8_t)cls;
        in.Column = c;
        p.Result = 0;
    };
    int k = 0;
    for (int key = 0; key < a.NumKeys; ++key, ++k) {
        const int c = a.KeyColumns[key];
        col(ps->Outputs[k], c, a.Input.Types[c].Kind == QDB_KIND_FLOAT ? QDB_REG_F64 : QDB_REG_INT);
        ps->OutputTypes[k] = P->out_types[k];
    }
    for (int g = 0; g < a.NumAggs; ++g, ++k) {
        const QdbAggregateFn& fn = a.Aggs[g];
        QdbExprProgram& p = ps->Outputs[k];
        ps->OutputTypes[k] = P->out_types[k];
        if (fn.Func == QDB_AGG_COUNT_STAR || (fn.Func == QDB_AGG_COUNT && !a.Input.Types[fn.Column].Nullable)) {
            QdbExprInsn& in = p.Insns[p.NumInsns++];
            in.Op = QDB_EXPR_CONST_INT;
            in.Class = QDB_REG_INT;
            in.ImmLo = 1;
            p.Result = 0;
        } else if (fn.Func == QDB_AGG_COUNT) {
            col(p, fn.Column, a.Input.Types[fn.Column].Kind == QDB_KIND_FLOAT ? QDB_REG_F64 : QDB_REG_INT);
            QdbExprInsn& isnull = p.Insns[p.NumInsns++];
            isnull.Op = QDB_EXPR_IS_NULL;
            isnull.Dst = 1;
            isnull.A = 0;
            isnull.Class = QDB_REG_BOOL;
            QdbExprInsn& nt = p.Insns[p.NumInsns++];
            nt.Op = QDB_EXPR_NOT;
            nt.Dst = 2;
            nt.A = 1;
            nt.Class = QDB_REG_BOOL;
            p.Result = 2;
        } else {
            col(p, fn.Column, a.Input.Types[fn.Column].Kind == QDB_KIND_FLOAT ? QDB_REG_F64 : QDB_REG_INT);
        }
    }
    return true;
}

}  // namespace

extern "C" int qdb_gpu_agg_partial_states(const QdbAggregateKernels* kernels, const TRowSet* batch, TRowSet* out)
{
    clear_error();
    if (!kernels || !batch || !out) return fail("null argument"), -1;
    memset(out, 0, sizeof(*out));
    if (kernels->Slot < 0 || kernels->Slot >= kSlots) return fail("bad aggregate slot %lld", (long long)kernels->Slot), -1;
    AggPlan* P = agg_slots().get((int)kernels->Slot);
    if (!P) return fail("aggregate kernels are not compiled"), -1;
    {
        std::lock_guard<std::mutex> lock(P->partial_mu);
        if (!P->partial_tried) {
            P->partial_tried = true;
            QdbProjectSpec ps;
            if (partial_spec_of(P, &ps)) {
                QdbProjectFn fn = nullptr;
                const int slot = qdb_gpu_compile_project(&ps, &fn);
                if (slot >= 0) {
                    P->partial_fn = fn;
                    P->partial_slot = slot;
                }
            }
        }
        if (!P->partial_fn) return fail("partial states: no projection for this aggregate"), -1;
    }
    const int64_t n = batch->RowCount;
    OwnedRowSet* o = init_owned_rowset(out, n, P->nout);
    std::vector<void*> bufs((size_t)P->nout * 2, nullptr);
    const size_t mask_bytes = (size_t)((n + 7) / 8) + 8;
    bool ok = true;
    for (int c = 0; ok && c < P->nout; ++c) {
        const QdbValueType& t = P->out_types[c];
        const size_t bytes = t.Kind == QDB_KIND_BOOL ? mask_bytes : (size_t)(n > 0 ? n : 1) * (size_t)t.Width + 8;
        char* data = (char*)dalloc(bytes);
        ok = data != nullptr;
        if (!ok) break;
        o->dev.push_back(data);
        o->cols[(size_t)c].Data = data;
        bufs[(size_t)c] = data;
        if (t.Nullable) {
            uint8_t* m = (uint8_t*)dalloc(mask_bytes);
            ok = m != nullptr && zero_fill(m, mask_bytes);
            if (!ok) break;
            o->dev.push_back(m);
            o->cols[(size_t)c].Mask = m;
            bufs[(size_t)(P->nout + c)] = m;
        }
    }
    // The states of the selected rows; the batch's selection comes along.
    if (ok && batch->Selection && n > 0) {
        uint8_t* sel = (uint8_t*)dalloc((size_t)n);
        ok = sel && hip_ok(hipMemcpyAsync(sel, batch->Selection, (size_t)n, hipMemcpyDeviceToDevice, stream()), "selection copy");
        if (sel) o->dev.push_back(sel);
        out->Selection = sel;
    }
    if (ok && n > 0) {
        TRowSet view = *batch;
        P->partial_fn(&view, bufs.data(), nullptr, nullptr, batch->Selection);
        ok = last_error_text().empty();
    }
    ok = ok && sync("partial states");
    if (!ok || !publish_rowset(out)) {
        destroy_kernel_rowset(out);
        memset(out, 0, sizeof(*out));
        return -1;
    }
    return 0;
}

extern "C" int64_t qdb_gpu_aggregate_kernel_source(const QdbAggregateSpec* spec, char* out, int64_t cap)
{
    clear_error();
    if (!spec) return fail("null argument"), -1;
    AggPlan* P = new AggPlan();
    if (!compile_plan(*spec, P, false)) {
        delete P;
        return -1;
    }
    const std::string src = agg_jit_source(P);
    delete P;
    if (out && cap > 0) {
        const size_t n = src.size() < (size_t)cap - 1 ? src.size() : (size_t)cap - 1;
        memcpy(out, src.data(), n);
        out[n] = '\0';
    }
    return (int64_t)src.size();
}

extern "C" int qdb_gpu_jit_check(const char* src, const char* arch, char* err, int64_t cap)
{
    clear_error();
    if (!src || !arch) return fail("null argument"), -1;
    std::string e;
    const bool ok = jit_compile_only(src, arch, &e);
    if (err && cap > 0) {
        const size_t n = e.size() < (size_t)cap - 1 ? e.size() : (size_t)cap - 1;
        memcpy(err, e.data(), n);
        err[n] = '\0';
    }
    return ok ? 0 : -1;
}

extern "C" void qdb_gpu_aggregate_destroy(void* ht)
{
    HtBlob* B = (HtBlob*)ht;
    if (!B) return;
    table_free(B->table);
    memset(B, 0, sizeof(*B));
}

namespace qdbgpu {
void release_aggregate_slot(int slot)
{
    agg_slots().release(slot);
}
}  // namespace qdbgpu

namespace qdbgpu {

// The filter kernel generated for one predicate (the reference's compiler
// generates one per filter, qdb/kernel/gen.cpp:911-1091): the program as
// straight-line code over columns of the given slot types; a row is kept
// when its input selection keeps it and the predicate is TRUE (NULL is not
// TRUE).
std::string filter_jit_source(const DevProgram& prog, const QdbValueType* slot_type)
{
    std::string o;
    o += "#define QDB_SPECIALIZED 1\n#include \"aggregate_kernels.h\"\n";
    o += "namespace qdbgpu {\nnamespace agg {\n";
    o += "__device__ __forceinline__ DVal dval_zero()\n{\n    DVal v;\n    v.i = 0;\n    v.f = 0.0;\n    v.cls = 0;\n    v.valid = 0;\n    return v;\n}\n";
    o += "}  // namespace agg\n}  // namespace qdbgpu\n";
    o += "using namespace qdbgpu;\nusing namespace qdbgpu::agg;\n";
    o += "extern \"C\" __global__ void __launch_bounds__(256) qdb_filter(DevCols C, int64_t n, const uint8_t* in_sel, uint8_t* out_sel)\n{\n";
    o += "    const int64_t stride = (int64_t)gridDim.x * blockDim.x;\n";
    o += "    for (int64_t row = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; row < n; row += stride) {\n";
    o += "        uint8_t keep = 0;\n        if (!in_sel || in_sel[row]) {\n";
    const std::string r = emit_program_typed(o, "p", prog, slot_type);
    o += "            keep = (" + r + ".valid && " + r + ".i != 0) ? 1 : 0;\n        }\n";
    o += "        out_sel[row] = keep;\n    }\n}\n";
    return o;
}

// The project kernel generated for one projection (the reference's compiler
// generates one per project, qdb/kernel/gen.cpp:1373-1613): each output's
// program as straight-line code and its store specialized to the output's
// kind and width. Waves cover 64 consecutive rows from a multiple of 64, so
// that one lane writes each word of a BOOL column or a validity bitmap.
// Rows the input selection drops keep their previous output values.
std::string project_jit_source(const DevProgram* progs, int nout, const QdbValueType* slot_type, const QdbValueType* out_types)
{
    std::string o;
    o += "#define QDB_SPECIALIZED 1\n#include \"aggregate_kernels.h\"\n";
    o += "namespace qdbgpu {\nnamespace agg {\n";
    o += "__device__ __forceinline__ DVal dval_zero()\n{\n    DVal v;\n    v.i = 0;\n    v.f = 0.0;\n    v.cls = 0;\n    v.valid = 0;\n    return v;\n}\n";
    o += "}  // namespace agg\n}  // namespace qdbgpu\n";
    o += "using namespace qdbgpu;\nusing namespace qdbgpu::agg;\n";
    o += "extern \"C\" __global__ void __launch_bounds__(256) qdb_project(DevCols C, ProjectOut O, int64_t n, const uint8_t* in_sel)\n{\n";
    o += "    const int lane = lane_id();\n";
    o += "    const int64_t stride = (int64_t)gridDim.x * blockDim.x;\n";
    o += "    for (int64_t base = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) - lane; base < n; base += stride) {\n";
    o += "        const int64_t row = base + lane;\n";
    o += "        const bool eval = row < n && (!in_sel || in_sel[row]);\n";
    o += "        const uint64_t evaluated = __ballot(eval);\n";
    for (int k = 0; k < nout; ++k) {
        const std::string ks = std::to_string(k);
        const QdbValueType& t = out_types[k];
        o += "        {\n            DVal v = dval_zero();\n            if (eval) {\n";
        const std::string r = emit_program_typed(o, "o" + ks, progs[k], slot_type);
        o += "                v = " + r + ";\n            }\n";
        if (t.Kind == QDB_KIND_BOOL) {
            o += "            {\n                const uint64_t bits = __ballot(eval && v.valid && v.i != 0);\n";
            o += "                if (lane == 0) store_bits((uint8_t*)O.data[" + ks + "], base, n, bits, evaluated);\n            }\n";
        } else if (t.Kind == QDB_KIND_FLOAT) {
            o += "            if (eval) {\n                double f = v.cls == QDB_REG_F64 ? v.f : (double)v.i;\n";
            o += "                if (!v.valid) f = 0.0;\n";
            if (t.Width == 4) o += "                ((float*)O.data[" + ks + "])[row] = (float)f;\n";
            else o += "                ((double*)O.data[" + ks + "])[row] = f;\n";
            o += "            }\n";
        } else {
            o += "            if (eval) {\n                u128 x = v.valid ? (u128)v.i : 0;\n";
            if (t.Kind == QDB_KIND_DECIMAL && t.Width == 8) {
                o += "                if (v.valid && narrow_overflows(v.i, O.prec[" + ks + "])) {\n";
                o += "                    x = 0;\n                    atomicOr(O.err, 1u << " + ks + ");\n                }\n";
            }
            switch (t.Width) {
            case 1: o += "                ((uint8_t*)O.data[" + ks + "])[row] = (uint8_t)x;\n"; break;
            case 2: o += "                ((uint16_t*)O.data[" + ks + "])[row] = (uint16_t)x;\n"; break;
            case 4: o += "                ((uint32_t*)O.data[" + ks + "])[row] = (uint32_t)x;\n"; break;
            case 8: o += "                ((uint64_t*)O.data[" + ks + "])[row] = (uint64_t)x;\n"; break;
            default:
                o += "                uint64_t* p = (uint64_t*)(O.data[" + ks + "] + 16 * row);\n";
                o += "                p[0] = (uint64_t)x;\n                p[1] = (uint64_t)(x >> 64);\n";
            }
            o += "            }\n";
        }
        if (t.Nullable) {
            o += "            {\n                const uint64_t bits = __ballot(eval && v.valid);\n";
            o += "                if (lane == 0) store_bits(O.mask[" + ks + "], base, n, bits, evaluated);\n            }\n";
        }
        o += "        }\n";
    }
    o += "    }\n}\n";
    return o;
}

}  // namespace qdbgpu

namespace qdbgpu {
namespace agg {

void launch_update_kw1(int S, const AggArgs& A, int grid, size_t lds);
void launch_update_kw2(int S, const AggArgs& A, int grid, size_t lds);
void launch_update_kw3(int S, const AggArgs& A, int grid, size_t lds);
void launch_update_kw4(int S, const AggArgs& A, int grid, size_t lds);
void launch_batch_kw1(int S, const AggArgs* insts, const BatchItem* items, int nitems, unsigned int* head, int grid, size_t lds);
void launch_batch_kw2(int S, const AggArgs* insts, const BatchItem* items, int nitems, unsigned int* head, int grid, size_t lds);
void launch_batch_kw3(int S, const AggArgs* insts, const BatchItem* items, int nitems, unsigned int* head, int grid, size_t lds);
void launch_batch_kw4(int S, const AggArgs* insts, const BatchItem* items, int nitems, unsigned int* head, int grid, size_t lds);
void launch_records_kw1(int S, const AggArgs& A, const uint64_t* recs, unsigned int n);
void launch_records_kw2(int S, const AggArgs& A, const uint64_t* recs, unsigned int n);
void launch_records_kw3(int S, const AggArgs& A, const uint64_t* recs, unsigned int n);
void launch_records_kw4(int S, const AggArgs& A, const uint64_t* recs, unsigned int n);
void launch_rehash_kw1(const AggArgs& A, const uint64_t* old_recs, const unsigned long long* old_ctrl, uint64_t old_slots);
void launch_rehash_kw2(const AggArgs& A, const uint64_t* old_recs, const unsigned long long* old_ctrl, uint64_t old_slots);
void launch_rehash_kw3(const AggArgs& A, const uint64_t* old_recs, const unsigned long long* old_ctrl, uint64_t old_slots);
void launch_rehash_kw4(const AggArgs& A, const uint64_t* old_recs, const unsigned long long* old_ctrl, uint64_t old_slots);

void launch_update(int KW, int S, const AggArgs& A, int grid, size_t lds)
{
    switch (KW) {
    case 1: launch_update_kw1(S, A, grid, lds); break;
    case 2: launch_update_kw2(S, A, grid, lds); break;
    case 3: launch_update_kw3(S, A, grid, lds); break;
    default: launch_update_kw4(S, A, grid, lds); break;
    }
}

void launch_batch(int KW, int S, const AggArgs* insts, const BatchItem* items, int nitems, unsigned int* head, int grid, size_t lds)
{
    switch (KW) {
    case 1: launch_batch_kw1(S, insts, items, nitems, head, grid, lds); break;
    case 2: launch_batch_kw2(S, insts, items, nitems, head, grid, lds); break;
    case 3: launch_batch_kw3(S, insts, items, nitems, head, grid, lds); break;
    default: launch_batch_kw4(S, insts, items, nitems, head, grid, lds); break;
    }
}

void launch_records(int KW, int S, const AggArgs& A, const uint64_t* recs, unsigned int n)
{
    switch (KW) {
    case 1: launch_records_kw1(S, A, recs, n); break;
    case 2: launch_records_kw2(S, A, recs, n); break;
    case 3: launch_records_kw3(S, A, recs, n); break;
    default: launch_records_kw4(S, A, recs, n); break;
    }
}

void launch_rehash(int KW, const AggArgs& A, const uint64_t* old_recs, const unsigned long long* old_ctrl, uint64_t old_slots)
{
    switch (KW) {
    case 1: launch_rehash_kw1(A, old_recs, old_ctrl, old_slots); break;
    case 2: launch_rehash_kw2(A, old_recs, old_ctrl, old_slots); break;
    case 3: launch_rehash_kw3(A, old_recs, old_ctrl, old_slots); break;
    default: launch_rehash_kw4(A, old_recs, old_ctrl, old_slots); break;
    }
}

}  // namespace agg
}  // namespace qdbgpu
