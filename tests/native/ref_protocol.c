/*
 * ref_protocol.c - TEST INFRASTRUCTURE: the reference's side of the aggregate
 * kernel boundary, restated in C so that the HIP library's entry points can
 * be driven exactly as the reference's unchanged callers drive them.
 *
 *   task heap   blocks carry an owner header; qdb_alloc links a block into the
 *               current task's heap, qdb_free unlinks and frees it, Detach
 *               moves it out of the heap (TKernelHeap,
 *               qdb/kernel/execution_context.cpp:20-120; qdb_alloc/qdb_free,
 *               qdb/modules/qumirdb_runtime.cpp:212-228).
 *   ownership   AdoptKernelOwnedRowSet detaches every owner listed in Private
 *               (owners[0] = count) and the list itself, then installs
 *               DestroyKernelOwnedRowSet, which qdb_free's them
 *               (qdb/exec/kernel_rowset.h:14-38).
 *   processor   TAggregateProcessor: Init with a power-of-two capacity >= 8,
 *               Update per batch (< 0 fails), Size read from the HashTable
 *               blob (offset 72) before agg_finish_rowset, and the finalized
 *               row count must equal it (qdb/exec/aggregate_exec.cpp:67-136).
 *
 * Every block header holds a magic word: Detach and qdb_free of a pointer that
 * did not come from qdb_alloc (a device pointer, a malloc'd block) is counted
 * as a violation instead of being followed.
 */
#include "../../include/qdb_gpu.h"

#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define BLOCK_MAGIC 0x48454150424c4b31ull

typedef struct Block {
    uint64_t magic;
    struct Heap* owner;
    struct Block* prev;
    struct Block* next;
    int64_t size;
    int64_t pad;
} Block;

typedef struct Heap {
    Block* head;
    int64_t linked;
} Heap;

static Heap g_task;           /* the current task's heap */
static int64_t g_outstanding; /* blocks allocated and not yet freed */
static int64_t g_violations;  /* frees / detaches of foreign pointers */

static Block* header(void* p)
{
    Block* b = (Block*)p - 1;
    if (b->magic != BLOCK_MAGIC) {
        ++g_violations;
        return NULL;
    }
    return b;
}

static void unlink_block(Block* b)
{
    if (b->prev) b->prev->next = b->next;
    else b->owner->head = b->next;
    if (b->next) b->next->prev = b->prev;
    b->owner->linked -= 1;
    b->owner = NULL;
    b->prev = b->next = NULL;
}

static void* h_alloc(int64_t size)
{
    if (size < 0) return NULL;
    Block* b = (Block*)calloc(1, sizeof(Block) + (size_t)(size ? size : 1));
    if (!b) return NULL;
    b->magic = BLOCK_MAGIC;
    b->owner = &g_task;
    b->next = g_task.head;
    if (g_task.head) g_task.head->prev = b;
    g_task.head = b;
    g_task.linked += 1;
    b->size = size;
    g_outstanding += 1;
    return b + 1;
}

static void h_free(void* p)
{
    if (!p) return;
    Block* b = header(p);
    if (!b) return;
    if (b->owner) unlink_block(b);
    b->magic = 0;
    free(b);
    g_outstanding -= 1;
}

static void h_detach(void* p)
{
    if (!p) return;
    Block* b = header(p);
    if (b && b->owner) unlink_block(b);
}

/* DestroyKernelOwnedRowSet. */
static void destroy_kernel_owned(TRowSet* rs)
{
    int64_t* owners = (int64_t*)rs->Private;
    if (!owners) return;
    const int64_t count = owners[0];
    for (int64_t i = 0; i < count; ++i) h_free((void*)(intptr_t)owners[i + 1]);
    h_free(owners);
    rs->Private = NULL;
}

/* AdoptKernelOwnedRowSet. */
static void adopt_kernel_owned(TRowSet* rs)
{
    int64_t* owners = (int64_t*)rs->Private;
    if (owners) {
        for (int64_t i = 0; i < owners[0]; ++i) h_detach((void*)(intptr_t)owners[i + 1]);
        h_detach(owners);
    }
    rs->Destroy = destroy_kernel_owned;
}

typedef void (*SetHeapFn)(QdbHeapAllocFn, QdbHeapFreeFn, QdbHeapDetachFn);

/* Registers this heap with the library (its qdb_gpu_set_host_heap). */
void harness_install(SetHeapFn set_heap)
{
    set_heap(h_alloc, h_free, h_detach);
}

/* [outstanding blocks, blocks still linked in the task heap, violations]. */
void harness_counters(int64_t* out)
{
    out[0] = g_outstanding;
    out[1] = g_task.linked;
    out[2] = g_violations;
}

/* TAggregateProcessor over one compiled aggregate: returns the finalized row
 * count (out adopted), -1 when Init / Update / finish failed, -2 when the
 * finalized count differs from the table's Size. */
int64_t harness_aggregate(const QdbAggregateKernels* k, TRowSet* const* batches, int64_t n, int64_t capacity,
                          TRowSet* out)
{
    uint8_t blob[QDB_HASH_TABLE_SIZE];
    memset(blob, 0, sizeof(blob));
    int64_t cap = 8;
    while (cap < capacity) cap <<= 1;
    if (k->Dispatch(blob, NULL, cap, QDB_AGG_OP_INIT) == 0) return -1;
    for (int64_t i = 0; i < n; ++i) {
        if (k->Dispatch(blob, batches[i], 0, QDB_AGG_OP_UPDATE) < 0) {
            qdb_gpu_aggregate_destroy(blob);
            return -1;
        }
    }
    int64_t size;
    memcpy(&size, blob + 72, 8);
    memset(out, 0, sizeof(*out));
    const int64_t finalized = k->FinishRowSet(blob, out);
    if (finalized < 0) return -1;
    if (finalized != size) return -2;
    adopt_kernel_owned(out);
    return finalized;
}
