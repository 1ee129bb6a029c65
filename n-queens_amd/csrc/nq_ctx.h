// nq_ctx.h -- the context object behind the C ABI and the helpers shared by
// nq_host.cpp (count path) and nq_mlp.hip (cost model).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <algorithm>
#include <vector>

#include "../../include/nq.h"
#include "nq_kernels.h"

// Records the message for nq_last_error() and returns `code`.
int nq_fail(int code, const char *fmt, ...);

#define NQ_HIPCHK(expr)                                                                    \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess)                                                              \
            return nq_fail(NQ_EHIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                           __FILE__, __LINE__);                                            \
    } while (0)

namespace nqk_buf {

// Device buffer owned by a context or a call; freed on destruction.
template <class T>
struct Buf {
    T *p = nullptr;
    size_t n = 0;
    Buf() = default;
    Buf(const Buf &) = delete;
    Buf &operator=(const Buf &) = delete;
    ~Buf() { release(); }
    hipError_t reserve(size_t want)
    {
        if (want <= n && p) return hipSuccess;
        release();
        const size_t cap = want ? want : 1;
        hipError_t e = hipMalloc((void **)&p, cap * sizeof(T));
        if (e == hipSuccess) n = cap;
        else p = nullptr;
        return e;
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
};

}  // namespace nqk_buf

struct nq_graph_key {
    int split = -1, block = 0, lds = 0, chunk = 0, grid = 0, nodes = 0;
    bool operator==(const nq_graph_key &o) const
    {
        return split == o.split && block == o.block && lds == o.lds && chunk == o.chunk && grid == o.grid &&
               nodes == o.nodes;
    }
};

struct nq_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    int cu = 0, lds_max = 0, wave = 64, clock_khz = 0;
    // resident task set
    int N = 0;                    // board size of the set, 0 for a mixed set
    uint64_t m = 0;
    uint32_t Tmax = 0, Tmin = 0;
    uint32_t fan = 1;             // expansion lanes per record: the set's largest N
    bool loaded = false;
    nqk_buf::Buf<uint4> rec, root, sched, items;
    nqk_buf::Buf<uint32_t> order, nitems, off, err, head;
    nqk_buf::Buf<uint8_t> nsz;    // board size per record (mixed sets)
    nqk_buf::Buf<uint8_t> splits; // split depth per record (nq_ctx_set_splits)
    bool has_splits = false;
    uint32_t split_min = 0;       // shallowest depth a per-record split starts an item at
    std::vector<uint8_t> Trec;    // terminal depth per record (host copy, from the decode)
    nqk_buf::Buf<unsigned long long> weighted, nodes, raw, total, item_count, sym;
    nqk_buf::Buf<unsigned long long> scan_tmp;   // tile sums of the item-offset scan, (m * fan + 1023) / 1024
    // cost-model ordering on the device (nq_ctx_order_by_cost)
    nqk_buf::Buf<float> cm_params, cm_x, cm_y, cm_ys;
    nqk_buf::Buf<uint32_t> cm_idx;
    nqk_buf::Buf<uint8_t> cm_tmp;
    int sized_split = -2;         // split depth the item buffer is sized for (-1 per-record, -2 none)
    uint64_t item_total = 0;
    void *scratch = nullptr;      // this device's static scratch array (nq_count_noalloc)
    // captured batched sequence
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    nq_graph_key gkey;
};

// Validates the handle and makes its device current.
int nq_check_ctx(nq_ctx *c);

// Device buffers and shape of one resident set as the runs read them: a
// context's own buffers, or views into the static scratch array
// (nq_count_noalloc).
struct nq_run {
    uint32_t m = 0, Tmax = 0, Tmin = 0;
    uint32_t fan = 1;             // expansion lanes per record (nitems / off hold m * fan entries)
    uint32_t split_min = 0;       // shallowest item start depth of the per-record splits
    const uint4 *root = nullptr, *sched = nullptr;
    const uint32_t *order = nullptr;   // processing order, or NULL for record order
    const uint8_t *splits = nullptr;   // per-record split depths, or NULL
    uint32_t *nitems = nullptr, *off = nullptr, *head = nullptr;
    uint4 *items = nullptr;
    unsigned long long *raw = nullptr, *nodes = nullptr, *weighted = nullptr, *total = nullptr;
    unsigned long long *item_count = nullptr, *scan_tmp = nullptr;
};

// Checks and completes a batched schedule for the set (block size, claim
// chunk, split depth clamped to T-2, LDS stack within lds_max).
int nq_prepare_sched(const nq_run &r, int lds_max, nq_sched *s);
// LDS layout of the batched walk for split depth `split` (-1: per record).
nqk::WalkLds nq_batched_lds(const nq_run &r, int split);
// Persistent grid size (occupancy, waves_per_eu cap) and the u32 queue-head bound.
int nq_batched_grid(int cu, const nq_sched &s, size_t lds_bytes, int nodes, uint64_t items, int *grid);
// The batched launch sequence on stream st: clear, expand (count, scan,
// write), persistent DFS, finalize.  ev (4 events, may be NULL) brackets the
// sequence (0, 3) and the DFS kernel (1, 2).
int nq_enqueue_batched(hipStream_t st, const nq_run &r, const nq_sched &s, int nodes, int grid,
                       hipEvent_t *ev);
// Expansion count pass + scan of the set at split depth `split`, with the item
// total read back into *count.
int nq_count_items(hipStream_t st, const nq_run &r, int split, int block, uint64_t *count);

// The static scratch array of this device (nq_noalloc.hip): its address,
// looked up once per device by nq_ctx_create.
int nq_scratch_attach(int device, void **ptr);

#define NQ_CTX_CHECK(c)                       \
    do {                                      \
        if (int r_ = nq_check_ctx(c)) return r_; \
    } while (0)
