// nq_ctx.h -- the context object behind the C ABI and the helpers shared by
// nq_host.cpp (count path) and nq_mlp.hip (cost model).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <algorithm>
#include <vector>

#include "../../include/nq.h"

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
    bool loaded = false;
    nqk_buf::Buf<uint4> rec, root, sched, items;
    nqk_buf::Buf<uint32_t> order, nitems, off, err, head;
    nqk_buf::Buf<uint8_t> nsz;    // board size per record (mixed sets)
    nqk_buf::Buf<uint8_t> splits; // split depth per record (nq_ctx_set_splits)
    bool has_splits = false;
    uint32_t split_min = 0;       // shallowest depth a per-record split starts an item at
    std::vector<uint8_t> Trec;    // terminal depth per record (host copy, from the decode)
    nqk_buf::Buf<unsigned long long> weighted, nodes, raw, total, item_count, sym;
    nqk_buf::Buf<unsigned long long> scan_tmp;   // tile sums of the item-offset scan, (m + 1023) / 1024
    // cost-model ordering on the device (nq_ctx_order_by_cost)
    nqk_buf::Buf<float> cm_params, cm_x, cm_y, cm_ys;
    nqk_buf::Buf<uint32_t> cm_idx;
    nqk_buf::Buf<uint8_t> cm_tmp;
    int sized_split = -2;         // split depth the item buffer is sized for (-1 per-record, -2 none)
    uint64_t item_total = 0;
    // captured batched sequence
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    nq_graph_key gkey;
};

// Validates the handle and makes its device current.
int nq_check_ctx(nq_ctx *c);

#define NQ_CTX_CHECK(c)                       \
    do {                                      \
        if (int r_ = nq_check_ctx(c)) return r_; \
    } while (0)
