// nq_kernels.h -- launch wrappers of the gfx950 kernels (nq_kernels.hip)
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace nqk {

constexpr int kStackCap = 16;  // private-memory DFS stack entries (>= T-1, T <= 16)
constexpr uint32_t kSchedU4 = 2;   // uint4 words of a record's schedule (nqt::sched_words)
// A work item, 3 x uint4: {ld, rd, free columns, frame word}, {record, T | wshift << 8 |
// N << 16, terminal mask, 0}, and the record's parameter bytes (schedule words 0-3).
constexpr uint32_t kItemU4 = 3;

// LDS of a walking workgroup: the DFS stacks, `slots` entries of 16 bytes per
// lane: the inert frame and the T-1-d0 frames a walk from depth d0 pushes at
// most.
struct WalkLds {
    uint32_t slots;
    size_t bytes(int block) const { return (size_t)block * 16u * slots; }
};
inline WalkLds walk_lds(uint32_t Tmax, uint32_t d0)
{
    return WalkLds{1u + (Tmax > d0 + 1u ? Tmax - 1u - d0 : 1u)};
}

// Depth at which the expansion starts a record's work items for split depth
// d: d itself when 0 < d <= T-2, else the record's root (k_expand makes the
// root the single item when d == 0 or d > T-2).  The walk stacks of a set are
// sized by walk_lds(max T, min over its records of this depth).
inline uint32_t item_depth(uint32_t d, uint32_t T)
{
    return (d > 0u && T >= 2u && d <= T - 2u) ? d : 0u;
}

hipError_t launch_decode(hipStream_t s, int N, const uint8_t *Ns, const uint4 *rec, uint32_t m, uint4 *root,
                         uint4 *sched, uint32_t *err);
hipError_t launch_plain(hipStream_t s, bool nodes, int block, WalkLds lay, const uint4 *root,
                        const uint4 *sched, uint32_t m, unsigned long long *weighted,
                        unsigned long long *nodes_out, unsigned long long *total);
// one lane per (record position, root candidate < fan): nitems (by record)
// and off (by position) hold m * fan entries (fan >= the set's largest N,
// m * fan < 2^32)
hipError_t launch_expand(hipStream_t s, bool nodes, bool write, int block, uint32_t stack_depth,
                         const uint4 *root, const uint4 *sched, const uint32_t *order, const uint8_t *splits,
                         uint32_t m, uint32_t split, uint32_t fan,
                         uint32_t *nitems, const uint32_t *off, uint4 *items,
                         unsigned long long *raw, unsigned long long *nodes_out);
// exclusive scan of nitems[0..m) into off (u32) and the total into *count;
// tsum: scratch of (m + 1023) / 1024 u64 entries.  With an order (records
// of fan entries each), entry a is nitems[order[a / fan] * fan + a % fan].
hipError_t launch_scan(hipStream_t s, const uint32_t *nitems, uint32_t m, uint32_t *off, unsigned long long *count,
                       unsigned long long *tsum, const uint32_t *order = nullptr, uint32_t fan = 0);
hipError_t launch_preset(hipStream_t s, bool write, int N, const uint4 *in, uint32_t m, uint32_t *cnt,
                         const uint32_t *off, uint4 *out, uint32_t *err);
hipError_t launch_batched(hipStream_t s, bool nodes, bool lds_stack, int grid, int block, WalkLds lay,
                          const uint4 *items,
                          const unsigned long long *n_items, unsigned int *head, uint32_t chunk,
                          unsigned long long *raw, unsigned long long *nodes_out);
hipError_t batched_occupancy(bool nodes, bool lds_stack, int block, size_t lds_bytes, int *blocks_per_cu);
hipError_t launch_finalize(hipStream_t s, const uint4 *sched, uint32_t m, const unsigned long long *raw,
                           unsigned long long *weighted, unsigned long long *total);
hipError_t launch_fill_u64(hipStream_t s, unsigned long long *p, uint64_t n, unsigned long long v);
hipError_t launch_fill_u32(hipStream_t s, unsigned int *p, uint64_t n, unsigned int v);
hipError_t launch_symmetric(hipStream_t s, int N, unsigned long long *out2);
// *out += every solution of an N x N board, 1 <= N <= 12 (plain DFS)
hipError_t launch_small_total(hipStream_t s, int N, unsigned long long *out);

// cost model (nq_mlp.hip)
hipError_t launch_features(hipStream_t s, const uint4 *rec, const uint4 *root, const uint4 *sched,
                           uint32_t m, float *feat);
hipError_t launch_mlp_forward(hipStream_t s, const float *params, const float *x, uint32_t m, float *y);

}  // namespace nqk
