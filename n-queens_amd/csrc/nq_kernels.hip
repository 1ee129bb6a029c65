// nq_kernels.hip -- gfx950 kernels of libnq: record decode, the per-lane
// bitboard DFS (plain and persistent/batched forms), work-item expansion,
// the weighted-count reduction and the symmetric-solution enumerator.
//
// Work unit: one DFS node is one extraction bit = avail & -avail of the
// reference dfs_iter (13Bit_codon/18Py_constellations_GPU_cuda_codon.py:368-562).
// All masks are u32 (N <= 24); counts are u64.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "nq_task.h"
#include "nq_kernels.h"
#include "nq_dfs.h"

namespace nqk {

// ---------------------------------------------------------------------------
// wave64 helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

__device__ __forceinline__ uint64_t wave_ballot(bool p) { return __builtin_amdgcn_ballot_w64(p); }

// Steps between two looks at the wave's lanes: a lane whose walk ends sits
// on the inert frame and waits at most this many steps (for its next item,
// or for the wave's other lanes); the steps in between carry no wave-wide
// check.
constexpr int kStepsPerCheck = 64;

// Count + nodes of a whole constellation from its root frame (depth 0).  A
// lane whose walk has ended steps on the inert frame until every walk of its
// wave has ended.
template <bool NODES, class Stack>
__device__ __forceinline__ void run_root(const Sched &S, uint4 root, Stack &st, uint64_t &cnt, uint64_t &nodes)
{
    cnt = 0; nodes = 0;
    if (root.w == 0u) return;
    if (S.T == 0u) { cnt = (root.w & S.tm) ? 1u : 0u; return; }
    const LaneK L = lane_consts(S);
    const PairK K = pair_consts(st.S);
    Walk w{root.x, root.y, free_cols(S, root.z), start_word(S, root.w, 0u), st.empty()};
    do {
        uint32_t c32 = 0;
#pragma unroll
        for (int u = 0; u < kStepsPerCheck; u++) walk_step_paired<NODES>(L, K, w, st, c32, nodes);
        cnt += c32;
    } while (wave_ballot(w.W != 0u) != 0ull);
}

// ---------------------------------------------------------------------------
// decode: records -> root frames + schedule descriptors
// sched layout per record: nqt::sched_words
// ---------------------------------------------------------------------------
__global__ void k_decode(int N, const uint8_t *__restrict__ Ns, const uint4 *__restrict__ rec, uint32_t m,
                         uint4 *__restrict__ root, uint4 *__restrict__ sched, uint32_t *__restrict__ err)
{
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= m) return;
    const uint4 r = rec[c];
    const int n = Ns ? (int)Ns[c] : N;
    nqt::Desc D;
    nqt::decode(n, r.x, r.y, r.z, r.w, D);
    if (D.status != nqt::kOk) atomicOr(err, 1u << D.status);
    root[c] = make_uint4(D.ld, D.rd, D.col, D.avail);
    uint32_t w[8];
    nqt::sched_words(D, n, w);
    sched[kSchedU4 * c] = make_uint4(w[0], w[1], w[2], w[3]);
    sched[kSchedU4 * c + 1] = make_uint4(w[4], w[5], w[6], w[7]);
}

// ---------------------------------------------------------------------------
// plain kernel: one lane per constellation, LDS stack, u64 wave reduction
// ---------------------------------------------------------------------------
constexpr uint32_t kPlainBlock = 64;

template <bool NODES>
__global__ void __launch_bounds__(kPlainBlock)
k_plain(const uint4 *__restrict__ root, const uint4 *__restrict__ sched, uint32_t m,
        unsigned long long *__restrict__ weighted, unsigned long long *__restrict__ nodes_out,
        unsigned long long *__restrict__ total)
{
    extern __shared__ uint4 lds[];
    const uint32_t lb = lds_addr(lds);
    const uint32_t c = blockIdx.x * kPlainBlock + threadIdx.x;
    uint64_t wsum = 0;
    if (c < m) {
        const Sched S = load_sched(sched, c);
        const uint32_t wsh = (sched[kSchedU4 * c + 1].x >> 8) & 3u;
        LdsStack st{LdsMem{}, 16u * kPlainBlock, lb + 16u * (kPlainBlock + threadIdx.x)};
        st.init();
        uint64_t cnt, nn;
        run_root<NODES>(S, root[c], st, cnt, nn);
        wsum = cnt << wsh;
        weighted[c] = wsum;
        if (NODES) nodes_out[c] = nn;
    }
    wsum = wave_sum_u64(wsum);
    if (lane_id() == 0 && wsum) atomicAdd(total, (unsigned long long)wsum);
}

// ---------------------------------------------------------------------------
// expansion: frames at depth `split` (one depth for all records, or a depth
// per record) become work items
//   one lane per (processing position q, root candidate k < fan): the lane
//   walks the root frame restricted to its record's k-th candidate, so a
//   record's items come out in DFS order, k by k (fan 1: one lane walks all
//   of a record's candidates, with per-record arrays).  Pass 0 counts items per
//   lane into nitems[c * fan + k] (by record, so the counts hold for any
//   processing order), pass 1 writes them at off[q * fan + k]..; nodes above
//   the split depth are counted here.
// ---------------------------------------------------------------------------
template <bool NODES, bool WRITE>
__global__ void k_expand(const uint4 *__restrict__ root, const uint4 *__restrict__ sched,
                         const uint32_t *__restrict__ order, const uint8_t *__restrict__ splits, uint32_t m,
                         uint32_t split_all, uint32_t fan,
                         uint32_t *__restrict__ nitems, const uint32_t *__restrict__ off,
                         uint4 *__restrict__ items,
                         unsigned long long *__restrict__ raw, unsigned long long *__restrict__ nodes_out)
{
    extern __shared__ uint4 lds[];
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t q = g / fan, k = g - q * fan;
    if (q >= m) return;
    const uint32_t c = order ? order[q] : q;
    const Sched S = load_sched(sched, c);
    const uint4 sa = sched[kSchedU4 * c], si = sched[kSchedU4 * c + 1];
    const uint4 meta = make_uint4(c, si.x, si.y, 0u);
    const uint32_t split = splits ? (uint32_t)splits[c] : split_all;
    const uint4 r = root[c];
    SlotStack st{lds + threadIdx.x, (uint32_t)__builtin_ctz(blockDim.x)};
    uint64_t cnt = 0, nn = 0;
    uint32_t n = 0;
    uint4 *const out = WRITE ? items + (size_t)kItemU4 * off[g] : nullptr;
    auto put = [&](uint32_t a, uint32_t b, uint32_t cc, uint32_t w) {
        if (WRITE) {
            uint4 *it = out + (size_t)kItemU4 * n;
            it[0] = make_uint4(a, b, cc, w);
            it[1] = meta;
            it[2] = sa;
        }
        n++;
    };
    if (r.w != 0u) {
        if (S.T == 0u) {
            // a root on the terminal row: its one solution needs no walk
            if (WRITE && k == 0u) raw[c] = (r.w & S.tm) ? 1u : 0u;
        } else if (S.T < 2u || split == 0u || split > S.T - 2u) {
            // the root itself is the single item (T = 1 roots are resolved by the DFS kernel)
            if (k == 0u) put(r.x, r.y, free_cols(S, r.z), start_word(S, r.w, 0u));
        } else {
            uint32_t av = r.w;
            for (uint32_t i = 0; i < k && av; i++) av &= av - 1u;
            if (av) {
                const uint32_t upw = start_word(S, 0u, split);
                auto emit = [&](uint32_t a, uint32_t b, uint32_t cc, uint32_t f) {
                    put(a, b, free_cols(S, cc), f | upw);
                };
                // fan 1: the lane walks every candidate of its root
                expand_subtree<NODES>(S, r.x, r.y, r.z, fan == 1u ? av : av & (0u - av), 0u, split, st, cnt, nn,
                                      emit);
            }
        }
    }
    if (!WRITE) nitems[c * fan + k] = n;
    if (WRITE && NODES && nn) atomicAdd(nodes_out + c, (unsigned long long)nn);   // DFS nodes are added too
}

// ---------------------------------------------------------------------------
// preset expansion: each record of preset p becomes its records of preset
// p + 1 (nqt::preset_children); pass 0 counts them per record, pass 1 writes
// them at off[c].. in record order
// ---------------------------------------------------------------------------
template <bool WRITE>
__global__ void k_preset(int N, const uint4 *__restrict__ in, uint32_t m, uint32_t *__restrict__ cnt,
                         const uint32_t *__restrict__ off, uint4 *__restrict__ out, uint32_t *__restrict__ err)
{
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= m) return;
    const uint4 r = in[c];
    uint32_t n = 0;
    uint4 *const o = WRITE ? out + off[c] : nullptr;
    const int32_t st = nqt::preset_children(N, r.x, r.y, r.z, r.w, [&](uint32_t a, uint32_t b, uint32_t cc, uint32_t sij) {
        if (WRITE) o[n] = make_uint4(a, b, cc, sij);
        n++;
    });
    if (st != nqt::kOk) atomicOr(err, 1u << st);
    if (!WRITE) cnt[c] = n;
}

// exclusive scan of nitems[0..m) -> off (mod 2^32, as the item indices), total
// -> *count, in three launches: each 1024-entry tile scanned by one block
// (coalesced; u64 sums), the tile sums scanned by one block, the tile offsets
// added back.  With an order, entry a is read from nitems[order[a / fan] * fan
// + a % fan] (record-indexed counts scanned in processing order).
constexpr uint32_t kScanTile = 1024;

__device__ __forceinline__ uint64_t wave_incl_scan_u64(uint64_t x, uint32_t lane)
{
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint64_t y = __shfl_up(x, d, 64);
        if (lane >= (uint32_t)d) x += y;
    }
    return x;
}

__global__ void __launch_bounds__(1024) k_scan_tile(const uint32_t *__restrict__ nitems, uint32_t m,
                                                    const uint32_t *__restrict__ order, uint32_t fan,
                                                    uint32_t *__restrict__ off, unsigned long long *__restrict__ tsum)
{
    __shared__ uint64_t wsum[kScanTile / 64];
    const uint32_t t = threadIdx.x, lane = t & 63u, w = t >> 6;
    const uint64_t a = (uint64_t)blockIdx.x * kScanTile + t;
    uint64_t v = 0;
    if (a < m) {
        if (order) {
            const uint32_t q = (uint32_t)a / fan;
            v = nitems[(uint64_t)order[q] * fan + ((uint32_t)a - q * fan)];
        } else {
            v = nitems[a];
        }
    }
    const uint64_t x = wave_incl_scan_u64(v, lane);
    if (lane == 63u) wsum[w] = x;
    __syncthreads();
    uint64_t before = 0;
    for (uint32_t k = 0; k < w; k++) before += wsum[k];
    if (a < m) off[a] = (uint32_t)(before + x - v);
    if (t == kScanTile - 1u) tsum[blockIdx.x] = before + x;
}

// exclusive scan of the n tile sums in place (one block of 1024), total -> *count
__global__ void __launch_bounds__(1024) k_scan_tops(unsigned long long *__restrict__ tsum, uint32_t n,
                                                    unsigned long long *__restrict__ count)
{
    __shared__ uint64_t part[1024];
    const uint32_t t = threadIdx.x;
    const uint32_t per = (n + 1023u) / 1024u;
    const uint32_t lo = min(n, t * per), hi = min(n, lo + per);
    uint64_t s = 0;
    for (uint32_t a = lo; a < hi; a++) s += tsum[a];
    part[t] = s;
    __syncthreads();
    for (uint32_t o = 1; o < 1024u; o <<= 1) {
        const uint64_t v = t >= o ? part[t - o] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    uint64_t run = part[t] - s;
    for (uint32_t a = lo; a < hi; a++) {
        const uint64_t v = tsum[a];
        tsum[a] = run;
        run += v;
    }
    if (t == 1023u) *count = part[1023];
}

__global__ void k_scan_add(uint32_t *__restrict__ off, uint32_t m, const unsigned long long *__restrict__ tsum)
{
    const uint64_t a = (uint64_t)blockIdx.x * kScanTile + threadIdx.x;
    if (a < m) off[a] += (uint32_t)tsum[blockIdx.x];
}

// ---------------------------------------------------------------------------
// persistent batched kernel: every lane runs its own DFS and takes the next
// work item from the wave's claimed range as soon as its subtree is done;
// the wave claims `chunk` items at a time from the global queue head.
// ---------------------------------------------------------------------------
template <bool LDS_STACK> struct BatchedStack;
template <> struct BatchedStack<true> { using type = LdsStack; };
template <> struct BatchedStack<false> { using type = PrivStack; };

// A lane counts the solutions of its current item in a u32 and hands it to
// the record's u64 when the item ends or when the count reaches 2^31 (each
// step adds at most one, so it never wraps).
constexpr uint32_t kCountFlush = 1u << 31;


// Each lane walks one item and holds the next one, loaded while it walks:
// when its walk ends the lane switches to the held item without waiting for
// memory, then asks for a new one.  Items are handed out from a range the
// wave claims `chunk` at a time from the global queue head.
template <bool NODES, bool LDS_STACK, uint32_t BLOCK>
__global__ void __launch_bounds__(BLOCK)
k_batched(const uint4 *__restrict__ items, const unsigned long long *__restrict__ nitems_total,
          unsigned int *__restrict__ head, uint32_t chunk,
          unsigned long long *__restrict__ raw, unsigned long long *__restrict__ nodes_out)
{
    extern __shared__ uint4 lds[];
    const uint32_t lb = lds_addr(lds);
    const uint32_t n_items = (uint32_t)*nitems_total;
    const uint32_t lane = lane_id();
    typename BatchedStack<LDS_STACK>::type st;
    if constexpr (LDS_STACK) st = LdsStack{LdsMem{}, 16u * BLOCK, lb + 16u * (BLOCK + threadIdx.x)};
    st.init();
    LaneK L = idle_lane();
    const uint4 f0 = inert_frame();
    Walk w{f0.x, f0.y, f0.z, f0.w, st.empty()};
    uint32_t next = 0, end = 0;       // wave-uniform claimed range [next, end)
    bool exhausted = false;           // wave-uniform: queue drained
    bool have = false;                // this lane walks an item
    bool held = false;                // this lane holds the next item (p0..p2)
    bool pend = false;                // this lane's ended walk still has counts to hand in
    uint4 p0{}, p1{}, p2{};
    uint32_t c = 0, cnt = 0, fc = 0, fcnt = 0;
    uint64_t nn = 0, fnn = 0;
    const PairK K = pair_consts(st.S);
    for (;;) {
#pragma unroll
        for (int u = 0; u < kStepsPerCheck; u++) walk_step_paired<NODES>(L, K, w, st, cnt, nn);
        const bool inert = w.W == 0u;   // a walking frame always has candidates
        if (wave_ballot(inert || cnt >= kCountFlush) == 0ull) continue;
        if (have && inert) {
            // the walk ended; its counts go on with the held item when that
            // is of the same record, else they are handed in after the switch
            have = false;
            if (!held || p1.x != c) {
                fc = c; fcnt = cnt; fnn = nn; pend = true;
                cnt = 0; nn = 0;
            }
        }
        // switch to the held item (its loads were issued at an earlier check);
        // every item starts a walk (roots on the terminal row make no item)
        if (!have && held) {
            held = false;
            c = p1.x;
            L = lane_k((p1.y >> 16) & 31u, p2.x, p2.y, p2.z, p2.w);
            w = Walk{p0.x, p0.y, p0.z, p0.w, st.empty()};
            have = true;
        }
        // hand out the next items of the wave's range to the lanes without a held item
        while (!exhausted) {
            const uint64_t need = wave_ballot(!held);
            if (need == 0ull) break;
            if (next >= end) {
                uint32_t b = 0;
                if (lane == 0) b = atomicAdd(head, chunk);
                b = __builtin_amdgcn_readfirstlane(__shfl(b, 0, 64));
                if (b >= n_items) { exhausted = true; break; }
                next = b;
                end = min(b + chunk, n_items);
            }
            const uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(need >> 32),
                                      __builtin_amdgcn_mbcnt_lo((uint32_t)need, 0u));
            const uint32_t idx = next + rank;
            if (!held && idx < end) {
                const uint4 *it = items + (size_t)kItemU4 * idx;
                p0 = it[0]; p1 = it[1]; p2 = it[2];
                held = true;
            }
            next = min(next + (uint32_t)__builtin_popcountll(need), end);
        }
        // counts (after the switch, which waits for older memory operations)
        if (pend) {
            if (fcnt) atomicAdd(raw + fc, (unsigned long long)fcnt);
            if (NODES && fnn) atomicAdd(nodes_out + fc, (unsigned long long)fnn);
            pend = false;
        }
        if (have && cnt >= kCountFlush) { atomicAdd(raw + c, (unsigned long long)cnt); cnt = 0; }
        if (wave_ballot(have || held) == 0ull) break;   // reached only once the queue is drained
    }
}

// weighted[c] = raw[c] << wshift[c]; total = sum
__global__ void k_finalize(const uint4 *__restrict__ sched, uint32_t m, const unsigned long long *__restrict__ raw,
                           unsigned long long *__restrict__ weighted, unsigned long long *__restrict__ total)
{
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t v = 0;
    if (c < m) {
        v = raw[c] << ((sched[kSchedU4 * c + 1].x >> 8) & 3u);
        weighted[c] = v;
    }
    v = wave_sum_u64(v);
    if (lane_id() == 0 && v) atomicAdd(total, (unsigned long long)v);
}

// fill kernel (used instead of hipMemsetAsync so that the sequence can be
// captured in a graph)
// Every solution of a small board (N <= kSmallMaxN) by a plain bitboard DFS
// over all rows: thread c puts the first row's queen in column c.  The
// reference counts N <= 5 this way instead of through constellations
// (_bit_total, 18Py:2071-2084).
constexpr int kSmallMaxN = 12;

__global__ void k_small_total(int N, unsigned long long *__restrict__ out)
{
    const int c = threadIdx.x;
    if (c >= N) return;
    const uint32_t mask = (1u << N) - 1u;
    uint32_t sl[kSmallMaxN], sr[kSmallMaxN], sc[kSmallMaxN], sa[kSmallMaxN];
    uint64_t n = 0;
    if (N == 1) {
        n = 1;
    } else {
        const uint32_t q = 1u << c;
        int d = 1;                       // row d: free columns sa[d]; rows 1..N-1
        sl[1] = q << 1;
        sr[1] = q >> 1;
        sc[1] = q;
        sa[1] = mask & ~(sl[1] | sr[1] | sc[1]);
        while (d >= 1) {
            if (sa[d] == 0u) { d--; continue; }
            const uint32_t bit = sa[d] & (0u - sa[d]);
            sa[d] ^= bit;
            if (d == N - 1) { n++; continue; }
            const uint32_t l = (sl[d] | bit) << 1, r = (sr[d] | bit) >> 1, cc = sc[d] | bit;
            d++;
            sl[d] = l;
            sr[d] = r;
            sc[d] = cc;
            sa[d] = mask & ~(l | r | cc);
        }
    }
    if (n) atomicAdd(out, (unsigned long long)n);
}

__global__ void k_fill_u64(unsigned long long *__restrict__ p, uint64_t n, unsigned long long v)
{
    for (uint64_t a = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; a < n; a += (uint64_t)gridDim.x * blockDim.x)
        p[a] = v;
}
__global__ void k_fill_u32(unsigned int *__restrict__ p, uint64_t n, unsigned int v)
{
    for (uint64_t a = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; a < n; a += (uint64_t)gridDim.x * blockDim.x)
        p[a] = v;
}

// ---------------------------------------------------------------------------
// symmetric solutions: F180 (half-turn invariant) and F90 (quarter-turn)
// ---------------------------------------------------------------------------
// A half-turn invariant solution is fixed by its top N/2 rows; a queen at
// (r,c) implies one at (N-1-r, N-1-c).  The attacked columns/diagonals of a
// symmetric partial placement are themselves symmetric, so the mirrored
// queen is free exactly when (r,c) is; only the pair's own conflicts
// (same column c == N-1-c, same diagonal r == c, same anti-diagonal
// r + c == N-1) need a separate test.  Each lane owns one prefix of P rows.
// At a leaf the quarter-turn test q[q[r]] == N-1-r is applied.
__device__ __forceinline__ uint32_t row_free(int N, int r, uint32_t full, uint32_t cols, uint64_t dg, uint64_t an)
{
    uint32_t f = full & ~(cols | (uint32_t)(dg >> (N - 1 - r)) | (uint32_t)(an >> r));
    f &= ~(1u << r);                 // r == c: pair on one diagonal
    f &= ~(1u << (N - 1 - r));       // r + c == N-1: pair on one anti-diagonal
    if (N & 1) f &= ~(1u << (N / 2)); // c == N-1-c
    return f;
}

// Columns of the rows r < h = N/2 of a half-turn symmetric placement, 5 bits
// each (13 rows for N <= 27).
typedef unsigned __int128 qrows_t;

__device__ __forceinline__ bool quarter_turn(int N, int h, qrows_t q)
{
    // q holds the column of row r < h in bits 5r..5r+4; rows >= N-h mirror them
    for (int r = 0; r < N; r++) {
        int c;
        if (r < h) c = (int)((q >> (5 * r)) & 31u);
        else if (r >= N - h) c = N - 1 - (int)((q >> (5 * (N - 1 - r))) & 31u);
        else c = N / 2;
        int cc;
        if (c < h) cc = (int)((q >> (5 * c)) & 31u);
        else if (c >= N - h) cc = N - 1 - (int)((q >> (5 * (N - 1 - c))) & 31u);
        else cc = N / 2;
        if (cc != N - 1 - r) return false;
    }
    return true;
}

__global__ void k_symmetric(int N, int P, uint32_t nprefix, unsigned long long *__restrict__ out2)
{
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t c180 = 0, c90 = 0;
    const int h = N / 2;
    if (g < nprefix) {
        const uint32_t full = (1u << N) - 1u;
        uint32_t cols = 0;
        uint64_t dg = 0, an = 0;
        qrows_t q = 0;
        if (N & 1) {
            const int mid = N / 2;
            cols = 1u << mid;
            dg = 1ull << (N - 1);
            an = 1ull << (2 * mid);
        }
        bool ok = true;
        uint32_t rest = g;
        for (int r = 0; r < P && ok; r++) {
            const int c = (int)(rest % (uint32_t)N);
            rest /= (uint32_t)N;
            const uint32_t f = row_free(N, r, full, cols, dg, an);
            if (!((f >> c) & 1u)) { ok = false; break; }
            cols |= (1u << c) | (1u << (N - 1 - c));
            dg |= (1ull << (c - r + N - 1)) | (1ull << (r - c + N - 1));
            an |= (1ull << (r + c)) | (1ull << (2 * N - 2 - r - c));
            q |= (qrows_t)c << (5 * r);
        }
        if (ok) {
            if (P == h) {
                c180 = 1;
                c90 = quarter_turn(N, h, q) ? 1 : 0;
            } else {
                // iterative DFS over rows P..h-1
                uint32_t fst[16];
                uint32_t cst[16];
                uint64_t dst[16], ast[16];
                qrows_t qst[16];
                int r = P;
                uint32_t f = row_free(N, r, full, cols, dg, an);
                for (;;) {
                    if (f == 0u) {
                        if (r == P) break;
                        r--;
                        f = fst[r]; cols = cst[r]; dg = dst[r]; an = ast[r]; q = qst[r];
                        continue;
                    }
                    const uint32_t bit = f & (0u - f);
                    f ^= bit;
                    const int c = __builtin_ctz(bit);
                    const uint32_t ncols = cols | bit | (1u << (N - 1 - c));
                    const uint64_t ndg = dg | (1ull << (c - r + N - 1)) | (1ull << (r - c + N - 1));
                    const uint64_t nan_ = an | (1ull << (r + c)) | (1ull << (2 * N - 2 - r - c));
                    const qrows_t nq = q | ((qrows_t)c << (5 * r));
                    if (r + 1 == h) {
                        c180++;
                        if (quarter_turn(N, h, nq)) c90++;
                        continue;
                    }
                    fst[r] = f; cst[r] = cols; dst[r] = dg; ast[r] = an; qst[r] = q;
                    r++;
                    cols = ncols; dg = ndg; an = nan_; q = nq;
                    f = row_free(N, r, full, cols, dg, an);
                }
            }
        }
    }
    c180 = wave_sum_u64(c180);
    c90 = wave_sum_u64(c90);
    if (lane_id() == 0) {
        if (c180) atomicAdd(out2 + 1, (unsigned long long)c180);
        if (c90) atomicAdd(out2 + 0, (unsigned long long)c90);
    }
}

// ---------------------------------------------------------------------------
// launch wrappers (host)
// ---------------------------------------------------------------------------
hipError_t launch_decode(hipStream_t s, int N, const uint8_t *Ns, const uint4 *rec, uint32_t m, uint4 *root,
                         uint4 *sched, uint32_t *err)
{
    if (m == 0) return hipSuccess;
    k_decode<<<(m + 255) / 256, 256, 0, s>>>(N, Ns, rec, m, root, sched, err);
    return hipGetLastError();
}

hipError_t launch_plain(hipStream_t s, bool nodes, int block, WalkLds lay, const uint4 *root,
                        const uint4 *sched, uint32_t m, unsigned long long *weighted, unsigned long long *nodes_out,
                        unsigned long long *total)
{
    if (m == 0) return hipSuccess;
    if (block != (int)kPlainBlock) return hipErrorInvalidValue;
    const size_t lds = lay.bytes(block);
    const dim3 grid((m + block - 1) / block);
    if (nodes) k_plain<true><<<grid, block, lds, s>>>(root, sched, m, weighted, nodes_out, total);
    else k_plain<false><<<grid, block, lds, s>>>(root, sched, m, weighted, nodes_out, total);
    return hipGetLastError();
}

hipError_t launch_expand(hipStream_t s, bool nodes, bool write, int block, uint32_t stack_depth,
                         const uint4 *root, const uint4 *sched, const uint32_t *order, const uint8_t *splits,
                         uint32_t m, uint32_t split, uint32_t fan,
                         uint32_t *nitems, const uint32_t *off, uint4 *items,
                         unsigned long long *raw, unsigned long long *nodes_out)
{
    if (m == 0) return hipSuccess;
    if (fan == 0 || (uint64_t)m * fan >= (1ull << 32)) return hipErrorInvalidValue;
    const uint32_t lanes = m * fan;
    const size_t lds = (size_t)block * stack_depth * sizeof(uint4);
    const dim3 grid((lanes + block - 1) / block);
#define NQ_EXPAND(NB, WB) \
    k_expand<NB, WB><<<grid, block, lds, s>>>(root, sched, order, splits, m, split, fan, nitems, off, items, raw, \
                                              nodes_out)
    if (write) { if (nodes) NQ_EXPAND(true, true); else NQ_EXPAND(false, true); }
    else { if (nodes) NQ_EXPAND(true, false); else NQ_EXPAND(false, false); }
#undef NQ_EXPAND
    return hipGetLastError();
}

hipError_t launch_scan(hipStream_t s, const uint32_t *nitems, uint32_t m, uint32_t *off, unsigned long long *count,
                       unsigned long long *tsum, const uint32_t *order, uint32_t fan)
{
    const uint32_t tiles = (uint32_t)(((uint64_t)m + kScanTile - 1u) / kScanTile);
    if (order && fan == 0) return hipErrorInvalidValue;
    if (tiles) {
        k_scan_tile<<<tiles, kScanTile, 0, s>>>(nitems, m, order, fan, off, tsum);
        if (hipError_t e = hipGetLastError()) return e;
    }
    k_scan_tops<<<1, 1024, 0, s>>>(tsum, tiles, count);
    if (hipError_t e = hipGetLastError()) return e;
    if (tiles) k_scan_add<<<tiles, kScanTile, 0, s>>>(off, m, tsum);
    return hipGetLastError();
}

hipError_t launch_small_total(hipStream_t s, int N, unsigned long long *out)
{
    if (N < 1 || N > kSmallMaxN) return hipErrorInvalidValue;
    k_small_total<<<1, 64, 0, s>>>(N, out);
    return hipGetLastError();
}

hipError_t launch_preset(hipStream_t s, bool write, int N, const uint4 *in, uint32_t m, uint32_t *cnt,
                         const uint32_t *off, uint4 *out, uint32_t *err)
{
    if (m == 0) return hipSuccess;
    const uint32_t blocks = (m + 255u) / 256u;
    if (write) k_preset<true><<<blocks, 256, 0, s>>>(N, in, m, cnt, off, out, err);
    else k_preset<false><<<blocks, 256, 0, s>>>(N, in, m, cnt, off, out, err);
    return hipGetLastError();
}

template <bool NB, bool LB>
static const void *batched_fn(int block)
{
    switch (block) {
    case 64: return (const void *)k_batched<NB, LB, 64>;
    case 128: return (const void *)k_batched<NB, LB, 128>;
    case 256: return (const void *)k_batched<NB, LB, 256>;
    default: return nullptr;
    }
}

static const void *batched_kernel(bool nodes, bool lds_stack, int block)
{
    if (lds_stack) return nodes ? batched_fn<true, true>(block) : batched_fn<false, true>(block);
    return nodes ? batched_fn<true, false>(block) : batched_fn<false, false>(block);
}

hipError_t launch_batched(hipStream_t s, bool nodes, bool lds_stack, int grid, int block, WalkLds lay,
                          const uint4 *items,
                          const unsigned long long *n_items, unsigned int *head, uint32_t chunk,
                          unsigned long long *raw, unsigned long long *nodes_out)
{
    const void *f = batched_kernel(nodes, lds_stack, block);
    if (!f) return hipErrorInvalidValue;
    const size_t lds = lds_stack ? lay.bytes(block) : 0u;
    void *args[] = {(void *)&items, (void *)&n_items, (void *)&head, (void *)&chunk, (void *)&raw,
                    (void *)&nodes_out};
    return hipLaunchKernel(f, dim3(grid), dim3(block), args, lds, s);
}

hipError_t batched_occupancy(bool nodes, bool lds_stack, int block, size_t lds_bytes, int *blocks_per_cu)
{
    const void *f = batched_kernel(nodes, lds_stack, block);
    if (!f) return hipErrorInvalidValue;
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, f, block, lds_bytes);
}

hipError_t launch_finalize(hipStream_t s, const uint4 *sched, uint32_t m, const unsigned long long *raw,
                           unsigned long long *weighted, unsigned long long *total)
{
    if (m == 0) return hipSuccess;
    k_finalize<<<(m + 255) / 256, 256, 0, s>>>(sched, m, raw, weighted, total);
    return hipGetLastError();
}

hipError_t launch_fill_u64(hipStream_t s, unsigned long long *p, uint64_t n, unsigned long long v)
{
    if (n == 0) return hipSuccess;
    uint64_t blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    k_fill_u64<<<(uint32_t)blocks, 256, 0, s>>>(p, n, v);
    return hipGetLastError();
}

hipError_t launch_fill_u32(hipStream_t s, unsigned int *p, uint64_t n, unsigned int v)
{
    if (n == 0) return hipSuccess;
    uint64_t blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    k_fill_u32<<<(uint32_t)blocks, 256, 0, s>>>(p, n, v);
    return hipGetLastError();
}

hipError_t launch_symmetric(hipStream_t s, int N, unsigned long long *out2)
{
    const int h = N / 2;
    const int P = h < 3 ? h : 3;
    uint32_t np = 1;
    for (int a = 0; a < P; a++) np *= (uint32_t)N;
    k_symmetric<<<(np + 255) / 256, 256, 0, s>>>(N, P, np, out2);
    return hipGetLastError();
}

}  // namespace nqk
