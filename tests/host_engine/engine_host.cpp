// Host build of the device DFS engine (n-queens_amd/csrc/nq_task.h, nq_dfs.h):
// the work-item expansion and the walk of every item, record by record, with
// the LDS stack emulated in a byte array.  Used by tests/test_engine_host.py
// to compare the kernels' logic with the oracle on the CPU; not product code.
#include "../../n-queens_amd/csrc/nq_task.h"
#include "../../n-queens_amd/csrc/nq_dfs.h"
#include <stdint.h>
#include <vector>

using namespace nqk;

namespace {

uint8_t g_lds[1 << 16];

// Count one item (a frame at depth d0); -1 on a stack overflow past the
// slots the kernels allocate (`lay`: the set-wide layout, or when its slot
// count is 0 the item's own bound walk_lds(T, d0)), -2 if the ended walk
// does not stay inert.
int64_t walk_item(const Sched &S, uint32_t ld, uint32_t rd, uint32_t col, uint32_t av, uint32_t d0,
                  uint64_t &nodes, bool priv, WalkLds lay = WalkLds{0u})
{
    if (S.T == 0) return (av & S.tm) ? 1 : 0;
    if (lay.slots == 0u) lay = walk_lds(S.T, d0);
    const uint32_t lane = 5, block = 64;   // lane 5 of a 64-lane workgroup
    LdsMem m{g_lds};
    for (size_t a = 0; a < lay.bytes(block) + 64; a++) g_lds[a] = 0xA5;   // poison
    const LaneK L = lane_consts(S);
    uint32_t cnt = 0;
    uint64_t total = 0;
    bool done;
    if (priv) {
        PrivStack st;
        st.init();
        Walk w{ld, rd, free_cols(S, col), start_word(S, av, d0), st.empty()};
        do {
            done = walk_step<true>(L, w, st, cnt, nodes);
            if (done || cnt >= (1u << 31)) { total += cnt; cnt = 0; }
        } while (!done);
    } else {
        LdsStack st{m, 16u * block, 16u * (block + lane)};
        st.init();
        Walk w{ld, rd, free_cols(S, col), start_word(S, av, d0), st.empty()};
        do {
            done = walk_step<true>(L, w, st, cnt, nodes);
            if (done || cnt >= (1u << 31)) { total += cnt; cnt = 0; }
            if (w.A >= st.empty() + lay.slots * st.S) return -1;   // past the slots the kernels allocate
        } while (!done);
        // the walk sits on the inert frame: further steps change nothing
        const Walk w0 = w;
        const uint64_t n0 = nodes;
        for (int k = 0; k < 3; k++) (void)walk_step<true>(L, w, st, cnt, nodes);
        if (cnt || nodes != n0 || w.W != w0.W || w.A != w0.A || w.fc != w0.fc) return -2;
    }
    return (int64_t)total;
}

}  // namespace

namespace {

// Weighted count and node count of every record, expanded at its split depth
// (depths[a], or `split` for all when depths is NULL; 0 = whole records) and
// walked with the LDS (priv = 0) or private stack.  With set_layout the LDS
// stacks get the layout the library allocates for the whole set
// (walk_lds(max T, min item start depth)), else each item its own bound.
int count_records(int N, const uint32_t *rec, long m, uint64_t *weighted, uint64_t *nodes, int split,
                  const uint8_t *depths, int priv, bool set_layout, int gate_bits = 32)
{
    WalkLds set_lay{0u};
    if (set_layout) {
        uint32_t Tmax = 0, d0 = 255;
        for (long a = 0; a < m; a++) {
            nqt::Desc D;
            const uint32_t *r = rec + 4 * a;
            nqt::decode(N, r[0], r[1], r[2], r[3], D, gate_bits);
            if (D.status) return -1 - D.status;
            Tmax = D.T > Tmax ? D.T : Tmax;
            const uint32_t d = item_depth(depths ? depths[a] : (uint32_t)split, D.T);
            d0 = d < d0 ? d : d0;
        }
        if (m) set_lay = walk_lds(Tmax, d0);
    }
    for (long a = 0; a < m; a++) {
        if (depths) split = depths[a];
        nqt::Desc D;
        const uint32_t *r = rec + 4 * a;
        nqt::decode(N, r[0], r[1], r[2], r[3], D, gate_bits);
        if (D.status) return -1 - D.status;
        uint32_t wd[8];
        nqt::sched_words(D, N, wd);
        const uint4 sc[2] = {make_uint4(wd[0], wd[1], wd[2], wd[3]), make_uint4(wd[4], wd[5], wd[6], wd[7])};
        const Sched S = load_sched(sc, 0);
        uint64_t c = 0, n = 0;
        if (split > 0 && D.T >= 2 && (uint32_t)split <= D.T - 2 && D.avail) {
            struct It { uint32_t a, b, c, f; };
            std::vector<It> items;
            uint64_t c0 = 0;
            auto emit = [&](uint32_t x, uint32_t y, uint32_t z, uint32_t f) { items.push_back({x, y, z, f}); };
            uint4 slots[kStackCap];
            SlotStack st{slots, 0};
            expand_subtree<true>(S, D.ld, D.rd, D.col, D.avail, 0u, (uint32_t)split, st, c0, n, emit);
            c = c0;
            for (const It &it : items) {
                const int64_t k = walk_item(S, it.a, it.b, it.c, it.f, (uint32_t)split, n, priv != 0, set_lay);
                if (k < 0) return (int)(-100 + k);
                c += (uint64_t)k;
            }
        } else if (D.avail) {
            const int64_t k = walk_item(S, D.ld, D.rd, D.col, D.avail, 0u, n, priv != 0, set_lay);
            if (k < 0) return (int)(-100 + k);
            c = (uint64_t)k;
        }
        weighted[a] = c << D.wshift;
        nodes[a] = n;
    }
    return 0;
}

}  // namespace

// Returns 0, -1 - decode status, -101 on a stack overflow, -102 if an ended
// walk does not stay inert.
extern "C" int engine_count(int N, const uint32_t *rec, long m, uint64_t *weighted, uint64_t *nodes, int split,
                            int priv)
{
    return count_records(N, rec, m, weighted, nodes, split, nullptr, priv, false);
}

// engine_count with build_soa's j-diagonal gate at gate_bits instead of 32
// (records beyond it take the P5 functions; any width >= N counts the same).
extern "C" int engine_count_gate(int N, const uint32_t *rec, long m, uint64_t *weighted, uint64_t *nodes,
                                 int split, int priv, int gate_bits)
{
    return count_records(N, rec, m, weighted, nodes, split, nullptr, priv, false, gate_bits);
}

// P5 target count of the records under a gate width (decode fids 8..11).
extern "C" long engine_p5_records(int N, const uint32_t *rec, long m, int gate_bits)
{
    long n = 0;
    for (long a = 0; a < m; a++) {
        nqt::Desc D;
        const uint32_t *r = rec + 4 * a;
        nqt::decode(N, r[0], r[1], r[2], r[3], D, gate_bits);
        n += D.fid >= 8 && D.fid <= 11;
    }
    return n;
}

// The same with a split depth per record and the LDS layout sized for the
// whole set as nq_ctx_set_splits + nq_ctx_run_batched size it.
extern "C" int engine_count_splits(int N, const uint32_t *rec, long m, uint64_t *weighted, uint64_t *nodes,
                                   const uint8_t *depths)
{
    return count_records(N, rec, m, weighted, nodes, 0, depths, 0, true);
}

// One level of preset expansion of every record (nqt::preset_children),
// written to out (capacity cap records); returns the record count, or
// -1 - status on a failed record, or -100 when out is too small.
extern "C" long engine_expand(int N, const uint32_t *rec, long m, uint32_t *out, long cap)
{
    long n = 0;
    for (long a = 0; a < m; a++) {
        const uint32_t *r = rec + 4 * a;
        bool full = false;
        const int32_t st = nqt::preset_children(N, r[0], r[1], r[2], r[3],
                                                [&](uint32_t x, uint32_t y, uint32_t z, uint32_t w) {
                                                    if (n >= cap) { full = true; return; }
                                                    uint32_t *o = out + 4 * n++;
                                                    o[0] = x; o[1] = y; o[2] = z; o[3] = w;
                                                });
        if (st) return -1 - st;
        if (full) return -100;
    }
    return n;
}
