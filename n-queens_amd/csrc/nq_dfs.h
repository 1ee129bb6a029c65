// nq_dfs.h -- the per-lane bitboard DFS engine shared by the plain, expand
// and batched kernels (nq_kernels.hip).  Host-compilable so that the engine
// logic can be exercised off the GPU.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "nq_kernels.h"

#ifndef NQ_HD
#define NQ_HD __host__ __device__
#endif

namespace nqk {

// ---------------------------------------------------------------------------
// per-record schedule (layout of the parameter bytes: nq_task.h)
// ---------------------------------------------------------------------------
struct Sched {
    uint32_t t0, t1, t2, t3;  // parameter byte of depth d in byte d of t0..t3 (d < 16)
    uint32_t T;               // terminal depth (<= 16)
    uint32_t tm;              // terminal free mask
    uint32_t mask;            // board mask (1 << N) - 1
    uint32_t nsh;             // N - 4: rd block constants are ARc << nsh
};

NQ_HD inline uint32_t tab_word(const Sched &S, uint32_t i)
{
    return i == 0u ? S.t0 : i == 1u ? S.t1 : i == 2u ? S.t2 : i == 3u ? S.t3 : 0u;
}

// Parameter byte of depth d (zero from depth 16 on).
NQ_HD inline uint32_t qbyte(const Sched &S, uint32_t d)
{
    return (tab_word(S, d >> 2) >> ((d & 3u) * 8u)) & 0xFFu;
}

// Parameter bytes of depths d..d+3 as one little-endian word.
NQ_HD inline uint32_t qbytes4(const Sched &S, uint32_t d)
{
    const uint32_t lo = tab_word(S, d >> 2), hi = tab_word(S, (d >> 2) + 1u);
    const uint32_t sh = (d & 3u) * 8u;
    return sh ? (lo >> sh) | (hi << (32u - sh)) : lo;
}

NQ_HD inline Sched load_sched(const uint4 *sched, uint32_t c)
{
    const uint4 a = sched[2 * c], b = sched[2 * c + 1];
    Sched S;
    S.t0 = a.x; S.t1 = a.y; S.t2 = a.z; S.t3 = a.w;
    S.T = b.y & 0xFFu;
    S.tm = b.z;
    const uint32_t N = (b.y >> 16) & 31u;
    S.mask = (1u << N) - 1u;
    S.nsh = N - 4u;
    return S;
}

// ---------------------------------------------------------------------------
// frames
// ---------------------------------------------------------------------------
// Frame word: avail (bits 0-18) | parameter byte of the frame's depth
// (19-26) | depth (27-31).  Depth in the top bits makes "child is on the
// terminal row" one unsigned compare of W + (1 << 27) against T << 27.
constexpr uint32_t kAvailMask = 0x7FFFFu;
constexpr uint32_t kQShift = 19u;
constexpr uint32_t kDShift = 27u;

NQ_HD inline uint32_t frame_word(uint32_t avail, uint32_t d, uint32_t q)
{
    return avail | (q << kQShift) | (d << kDShift);
}

// Per-lane constants of the walk.  An idle lane (mask 0, T27 all ones)
// steps without effect: no candidate, no count, no stack motion.
struct LaneK {
    uint32_t mask, tm;
    uint32_t T27;   // terminal depth << 27
    uint32_t nsh;
};
NQ_HD inline LaneK lane_consts(const Sched &S) { return LaneK{S.mask, S.tm, S.T << kDShift, S.nsh}; }
NQ_HD inline LaneK idle_lane() { return LaneK{0u, 0u, ~0u, 0u}; }

// ---------------------------------------------------------------------------
// LDS access by 32-bit LDS address (on the device the address goes straight
// into the ds instruction; on the host it indexes a byte array)
// ---------------------------------------------------------------------------
#if defined(__HIP_DEVICE_COMPILE__)
typedef __attribute__((address_space(3))) uint8_t lds_u8;
typedef __attribute__((address_space(3))) uint32_t lds_u32;
typedef __attribute__((address_space(3))) uint4 lds_u4;
struct LdsMem {
    NQ_HD inline uint4 ld4(uint32_t a) const { return *(const lds_u4 *)(size_t)a; }
    NQ_HD inline void st4(uint32_t a, uint4 v) const { *(lds_u4 *)(size_t)a = v; }
    NQ_HD inline uint32_t ld1(uint32_t a) const { return *(const lds_u8 *)(size_t)a; }
    NQ_HD inline void st32(uint32_t a, uint32_t v) const { *(lds_u32 *)(size_t)a = v; }
};
// LDS address of a pointer into the workgroup's shared memory
__device__ inline uint32_t lds_addr(const void *p) { return (uint32_t)(size_t)(const lds_u8 *)p; }
#else
struct LdsMem {
    uint8_t *mem;
    NQ_HD inline uint4 ld4(uint32_t a) const { return *(const uint4 *)(mem + a); }
    NQ_HD inline void st4(uint32_t a, uint4 v) const { *(uint4 *)(mem + a) = v; }
    NQ_HD inline uint32_t ld1(uint32_t a) const { return mem[a]; }
    NQ_HD inline void st32(uint32_t a, uint32_t v) const { *(uint32_t *)(mem + a) = v; }
};
// for the kernels' host pass only (kernels never run on the host)
NQ_HD inline uint32_t lds_addr(const void *) { return 0u; }
#endif

// ---------------------------------------------------------------------------
// stacks of 16-byte frames {ld, rd, col, frame word}, byte-addressed:
// slot i of a lane at base + i*S.  The walk keeps the address of its top
// entry, which is base - S for an empty stack.
// ---------------------------------------------------------------------------
struct LdsStack {
    LdsMem m;
    uint32_t S;       // slot stride: 16 bytes x workgroup size
    uint32_t base;    // this lane's slot 0
    NQ_HD inline uint32_t empty() const { return base - S; }
    NQ_HD inline uint4 read(uint32_t a) const { return m.ld4(a); }
    NQ_HD inline void write(uint32_t a, uint4 v) { m.st4(a, v); }
};
struct PrivStack {
    uint4 a[kStackCap];
    uint32_t S = 16u;
    uint32_t base = 0u;
    NQ_HD inline uint32_t empty() const { return base - S; }
    NQ_HD inline uint4 read(uint32_t o) const { return a[o >> 4]; }
    NQ_HD inline void write(uint32_t o, uint4 v) { a[o >> 4] = v; }
};

// A lane's parameter bytes, one byte per depth from the depth its walk
// starts at; off = table address minus that depth.
struct QTab {
    LdsMem m;
    uint32_t off;
    NQ_HD inline uint32_t get(uint32_t d) const { return m.ld1(off + d); }
};

// Write a lane's parameter bytes of depths d0 .. d0+ts-1 (ts a multiple of 4)
// at LDS address a.
NQ_HD inline void write_qtab(const LdsMem &m, uint32_t a, const Sched &S, uint32_t d0, uint32_t ts)
{
    for (uint32_t k = 0; k < ts; k += 4u) m.st32(a + k, qbytes4(S, d0 + k));
}

// A lane's DFS position: the current frame and the address of the top
// stack entry.
struct Walk {
    uint32_t ld, rd, col, W;
    uint32_t A;
};

// One DFS node, without divergent branches.  Every lane extracts one
// candidate of its frame and builds the child (18Py:368-562 one iteration):
// shift by the step, OR the block constants, drop column 0 when the child is
// the jmark frame, apply the one-row lookahead.  A child on the terminal row
// (depth T) counts one solution when its free set meets the terminal mask.
// The lane then descends (pushing the frame's remainder when it has one),
// keeps the remainder, or pops.  The top entry is read every step (used only
// when popping), so its load is issued a full step before its use.  Frames
// with remaining candidates are pushed from depths d0..T-2: T-1-d0 slots.
// Solutions are counted in a u32; `wrap` reports the step whose count
// wrapped to zero (the caller adds 2^32).  Returns true when the walk's
// subtree is exhausted.
template <bool NODES, class Stack>
NQ_HD inline bool walk_step(const LaneK &L, Walk &w, Stack &st, const QTab &qt, uint32_t &cnt, uint64_t &nodes,
                            bool &wrap)
{
    const uint32_t S = st.S;
    const uint4 top = st.read((uint32_t)max((int32_t)w.A, (int32_t)st.base));
    const uint32_t W = w.W;
    const uint32_t bit = W & (0u - W);                   // lowest candidate (avail is the low field)
    const uint32_t rest = W ^ bit;
    const uint32_t s = (W >> 19) & 3u;
    const uint32_t nld = ((w.ld | bit) << s) | ((W >> 21) & 3u);
    const uint32_t nrd = ((w.rd | bit) >> s) | (((W >> 23) & 3u) << L.nsh);
    const uint32_t ncol = w.col | bit;
    const uint32_t occ = nld | nrd | ((W >> 26) & 1u);   // JMN: column 0 of a jmark child
    const uint32_t nf = L.mask & ~(occ | ncol);
    const uint32_t nft = nf & L.tm;
    const uint32_t nla = (uint32_t)((int32_t)(W << 6) >> 31);   // all ones when the frame has no prune
    const bool lok = ((L.mask & ~((nld << 1) | (nrd >> 1) | ncol)) | nla) != 0u;
    const uint32_t Wd = W + (1u << kDShift);             // child depth in the top bits
    const uint32_t qc = qt.get(Wd >> kDShift);
    const bool leaf = Wd >= L.T27;
    uint32_t nc;
    wrap = __builtin_add_overflow(cnt, (uint32_t)(leaf & lok & (nft != 0u)), &nc);
    cnt = nc;
    if (NODES) nodes += 1u;
    const bool down = lok & (nf != 0u) & !leaf;
    const bool more = (rest & kAvailMask) != 0u;
    const bool push = down & more;
    const bool pop = !down & !more & (w.A != st.empty());
    if (push) st.write(w.A + S, make_uint4(w.ld, w.rd, w.col, rest));
    const uint32_t Wc = (Wd & (31u << kDShift)) | (qc << kQShift) | nf;
    w.ld = down ? nld : (pop ? top.x : w.ld);
    w.rd = down ? nrd : (pop ? top.y : w.rd);
    w.col = down ? ncol : (pop ? top.z : w.col);
    w.W = down ? Wc : (pop ? top.w : rest);
    w.A += push ? S : (pop ? 0u - S : 0u);
    return !down & !more & !pop;
}

// ---------------------------------------------------------------------------
// branchy DFS for the work-item expansion
// ---------------------------------------------------------------------------
// Index-addressed stack: slot i of a lane at p[i << shift].
struct SlotStack {
    uint4 *p;
    uint32_t shift;
    NQ_HD inline void put(int i, uint4 v) { p[(uint32_t)i << shift] = v; }
    NQ_HD inline uint4 get(int i) const { return p[(uint32_t)i << shift]; }
};

// One DFS from a frame at depth d (< T) until its subtree is exhausted,
// handing the children created at depth `split` (d < split < T) to emit()
// instead of expanding them.  Counts the solutions and nodes above the split.
template <bool NODES, class Stack, class Emit>
NQ_HD inline void expand_subtree(const Sched &S, uint32_t ld, uint32_t rd, uint32_t col, uint32_t avail,
                                 uint32_t d, uint32_t split, Stack &st, uint64_t &cnt, uint64_t &nodes,
                                 Emit &&emit)
{
    uint32_t q = qbyte(S, d);
    int sp = 0;
    for (;;) {
        const uint32_t bit = avail & (0u - avail);
        avail ^= bit;
        const uint32_t s = q & 3u;
        const uint32_t nld = ((ld | bit) << s) | ((q >> 2) & 3u);
        const uint32_t nrd = ((rd | bit) >> s) | (((q >> 4) & 3u) << S.nsh);
        const uint32_t ncol = col | bit;
        const uint32_t occ = nld | nrd | ncol | ((q >> 7) & 1u);
        const uint32_t nf = S.mask & ~occ;
        const bool lok = ((q >> 6) & 1u) != 0u || (S.mask & ~((nld << 1) | (nrd >> 1) | ncol)) != 0u;
        const uint32_t cd = d + 1u;
        if (NODES) nodes += 1u;
        bool down = false;
        if (cd == S.T) {
            if (lok && (S.tm & ~occ) != 0u) cnt += 1u;
        } else if (lok && nf != 0u) {
            if (cd == split) emit(nld, nrd, ncol, nf);
            else down = true;
        }
        if (down) {
            if (avail) { st.put(sp, make_uint4(ld, rd, col, avail | (d << kDShift))); sp++; }
            ld = nld; rd = nrd; col = ncol; avail = nf; d = cd; q = qbyte(S, cd);
        } else if (avail == 0u) {
            if (sp == 0) break;
            sp--;
            const uint4 e = st.get(sp);
            ld = e.x; rd = e.y; col = e.z;
            avail = e.w & kAvailMask;
            d = e.w >> kDShift;
            q = qbyte(S, d);
        }
    }
}

}  // namespace nqk
