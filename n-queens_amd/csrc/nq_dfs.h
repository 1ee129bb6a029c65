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
// per-lane schedule registers
// ---------------------------------------------------------------------------
struct Sched {
    uint32_t t0, t1, t2, t3, t4;  // parameter bytes for depths 0..15 (t4 unused by the walk)
    uint32_t T;                   // terminal depth
    uint32_t tm;                  // terminal free mask
    uint32_t mask;                // board mask (1 << N) - 1
    uint32_t nsh;                 // N - 4: rd block constants are ARc << nsh
};

// Byte d of the parameter table, for 0 <= d < 16; zero for d >= 16 (the
// terminal depth T <= 16 never needs its own byte).  v_perm_b32 picks byte
// (d & 7) of a register pair directly, which keeps the table in VGPRs (an
// indexed select over the words is otherwise lowered to a scratch array).
// Selector byte 0x0c yields zero, so only byte 0 of each result is non-zero;
// the pair choice uses masks so that no divergent branch is generated.
NQ_HD inline uint32_t pbyte(const Sched &S, uint32_t d)
{
#if defined(__HIP_DEVICE_COMPILE__)
    const uint32_t sel = 0x0c0c0c00u | (d & 7u);
    const uint32_t a = __builtin_amdgcn_perm(S.t1, S.t0, sel);
    const uint32_t b = __builtin_amdgcn_perm(S.t3, S.t2, sel);
    const uint32_t lo = 0u - (uint32_t)(d < 8u);
    return ((a & lo) | (b & ~lo)) & (0u - (uint32_t)(d < 16u));
#else
    const uint32_t w[4] = {S.t0, S.t1, S.t2, S.t3};
    return d < 16u ? (w[d >> 2] >> ((d & 3u) * 8u)) & 0xFFu : 0u;
#endif
}

NQ_HD inline Sched load_sched(const uint4 *sched, uint32_t c)
{
    uint4 a = sched[2 * c], b = sched[2 * c + 1];
    Sched S;
    S.t0 = a.x; S.t1 = a.y; S.t2 = a.z; S.t3 = a.w; S.t4 = b.x;
    S.T = b.y & 0xFFu;
    S.tm = b.z;
    const uint32_t N = (b.y >> 16) & 31u;
    S.mask = (1u << N) - 1u;
    S.nsh = N - 4u;
    return S;
}

// Solutions below a pre-terminal frame (depth T-1) without expanding it.
// For a placement `bit` the terminal free set is F & ~(bit | bit<<s | bit>>s),
// F = free set of the terminal row before that placement.  A bit b is a
// dead end exactly when every f in F is one of b, b<<s, b>>s, i.e. when b is
// covered by all |F| of the masks F<<s, F, F>>s at once.  The three masks
// cover a position at most once per f, so the dead bits are: any cover for
// |F| = 1, the majority of the three masks for |F| = 2, all three for
// |F| = 3, none for |F| >= 4, and every bit when F is empty.
NQ_HD inline uint32_t term_count(uint32_t ld, uint32_t rd, uint32_t col, uint32_t avail,
                                 uint32_t p, uint32_t tm, uint32_t nsh)
{
    const uint32_t s = (p & 3u) + 1u;
    const uint32_t X = (ld << s) | ((p >> 2) & 3u) | (rd >> s) | (((p >> 4) & 3u) << nsh) | col;
    const uint32_t F = tm & ~X;
    const uint32_t a = F << s, c = F >> s;
    const uint32_t pc = (uint32_t)__builtin_popcount(F);
    const uint32_t any = a | F | c;
    const uint32_t maj = (a & F) | (a & c) | (F & c);
    const uint32_t all = a & F & c;
    uint32_t B = (pc == 1u) ? any : ((pc == 2u) ? maj : ((pc == 3u) ? all : 0u));
    B = (pc == 0u) ? ~0u : B;
    return (uint32_t)__builtin_popcount(avail & ~B);
}

// ---------------------------------------------------------------------------
// DFS stacks: LDS (slot i of a lane at base + (i << shift) bytes, i.e. one
// row of 16-byte entries per slot across the workgroup) or private memory
// ---------------------------------------------------------------------------
struct LdsStack {
    uint4 *base;       // this lane's slot 0
    uint32_t shift;    // log2(workgroup size): slot stride in uint4 units
    NQ_HD inline uint32_t stride() const { return 1u << shift; }
    NQ_HD inline void put(int i, uint4 v) { base[(uint32_t)i << shift] = v; }
    NQ_HD inline uint4 get(int i) const { return base[(uint32_t)i << shift]; }
    // by offset (slot << shift), as the walk keeps it
    NQ_HD inline void put_at(uint32_t o, uint4 v) { base[o] = v; }
    NQ_HD inline uint4 get_at(uint32_t o) const { return base[o]; }
};
struct PrivStack {
    uint4 a[kStackCap];
    NQ_HD inline uint32_t stride() const { return 1u; }
    NQ_HD inline void put(int i, uint4 v) { a[i] = v; }
    NQ_HD inline uint4 get(int i) const { return a[i]; }
    NQ_HD inline void put_at(uint32_t o, uint4 v) { a[o] = v; }
    NQ_HD inline uint4 get_at(uint32_t o) const { return a[o]; }
};

// Frame word: avail (bits 0-18) | depth (19-23) | parameter byte (24-31).
NQ_HD inline uint32_t frame_word(uint32_t avail, uint32_t d, uint32_t p)
{
    return avail | (d << 19) | (p << 24);
}

// Stack entry: {ld, rd, col, frame word}.
NQ_HD inline uint4 pack_frame(uint32_t ld, uint32_t rd, uint32_t col, uint32_t avail,
                              uint32_t d, uint32_t p)
{
    return make_uint4(ld, rd, col, frame_word(avail, d, p));
}

// One DFS from a frame at depth d (d <= T-2, avail != 0 after jmark) until
// its subtree is exhausted, handing children created at depth `split` to
// emit() instead of expanding them.  Used by the work-item expansion.
template <bool NODES, bool SPLIT, class Stack, class Emit>
NQ_HD inline void dfs_subtree(uint32_t mask, uint32_t nsh, const Sched &S, uint32_t ld,
                              uint32_t rd, uint32_t col, uint32_t avail, uint32_t d,
                              uint32_t split, Stack &st, uint64_t &cnt, uint64_t &nodes,
                              Emit &&emit)
{
    uint32_t p = pbyte(S, d);
    const uint32_t Tm1 = S.T - 1u;
    int sp = 0;
    for (;;) {
        const uint32_t s = (p & 3u) + 1u;
        const uint32_t bit = avail & (0u - avail);
        avail ^= bit;
        uint32_t nld = ((ld | bit) << s) | ((p >> 2) & 3u);
        const uint32_t nrd = ((rd | bit) >> s) | (((p >> 4) & 3u) << nsh);
        const uint32_t ncol = col | bit;
        uint32_t nf = mask & ~(nld | nrd | ncol);
        const uint32_t look = mask & ~((nld << 1) | (nrd >> 1) | ncol);
        const uint32_t cd = d + 1u;
        const uint32_t cp = pbyte(S, cd);
        bool ok = (nf != 0u) && (((p & 0x40u) == 0u) || (look != 0u));
        if (cp & 0x80u) { nf &= ~1u; nld |= 1u; }
        ok = ok && (nf != 0u);
        if (NODES) nodes += 1u;
        if (cd == Tm1) {
            if (ok) {
                cnt += term_count(nld, nrd, ncol, nf, cp, S.tm, nsh);
                if (NODES) nodes += (uint32_t)__builtin_popcount(nf);
            }
            ok = false;
        }
        if (SPLIT && ok && cd == split) {
            emit(nld, nrd, ncol, nf);
            ok = false;
        }
        if (ok) {
            if (avail) { st.put(sp, pack_frame(ld, rd, col, avail, d, p)); sp++; }
            ld = nld; rd = nrd; col = ncol; avail = nf; d = cd; p = cp;
        } else if (avail == 0u) {
            if (sp == 0) break;
            sp--;
            const uint4 e = st.get(sp);
            ld = e.x; rd = e.y; col = e.z;
            avail = e.w & 0x7FFFFu;
            d = (e.w >> 19) & 31u;
            p = e.w >> 24;
        }
    }
}

// A lane's DFS position: the current frame {ld, rd, col, frame word} and
// the offset of the first free stack slot (entries below it x slot stride).
struct Walk {
    uint32_t ld, rd, col, W;
    uint32_t so;
};

// One DFS node, without divergent branches.  Every lane extracts one
// candidate of its frame and builds the child; a child on the terminal row
// (depth T) counts one solution when its free set meets the terminal mask
// (the base frame of 18Py:226-234: functionid 14 ignores column bit 0).
// The lane then descends, keeps its frame, or pops.  The frame remainder is
// written to the free slot `sp` every step (kept only when descending) and
// the entry `sp-1` is read every step (used only when popping), so the LDS
// traffic is uniform and the load is issued a full step ahead of its use.
// Needs one stack slot more than the deepest push.  Returns true when the
// walk's subtree is exhausted.
template <bool NODES, class Stack>
NQ_HD inline bool walk_step(uint32_t mask, uint32_t nsh, const Sched &S, Walk &w, Stack &st, uint64_t &cnt,
                            uint64_t &nodes)
{
    const uint32_t stride = st.stride();
    const uint4 top = st.get_at((uint32_t)max((int)w.so - (int)stride, 0));
    const uint32_t W = w.W;
    const uint32_t avail = W & 0x7FFFFu;
    const uint32_t bit = avail & (0u - avail);
    const uint32_t rest = W ^ bit;                       // remainder, same depth and parameters
    const uint32_t s = ((W >> 24) & 3u) + 1u;
    uint32_t nld = ((w.ld | bit) << s) | ((W >> 26) & 3u);
    const uint32_t nrd = ((w.rd | bit) >> s) | (((W >> 28) & 3u) << nsh);
    const uint32_t ncol = w.col | bit;
    uint32_t nf = mask & ~(nld | nrd | ncol);
    const uint32_t look = mask & ~((nld << 1) | (nrd >> 1) | ncol);
    const uint32_t cd = ((W >> 19) & 31u) + 1u;
    const uint32_t cp = pbyte(S, cd);
    const uint32_t nla = ((W >> 30) & 1u) - 1u;          // all ones when the frame has no lookahead
    const uint32_t jm = cp >> 7;                         // jmark frame: clear column 0, ld |= 1
    const bool ok0 = (nf != 0u) & ((look | nla) != 0u);
    nf &= ~jm;
    nld |= jm;
    const bool ok = ok0 & (nf != 0u);
    const bool leaf = (cd == S.T);
    cnt += (uint32_t)(leaf & ok & ((nf & S.tm) != 0u));
    if (NODES) nodes += 1u;
    st.put_at(w.so, make_uint4(w.ld, w.rd, w.col, rest));
    const bool down = ok & !leaf;
    const bool more = (rest & 0x7FFFFu) != 0u;
    const bool pop = !down & !more & (w.so != 0u);
    const uint32_t Wc = nf | (cd << 19) | (cp << 24);
    // next frame = child (down), popped entry (pop) or this frame's remainder,
    // selected with masks so that all lanes run the same instructions
    const uint32_t dm = 0u - (uint32_t)down, pm = 0u - (uint32_t)pop;
    w.ld = (nld & dm) | (((top.x & pm) | (w.ld & ~pm)) & ~dm);
    w.rd = (nrd & dm) | (((top.y & pm) | (w.rd & ~pm)) & ~dm);
    w.col = (ncol & dm) | (((top.z & pm) | (w.col & ~pm)) & ~dm);
    w.W = (Wc & dm) | (((top.w & pm) | (rest & ~pm)) & ~dm);
    w.so += (down & more) ? stride : (pop ? (0u - stride) : 0u);
    return !down && !more && !pop;
}

struct NoEmit {
    NQ_HD inline void operator()(uint32_t, uint32_t, uint32_t, uint32_t) const {}
};


}  // namespace nqk
