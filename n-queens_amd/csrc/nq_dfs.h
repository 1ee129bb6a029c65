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
// per-record schedule (layout: nqt::sched_words)
// ---------------------------------------------------------------------------
struct Sched {
    uint32_t t[4];            // byte r (0..15): parameter byte of depth T-1-r
    uint32_t T;               // terminal depth (<= 16)
    uint32_t tm;              // terminal free mask
    uint32_t mask;            // board mask (1 << N) - 1
    uint32_t nsh;             // N - 4: rd block constants are ARc << nsh
};

// Byte r (< 16) of a 16-byte table held in four words.
NQ_HD inline uint32_t table_byte(uint32_t t0, uint32_t t1, uint32_t t2, uint32_t t3, uint32_t r)
{
    const uint32_t w = r < 4u ? t0 : r < 8u ? t1 : r < 12u ? t2 : t3;
    return (w >> ((r & 3u) * 8u)) & 0xFFu;
}

// Parameter byte of depth d < T.
NQ_HD inline uint32_t qbyte(const Sched &S, uint32_t d)
{
    return table_byte(S.t[0], S.t[1], S.t[2], S.t[3], S.T - 1u - d);
}

NQ_HD inline Sched load_sched(const uint4 *sched, uint32_t c)
{
    const uint4 a = sched[kSchedU4 * c], i = sched[kSchedU4 * c + 1];
    Sched S;
    S.t[0] = a.x; S.t[1] = a.y; S.t[2] = a.z; S.t[3] = a.w;
    S.T = i.x & 0xFFu;
    S.tm = i.y;
    const uint32_t N = (i.x >> 16) & 31u;
    S.mask = (1u << N) - 1u;
    S.nsh = N - 4u;
    return S;
}

// ---------------------------------------------------------------------------
// frames
// ---------------------------------------------------------------------------
// A frame is {ld, rd, col, W}: the diagonals, the occupied columns (board
// bits 0..N-1, N <= 28) and the frame word W, which holds the candidates
// (avail, bits 0-27) and r = T-1-depth (bits 28-31).  The frame's children
// are on the terminal row exactly when r = 0, i.e. when W is below 2^28; a
// child's word is W's top bits less 2^28 plus its candidates; and the
// frame's parameter byte is byte r of the lane's schedule table, looked up
// from W by each step (frame_q).
constexpr uint32_t kAvailMask = 0x0FFFFFFFu;
constexpr uint32_t kQShift = 24u;
constexpr uint32_t kRShift = 28u;

NQ_HD inline uint32_t frame_word(uint32_t avail, uint32_t r) { return avail | (r << kRShift); }

// Frame word of a frame at depth d with candidates avail.
NQ_HD inline uint32_t start_word(const Sched &S, uint32_t avail, uint32_t d)
{
    return frame_word(avail, S.T - 1u - d);
}

// Any function of three words, bit by bit: bit i of IMM is the result for
// the input bits (a, b, c) = (i >> 2, i >> 1, i) & 1.  On gfx950 this is one
// v_bitop3_b32, which issues at the full VALU rate with three VGPR operands
// (v_or3_b32, v_and_or_b32, v_bfi_b32, v_perm_b32 and v_cndmask_b32 issue at
// half rate).
template <uint32_t IMM>
NQ_HD inline uint32_t bop3(uint32_t a, uint32_t b, uint32_t c)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_bitop3_b32(a, b, c, IMM);
#else
    uint32_t r = 0;
    for (uint32_t i = 0; i < 8u; i++)
        if ((IMM >> i) & 1u) r |= ((i & 4u) ? a : ~a) & ((i & 2u) ? b : ~b) & ((i & 1u) ? c : ~c);
    return r;
#endif
}
constexpr uint32_t kOr3 = 0xFEu;        // a | b | c
constexpr uint32_t kSel = 0xCAu;        // a ? b : c, bit by bit
constexpr uint32_t kAndNotOr = 0x10u;   // a & ~(b | c)
constexpr uint32_t kAndNotOrC = 0xBAu;  // (a & ~b) | c
constexpr uint32_t kOrAnd = 0xF8u;      // a | (b & c)

// A constant held in a VGPR: bitop3 reads its operands from VGPRs at the
// full rate, an SGPR or literal operand halves it.
NQ_HD inline uint32_t vconst(uint32_t v)
{
#if defined(__HIP_DEVICE_COMPILE__)
    asm volatile("" : "+v"(v));
#endif
    return v;
}

// Per-lane constants of the walk: board mask, rd shift, the parameter bytes
// by r and the masks below (in VGPRs).  An idle lane (mask 0) steps without
// effect: no candidate, no count, no stack motion.
// With the frame's parameter byte q in bits 24-31 of a word, the rd constant
// of the child is (q >> (32-N)) & km, km = 0b1011 << (N-4): ARc (bits 28-29)
// to N-4.., JD (bit 31) to N-1.
constexpr uint32_t kNlaBit = 1u << 30;   // the NLA flag of q
struct LaneK {
    uint32_t mask;
    uint32_t ksh, km;     // 32 - N and 0b1011 << (N-4)
    uint32_t t0, t1, t2, t3;
    uint32_t nlab, three;
};
NQ_HD inline LaneK lane_k(uint32_t N, uint32_t t0, uint32_t t1, uint32_t t2, uint32_t t3)
{
    return LaneK{(1u << N) - 1u, 32u - N, 0xBu << (N - 4u), t0, t1, t2, t3, vconst(kNlaBit), vconst(3u)};
}
NQ_HD inline LaneK lane_consts(const Sched &S) { return lane_k(S.nsh + 4u, S.t[0], S.t[1], S.t[2], S.t[3]); }
NQ_HD inline LaneK idle_lane() { return LaneK{0u, 28u, 3u, 0u, 0u, 0u, 0u, vconst(kNlaBit), vconst(3u)}; }

// Parameter byte of a frame, in bits 24-31 (bits 0-23 are not defined), from
// its frame word (top bits: the frame's r).  On the device two v_perm_b32
// place the byte from each register pair in byte 3: the selector is W's bits
// 4-30, whose byte 3 is r & 7 (its other bytes select arbitrary bytes); the
// sign of W (r >= 8) picks the pair.  An indexed select over the words would
// be lowered to a scratch array.
NQ_HD inline uint32_t frame_q(const LaneK &L, uint32_t W)
{
#if defined(__HIP_DEVICE_COMPILE__)
    const uint32_t sel = __builtin_amdgcn_ubfe(W, 4u, 27u);
    const uint32_t a = __builtin_amdgcn_perm(L.t1, L.t0, sel);
    const uint32_t b = __builtin_amdgcn_perm(L.t3, L.t2, sel);
    return bop3<kSel>((uint32_t)((int32_t)W >> 31), b, a);
#else
    return table_byte(L.t0, L.t1, L.t2, L.t3, W >> kRShift) << kQShift;
#endif
}

// ---------------------------------------------------------------------------
// LDS access by 32-bit LDS address (on the device the address goes straight
// into the ds instruction; on the host it indexes a byte array)
// ---------------------------------------------------------------------------
#if defined(__HIP_DEVICE_COMPILE__)
typedef __attribute__((address_space(3))) uint8_t lds_u8;
typedef __attribute__((address_space(3))) uint4 lds_u4;
struct LdsMem {
    NQ_HD inline uint4 ld4(uint32_t a) const { return *(const lds_u4 *)(size_t)a; }
    NQ_HD inline void st4(uint32_t a, uint4 v) const { *(lds_u4 *)(size_t)a = v; }
};
// LDS address of a pointer into the workgroup's shared memory
__device__ inline uint32_t lds_addr(const void *p) { return (uint32_t)(size_t)(const lds_u8 *)p; }
#else
struct LdsMem {
    uint8_t *mem;
    NQ_HD inline uint4 ld4(uint32_t a) const { return *(const uint4 *)(mem + a); }
    NQ_HD inline void st4(uint32_t a, uint4 v) const { *(uint4 *)(mem + a) = v; }
};
// for the kernels' host pass only (kernels never run on the host)
NQ_HD inline uint32_t lds_addr(const void *) { return 0u; }
#endif

// ---------------------------------------------------------------------------
// stacks of 16-byte frames {ld, rd, col, frame word}, byte-addressed:
// slot i of a lane at base + i*S.  The slot below base holds the inert frame
// {0, 0, all columns, no candidates}: the top entry of an empty stack.  The
// walk keeps the address of its top entry (base - S when empty); a walk that
// ends takes the inert frame and stays on it, stepping without effect.
// ---------------------------------------------------------------------------
NQ_HD inline uint4 inert_frame() { return make_uint4(0u, 0u, ~0u, 0u); }

struct LdsStack {
    LdsMem m;
    uint32_t S;       // slot stride: 16 bytes x workgroup size
    uint32_t base;    // this lane's slot 0 (the inert frame sits at base - S)
    NQ_HD inline uint32_t empty() const { return base - S; }
    NQ_HD inline uint4 read(uint32_t a) const { return m.ld4(a); }
    NQ_HD inline void write(uint32_t a, uint4 v) { m.st4(a, v); }
    NQ_HD inline void init() { write(empty(), inert_frame()); }
};
struct PrivStack {
    uint4 a[kStackCap + 1];
    uint32_t S = 16u;
    uint32_t base = 16u;
    NQ_HD inline uint32_t empty() const { return base - S; }
    NQ_HD inline uint4 read(uint32_t o) const { return a[o >> 4]; }
    NQ_HD inline void write(uint32_t o, uint4 v) { a[o >> 4] = v; }
    NQ_HD inline void init() { write(empty(), inert_frame()); }
};

// A lane's DFS position: the current frame and the address of the top
// stack entry.
struct Walk {
    uint32_t ld, rd, col, W;
    uint32_t A;
};

// One DFS node, without divergent branches.  Every lane extracts one
// candidate of its frame and builds the child (18Py:368-562 one iteration):
// shift by the step, OR the block constants (the jmark's ld bit and the P5
// j diagonal among them), apply the one-row lookahead to the shifted ld
// before the constants.  A child on the terminal row (depth T) counts one
// solution when its free set is not empty.
// The lane then descends (pushing the frame's remainder when it has one),
// keeps the remainder, or pops.  The top entry is read every step (used only
// when popping), so its load is issued a full step before its use.  Frames
// with remaining candidates are pushed from depths d0..T-2: T-1-d0 slots
// above the inert frame's.  Returns true when the walk's subtree is
// exhausted (the step took the inert frame).
// ISSUE: spend two full-rate VALU ops where one slower op would do (the
// child's frame word).  That pays where the SIMD is issue-bound (the
// persistent kernel, 16 waves per CU) and costs where a lane's dependency
// chain is the limit (one lane per constellation).
template <bool NODES, bool ISSUE = true, class Count, class Stack>
NQ_HD inline bool walk_step(const LaneK &L, Walk &w, Stack &st, Count &cnt, uint64_t &nodes)
{
    const uint32_t S = st.S;
    const uint4 top = st.read(w.A);
    const uint32_t W = w.W, C = w.col;
    const uint32_t q = frame_q(L, W);                    // this frame's parameter byte (bits 24-31)
    const uint32_t bit = W & (0u - W);                   // lowest candidate (avail is the low field)
    const uint32_t rest = W ^ bit;
    const uint32_t s = (q >> 24) & 3u;
    const uint32_t x = (w.ld | bit) << s;                // the child's ld before its constants
    const uint32_t nld = bop3<kOrAnd>(x, q >> 26, L.three);          // | AL
    const uint32_t nrd = bop3<kOrAnd>((w.rd | bit) >> s, q >> L.ksh, L.km);
    const uint32_t ncol = C | bit;
    uint32_t nf = bop3<kAndNotOr>(L.mask, nld | nrd, ncol);
    // nonzero when the frame has no prune
    const uint32_t nla = q & L.nlab;
#if defined(__HIP_DEVICE_COMPILE__)
    asm("" : "+v"(nf));   // the child's word below reuses nf instead of refolding it
    uint32_t x2;
    // x << 1 as an add: v_add_u32 issues at the full rate, v_lshlrev_b32
    // at half (the compiler turns x + x back into a shift)
    asm("v_add_u32 %0, %1, %1" : "=v"(x2) : "v"(x));
#else
    const uint32_t x2 = x << 1;
#endif
    const uint32_t look = bop3<kOr3>(x2, nrd >> 1, ncol);
    const bool lok = bop3<kAndNotOrC>(L.mask, look, nla) != 0u;
    const uint32_t hiW = W & ~kAvailMask;                // r << 28
    const bool inner = W >= (1u << kRShift);             // children not on the terminal row
    const bool ok = lok & (nf != 0u);
    const bool down = ok & inner;
    const bool sol = ok ^ down;
    cnt += (Count)sol;
#if defined(__HIP_DEVICE_COMPILE__)
    asm volatile("" : "+v"(cnt));   // keeps the add beside its condition (one add-with-carry)
#endif
    if (NODES) nodes += (uint64_t)(W != 0u);             // the inert frame is no node
    const bool more = rest > hiW;                        // candidates left after this one
    const bool push = down & more;
    const bool up = !down & !more;                       // take the top entry
    const bool nonempty = w.A != st.empty();
    const bool pop = up & nonempty;                      // ... and drop it from the stack
    const uint4 rem = make_uint4(w.ld, w.rd, C, rest);
    const uint32_t A = w.A;
    // the child: r - 1, candidates nf
    uint32_t Wc = hiW + nf - (1u << kRShift);
#if defined(__HIP_DEVICE_COMPILE__)
    if constexpr (ISSUE)   // two two-operand adds, not a v_add3 with a scalar operand
        asm("v_add_u32 %0, %1, %2" : "=v"(Wc) : "v"(hiW - (1u << kRShift)), "v"(nf));
#endif
    w.ld = down ? nld : (up ? top.x : w.ld);
    w.rd = down ? nrd : (up ? top.y : w.rd);
    w.col = down ? ncol : (up ? top.z : C);
    w.W = down ? Wc : (up ? top.w : rest);
    w.A += push ? S : (pop ? 0u - S : 0u);
    // written after the selects: the wait for the top entry does not also
    // wait for this write
#if defined(__HIP_DEVICE_COMPILE__)
    asm volatile("" ::"v"(w.ld), "v"(w.rd), "v"(w.col), "v"(w.W));
#endif
    if (push) st.write(A + S, rem);
    return up & !nonempty;
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
        const uint32_t x = (ld | bit) << s;
        const uint32_t nld = x | ((q >> 2) & 3u);
        const uint32_t nrd = ((rd | bit) >> s) | (((q >> 4) & 3u) << S.nsh) | (((q >> 7) & 1u) << (S.nsh + 3u));
        const uint32_t ncol = col | bit;
        const uint32_t occ = nld | nrd | ncol;
        const uint32_t nf = S.mask & ~occ;
        const bool lok = ((q >> 6) & 1u) != 0u || (S.mask & ~((x << 1) | (nrd >> 1) | ncol)) != 0u;
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
            if (avail) { st.put(sp, make_uint4(ld, rd, col, avail | (d << kRShift))); sp++; }
            ld = nld; rd = nrd; col = ncol; avail = nf; d = cd; q = qbyte(S, cd);
        } else if (avail == 0u) {
            if (sp == 0) break;
            sp--;
            const uint4 e = st.get(sp);
            ld = e.x; rd = e.y; col = e.z;
            avail = e.w & kAvailMask;   // avail < 2^28 sits below the depth
            d = e.w >> kRShift;
            q = qbyte(S, d);
        }
    }
}

}  // namespace nqk
