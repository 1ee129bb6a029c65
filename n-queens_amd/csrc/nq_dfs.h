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
// A frame is {ld, rd, fc, W}: the diagonals, the free columns (the board's
// columns 0..N-1, N <= 28, that hold no queen: mask & ~col) and the frame
// word W, which holds the candidates
// (avail, bits 0-27) and r = T-1-depth (bits 28-31).  The frame's children
// are on the terminal row exactly when r = 0, i.e. when W is below 2^28; a
// child's word is W's top bits less 2^28 plus its candidates; and the
// frame's parameter byte is byte r of the lane's schedule table, looked up
// from W by each step (frame_q).
constexpr uint32_t kAvailMask = 0x0FFFFFFFu;
constexpr uint32_t kQShift = 24u;
constexpr uint32_t kRShift = 28u;

NQ_HD inline uint32_t frame_word(uint32_t avail, uint32_t r) { return avail | (r << kRShift); }

// The frame's free columns from the occupied ones.
NQ_HD inline uint32_t free_cols(const Sched &S, uint32_t col) { return S.mask & ~col; }

// Frame word of a frame at depth d with candidates avail.
NQ_HD inline uint32_t start_word(const Sched &S, uint32_t avail, uint32_t d)
{
    return frame_word(avail, S.T - 1u - d);
}

// Any function of three words, bit by bit: bit i of IMM is the result for
// the input bits (a, b, c) = (i >> 2, i >> 1, i) & 1.  On gfx950 this is one
// v_bitop3_b32, which issues at the full VALU rate with three VGPR operands
// (v_or3_b32, v_and_or_b32, v_bfi_b32 and v_perm_b32 take about twice as
// long and pair with no neighbour).
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
constexpr uint32_t kSel = 0xCAu;        // a ? b : c, bit by bit
constexpr uint32_t kAndNotOr = 0x10u;   // a & ~(b | c)
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

// Per-lane constants of the walk: rd shift, the parameter bytes by r and the
// masks below (in VGPRs).  A lane on the inert frame steps without effect:
// no candidate, no count, no stack motion.
// With the frame's parameter byte q in bits 24-31 of a word, the rd constant
// of the child is (q >> (32-N)) & km, km = 0b1011 << (N-4): ARc (bits 28-29)
// to N-4.., JD (bit 31) to N-1.
constexpr uint32_t kNlaBit = 1u << 30;   // the NLA flag of q
struct LaneK {
    uint32_t ksh, km;     // 32 - N and 0b1011 << (N-4)
    uint32_t t0, t1, t2, t3;
    uint32_t nlab, three;
};
NQ_HD inline LaneK lane_k(uint32_t N, uint32_t t0, uint32_t t1, uint32_t t2, uint32_t t3)
{
    return LaneK{32u - N, 0xBu << (N - 4u), t0, t1, t2, t3, vconst(kNlaBit), vconst(3u)};
}
NQ_HD inline LaneK lane_consts(const Sched &S) { return lane_k(S.nsh + 4u, S.t[0], S.t[1], S.t[2], S.t[3]); }
NQ_HD inline LaneK idle_lane() { return LaneK{28u, 3u, 0u, 0u, 0u, 0u, vconst(kNlaBit), vconst(3u)}; }

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
// stacks of 16-byte frames {ld, rd, fc, frame word}, byte-addressed:
// slot i of a lane at base + i*S.  The slot below base holds the inert frame
// {0, 0, no free column, no candidates}: the top entry of an empty stack.  The
// walk keeps the address of its top entry (base - S when empty); a walk that
// ends takes the inert frame and stays on it, stepping without effect.
// ---------------------------------------------------------------------------
NQ_HD inline uint4 inert_frame() { return make_uint4(0u, 0u, 0u, 0u); }

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
    uint32_t ld, rd, fc, W;
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
// This is the step's definition, which the host engine runs as written; the
// kernels run the same step with a fixed instruction order
// (walk_step_paired below).
template <bool NODES, class Count, class Stack>
NQ_HD inline bool walk_step(const LaneK &L, Walk &w, Stack &st, Count &cnt, uint64_t &nodes)
{
    const uint32_t S = st.S;
    const uint4 top = st.read(w.A);
    const uint32_t W = w.W, F = w.fc;
    const uint32_t q = frame_q(L, W);                    // this frame's parameter byte (bits 24-31)
    const uint32_t bit = W & (0u - W);                   // lowest candidate (avail is the low field)
    const uint32_t rest = W ^ bit;
    const uint32_t s = (q >> 24) & 3u;
    const uint32_t x = (w.ld | bit) << s;                // the child's ld before its constants
    const uint32_t nld = bop3<kOrAnd>(x, q >> 26, L.three);          // | AL
    const uint32_t nrd = bop3<kOrAnd>((w.rd | bit) >> s, q >> L.ksh, L.km);
    const uint32_t nfc = F & ~bit;                       // the child's free columns
    uint32_t nf = bop3<kAndNotOr>(nfc, nld, nrd);
#if defined(__HIP_DEVICE_COMPILE__)
    asm("" : "+v"(nf));   // the child's word below reuses nf instead of refolding it
    uint32_t x2;
    // x << 1 as an add: v_add_u32 issues at the full rate, v_lshlrev_b32
    // at half (the compiler turns x + x back into a shift)
    asm("v_add_u32 %0, %1, %1" : "=v"(x2) : "v"(x));
#else
    const uint32_t x2 = x << 1;
#endif
    // the next row keeps a free square, or the frame has no prune (q's NLA bit)
    const bool lok = bop3<kOrAnd>(bop3<kAndNotOr>(nfc, x2, nrd >> 1), q, L.nlab) != 0u;
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
    const uint4 rem = make_uint4(w.ld, w.rd, F, rest);
    const uint32_t A = w.A;
    // the child: r - 1, candidates nf
    const uint32_t Wc = hiW + nf - (1u << kRShift);
    w.ld = down ? nld : (up ? top.x : w.ld);
    w.rd = down ? nrd : (up ? top.y : w.rd);
    w.fc = down ? nfc : (up ? top.z : F);
    w.W = down ? Wc : (up ? top.w : rest);
    w.A += push ? S : (pop ? 0u - S : 0u);
    // written after the selects: the wait for the top entry does not also
    // wait for this write
#if defined(__HIP_DEVICE_COMPILE__)
    asm volatile("" ::"v"(w.ld), "v"(w.rd), "v"(w.fc), "v"(w.W));
#endif
    if (push) st.write(A + S, rem);
    return up & !nonempty;
}

// ---------------------------------------------------------------------------
// walk_step with a fixed instruction order (the kernels' step)
// ---------------------------------------------------------------------------
// On gfx950, with four waves per SIMD, the simple 32-bit ops (add, sub, and,
// or, xor, right shifts, bitop3, mov, cndmask, add/sub with carry) issue in
// about 2.6 clocks per wave instruction when their neighbours are simple ops
// too; v_perm, v_bfe, v_lshlrev, every v_cmp, v_min/v_max, v_mul_u24,
// v_bfi and the three-input adds take about 4.3, and placed one at a time
// between simple ops about twice that.  The compiler's schedule of
// walk_step scatters eleven such ops over the step.  This form keeps three
// of them, in one dependent cluster (perm -> bfe -> lshl), and writes every
// condition as an add/sub carry-out into an SGPR pair (x != 0: the carry of
// x + 0xFFFFFFFF; b > a: the borrow of a - b; r != 0: the carry of the add
// that forms the child's word, (r << 28) + 0xF0000000), which the scalar
// unit then combines.  The parameter byte comes from one perm: W + W drops
// r's top bit and carries it out (r >= 8), the carry picks the pair of table
// words by two selects, and (W + W) >> 5 is the selector, byte 3 = r & 7.
// The step is two asm blocks: the body, and the selects, which the compiler
// places after its wait for the top entry; one block per step would need
// the frame in a register quadruple for the stack write.  Operands and
// results are exactly walk_step's.
struct PairK {
    uint32_t hi;      // 0xF0000000: the r field
    uint32_t S, nS;   // the stack's slot stride and its negative
};
#if defined(__HIP_DEVICE_COMPILE__)
template <bool NODES, class Stack>
__device__ inline void walk_step_paired(const LaneK &L, const PairK &K, Walk &w, Stack &st, uint32_t &cnt,
                                        uint64_t &nodes)
{
    const uint4 top = st.read(w.A);
    const uint32_t W = w.W, ld = w.ld, rd = w.rd, F = w.fc, A = w.A, E = st.empty();
    uint32_t nld, nrd, nfc, rest, Wc, ta, tb, tc, td, te, tf;
    uint64_t m1, m2, m3, up, down, push, pop;
    asm volatile(
        // byte lookup and the ld shift: the one cluster of slow ops
        "v_add_co_u32_e64 %[a], %[m1], %[W], %[W]\n\t"           // W << 1; the carry: r >= 8
        "v_cndmask_b32_e64 %[b], %[t0], %[t2], %[m1]\n\t"        // table words of r's half
        "v_lshrrev_b32 %[a], 5, %[a]\n\t"                        // selector, byte 3 = r & 7
        "v_cndmask_b32_e64 %[e], %[t1], %[t3], %[m1]\n\t"
        "v_sub_u32 %[d], 0, %[W]\n\t"
        "v_and_b32 %[d], %[W], %[d]\n\t"                         // bit
        "v_or_b32 %[f], %[ld], %[d]\n\t"
        "v_perm_b32 %[c], %[e], %[b], %[a]\n\t"                  // q: the byte in bits 24-31
        "v_bfe_u32 %[b], %[c], 24, 2\n\t"                        // s
        "v_lshlrev_b32 %[a], %[b], %[f]\n\t"                     // x = (ld | bit) << s
        // the child
        "v_or_b32 %[e], %[rd], %[d]\n\t"
        "v_lshrrev_b32 %[e], %[b], %[e]\n\t"
        "v_lshrrev_b32 %[b], 26, %[c]\n\t"
        "v_bitop3_b32 %[nld], %[a], %[b], %[three] bitop3:0xf8\n\t"
        "v_lshrrev_b32 %[b], %[ksh], %[c]\n\t"
        "v_bitop3_b32 %[nrd], %[e], %[b], %[km] bitop3:0xf8\n\t"
        "v_bitop3_b32 %[nfc], %[fc], %[d], %[d] bitop3:0x30\n\t"  // the child's free columns: fc & ~bit
        "v_xor_b32 %[rest], %[W], %[d]\n\t"
        // the one-row lookahead and the child's free set
        "v_add_u32 %[a], %[a], %[a]\n\t"
        "v_lshrrev_b32 %[b], 1, %[nrd]\n\t"
        "v_bitop3_b32 %[a], %[nfc], %[a], %[b] bitop3:0x10\n\t"   // free squares of the next row
        "v_bitop3_b32 %[a], %[a], %[c], %[nlab] bitop3:0xf8\n\t"  // | the no-prune flag: lookahead passes
        "v_bitop3_b32 %[b], %[nfc], %[nld], %[nrd] bitop3:0x10\n\t"  // nf
        "v_and_b32 %[d], %[W], %[hi]\n\t"                        // r << 28
        "v_add_co_u32_e64 %[e], %[m3], %[d], %[hi]\n\t"         // (r - 1) << 28; carries when r != 0 (inner)
        "v_add_u32 %[Wc], %[e], %[b]\n\t"                        // the child's word
        // conditions
        "v_add_co_u32_e64 %[e], %[m1], %[b], -1\n\t"             // nf != 0
        "v_add_co_u32_e64 %[e], %[m2], %[a], -1\n\t"             // lookahead
        "s_and_b64 %[m1], %[m1], %[m2]\n\t"                      // ok
        "v_sub_co_u32_e64 %[e], %[m2], %[d], %[rest]\n\t"        // more: rest > r << 28
        "s_and_b64 %[down], %[m1], %[m3]\n\t"
        "s_andn2_b64 %[m1], %[m1], %[m3]\n\t"                    // sol
        "v_sub_co_u32_e64 %[e], %[m3], %[E], %[A]\n\t"           // nonempty: A > E
        "s_and_b64 %[push], %[down], %[m2]\n\t"
        "s_nor_b64 %[up], %[down], %[m2]\n\t"
        "s_and_b64 %[pop], %[up], %[m3]\n\t"
        "v_addc_co_u32_e64 %[cnt], %[m2], 0, %[cnt], %[m1]"     // cnt += sol
        : [nld] "=&v"(nld), [nrd] "=&v"(nrd), [nfc] "=&v"(nfc), [rest] "=&v"(rest), [Wc] "=&v"(Wc),
          [a] "=&v"(ta), [b] "=&v"(tb), [c] "=&v"(tc), [d] "=&v"(td), [e] "=&v"(te), [f] "=&v"(tf),
          [m1] "=&s"(m1), [m2] "=&s"(m2), [m3] "=&s"(m3), [up] "=&s"(up), [down] "=&s"(down), [push] "=&s"(push),
          [pop] "=&s"(pop), [cnt] "+v"(cnt)
        : [W] "v"(W), [ld] "v"(ld), [rd] "v"(rd), [fc] "v"(F), [A] "v"(A), [E] "v"(E), [t0] "v"(L.t0),
          [t1] "v"(L.t1), [t2] "v"(L.t2), [t3] "v"(L.t3), [ksh] "v"(L.ksh), [km] "v"(L.km), [three] "v"(L.three),
          [nlab] "v"(L.nlab), [hi] "v"(K.hi)
        : "scc");
    if (NODES) nodes += (uint64_t)(W != 0u);             // the inert frame is no node
    uint32_t l1, r1, c1, w1, dA;
    asm volatile(
        "v_cndmask_b32_e64 %[dA], 0, %[nS], %[pop]\n\t"
        "v_cndmask_b32_e64 %[l1], %[ld], %[tx], %[up]\n\t"
        "v_cndmask_b32_e64 %[dA], %[dA], %[S], %[push]\n\t"
        "v_cndmask_b32_e64 %[r1], %[rd], %[ty], %[up]\n\t"
        "v_cndmask_b32_e64 %[lo], %[l1], %[nld], %[down]\n\t"
        "v_cndmask_b32_e64 %[c1], %[fc], %[tz], %[up]\n\t"
        "v_cndmask_b32_e64 %[ro], %[r1], %[nrd], %[down]\n\t"
        "v_cndmask_b32_e64 %[w1], %[rest], %[tw], %[up]\n\t"
        "v_cndmask_b32_e64 %[co], %[c1], %[nfc], %[down]\n\t"
        "v_add_u32 %[Ao], %[A], %[dA]\n\t"
        "v_cndmask_b32_e64 %[Wo], %[w1], %[Wc], %[down]"
        : [lo] "=&v"(w.ld), [ro] "=&v"(w.rd), [co] "=&v"(w.fc), [Wo] "=&v"(w.W), [Ao] "=&v"(w.A),
          [l1] "=&v"(l1), [r1] "=&v"(r1), [c1] "=&v"(c1), [w1] "=&v"(w1), [dA] "=&v"(dA)
        : [up] "s"(up), [down] "s"(down), [push] "s"(push), [pop] "s"(pop), [tx] "v"(top.x), [ty] "v"(top.y),
          [tz] "v"(top.z), [tw] "v"(top.w), [ld] "v"(ld), [rd] "v"(rd), [fc] "v"(F), [rest] "v"(rest),
          [nld] "v"(nld), [nrd] "v"(nrd), [nfc] "v"(nfc), [Wc] "v"(Wc), [A] "v"(A), [S] "v"(K.S), [nS] "v"(K.nS));
    if (__builtin_amdgcn_inverse_ballot_w64(push)) st.write(A + st.S, make_uint4(ld, rd, F, rest));
}
#else
// host pass of the kernels only (kernels never run on the host)
template <bool NODES, class Stack>
NQ_HD inline void walk_step_paired(const LaneK &L, const PairK &, Walk &w, Stack &st, uint32_t &cnt, uint64_t &nodes)
{
    (void)walk_step<NODES>(L, w, st, cnt, nodes);
}
#endif
inline NQ_HD PairK pair_consts(uint32_t S)
{
    return PairK{vconst(0xF0000000u), vconst(S), vconst(0u - S)};
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
