// nq_task.h -- decode of one 16-byte constellation record into the DFS
// descriptor the kernels run.  Compiled for both the host (tests) and gfx950.
//
// The descriptor replaces the reference's per-task SoA row
// (build_soa_for_range, 13Bit_codon/18Py_constellations_GPU_cuda_codon.py:698-1054)
// plus the per-frame func_meta dispatch of dfs_iter (18Py:368-562, table
// 18Py:1103-1156).  In the reference every frame looks its behaviour up from
// (functionid, row); but functionid changes only at fixed rows, so the
// behaviour of a frame is a function of its DFS depth alone.  The decode walks
// the functionid/row sequence once and stores one parameter byte per depth d,
// describing the move from a frame at depth d to its child:
//
//   bits 0-1  s     rows advanced by a placement (1, or 2/3 on a mark row)
//   bits 2-3  AL    constant OR-ed into the child's ld after the shift
//                   (blockL | add1); bit 0 is also set when the child is the
//                   jmark frame: the reference's "ld |= 1" on that frame
//                   (18Py:419-421), which takes column bit 0 from its
//                   candidates (its own children then see it shifted by
//                   their step; a jmark frame is never a mark frame); and on
//                   depth T-1 when the terminal functionid is 14 (see below).
//                   The one-row lookahead is computed before AL is OR-ed in,
//                   as the reference's parent does not see the jmark's bit
//   bits 4-5  ARc   constant OR-ed into the child's rd: ARc << (N-4), i.e.
//                   0, 1<<(N-4) or 1<<(N-3)
//   bit  6    NLA   no one-row lookahead prune at this depth (the prune runs
//                   when availptn is set and row+1 < endmark)
//   bit  7    JD    the child is the P5 switch frame: its rd gets bit N-1,
//                   the j queen's down-right diagonal entering the board
//                   (see the P5 note in decode)
//
// T is the depth of the terminal row (row == endmark on a base functionid);
// a child created at depth T counts one solution when its free set is
// non-empty (functionid 14 ignores column bit 0: "avail >> 1", 18Py:251-253).
#pragma once
#include <stdint.h>

#ifndef NQ_HD
#define NQ_HD __host__ __device__
#endif

namespace nqt {

constexpr int kMaxDepth = 24;        // parameter bytes per descriptor (6 words)
constexpr int kMinN = 6;
// u32 masks: the board fits the walk's 28-bit candidate field; build_soa's
// rd bit (N-1-j) + (N-2-start) is set only when j >= 2N-33-start, which keeps
// it at bit 30 or below (the P5 records set it later)
constexpr int kMaxN = 28;

enum : int32_t {
    kOk = 0,
    kErrP5 = 1,          // a P5 record whose switch row is not reached (never for a canonical record)
    kErrShift = 2,       // a shift of build_soa_for_range out of the u32 range
    kErrNoTerminal = 3,  // row sequence passes endmark without a base frame
    kErrDepth = 4,       // terminal depth beyond kMaxDepth
    kErrN = 5
};

// func_meta (next_funcid, funcptn, availptn), 18Py:1103-1132
NQ_HD inline int meta_next(int f)
{
    const uint8_t t[28] = {1, 2, 3, 3, 2, 6, 2, 2, 0, 4, 5, 7, 13, 14,
                           14, 14, 17, 14, 14, 20, 21, 21, 21, 25, 21, 21, 26, 26};
    return t[f];
}
NQ_HD inline int meta_ptn(int f)
{
    const uint8_t t[28] = {0, 1, 3, 5, 2, 0, 1, 2, 4, 4, 4, 4, 0, 1,
                           5, 2, 0, 1, 2, 0, 1, 5, 2, 0, 2, 1, 5, 0};
    return t[f];
}
NQ_HD inline int meta_avail(int f)
{
    const uint8_t t[28] = {0, 0, 1, 1, 0, 0, 0, 0, 1, 1, 1, 1, 0, 0,
                           1, 0, 0, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0};
    return t[f];
}
// blockL_by_funcid, 18Py:1227
NQ_HD inline int block_l(int f)
{
    const uint8_t t[28] = {0, 1, 0, 0, 1, 1, 0, 2, 0, 0, 0, 0, 0, 1,
                           0, 1, 1, 0, 2, 0, 0, 0, 1, 1, 2, 0, 0, 0};
    return t[f];
}
// FUNC_CATEGORY (18Py:1080-1089, 1228-1230) as the multiplier of 1<<(N-4):
// category 3 -> 2 (= 1<<(N-3)), category 4 -> 1 (= 1<<(N-4)), other -> 0
NQ_HD inline int block_k_code(int f)
{
    const uint8_t t[28] = {2, 0, 0, 0, 1, 0, 2, 2, 0, 0, 0, 0, 2, 0,
                           0, 1, 0, 2, 2, 2, 0, 0, 1, 0, 2, 2, 0, 2};
    return t[f];
}

struct Desc {
    uint32_t ld, rd, col, avail;  // root frame (depth 0), jmark already applied
    uint32_t tab[kMaxDepth / 4];  // parameter byte per depth
    uint32_t T;                   // terminal depth
    uint32_t tm;                  // terminal free mask: board, minus bit 0 for fid 14
    uint32_t wshift;              // weight = 1 << wshift (1, 2 or 3)
    int32_t status;
    // decoded reference fields, kept for checks and cost-model features
    int32_t fid, start, jmark, endmark, mark1, mark2;
    int32_t i, j, k, l;
    uint32_t free0;               // build_soa free set, before the jmark mask
};

NQ_HD inline uint32_t param_byte(const uint32_t *tab, uint32_t d)
{
    return (tab[d >> 2] >> ((d & 3u) * 8u)) & 0xFFu;
}

// symmetry weight, 18Py:1407-1408
NQ_HD inline uint32_t weight_shift(int ijkl, int N)
{
    int i = (ijkl >> 15) & 31, j = (ijkl >> 10) & 31, k = (ijkl >> 5) & 31, l = ijkl & 31;
    int r90 = ((N - 1 - k) << 15) + ((N - 1 - l) << 10) + (j << 5) + i;
    if (r90 == ijkl) return 1;
    if (i == N - 1 - j && k == N - 1 - l) return 2;
    return 3;
}

// gate_bits: the rd word width build_soa's j-diagonal gate assumes (32, the
// u32 kernels; the host tests lower it to put P5 records on small boards --
// any width >= N counts the same).
NQ_HD inline void decode(int N, uint32_t r_ld, uint32_t r_rd, uint32_t r_col, uint32_t r_sij, Desc &D,
                         int gate_bits = 32)
{
    for (int a = 0; a < kMaxDepth / 4; a++) D.tab[a] = 0;
    D.status = kOk;
    D.T = 0;
    D.ld = D.rd = D.col = D.avail = 0;
    D.free0 = 0;
    D.fid = D.start = D.jmark = D.endmark = D.mark1 = D.mark2 = 0;
    D.i = D.j = D.k = D.l = 0;
    D.tm = 0;
    D.wshift = 0;
    if (N < kMinN || N > kMaxN) { D.status = kErrN; return; }
    const uint32_t mask = (1u << N) - 1u;
    const uint32_t small = (1u << (N - 2)) - 1u;
    const int N1 = N - 1, N2 = N - 2;
    const int start = (int)(r_sij >> 20);
    const int ijkl = (int)(r_sij & 0xFFFFFu);
    const int i = (ijkl >> 15) & 31, j = (ijkl >> 10) & 31, k = (ijkl >> 5) & 31, l = ijkl & 31;
    D.i = i; D.j = j; D.k = k; D.l = l; D.start = start;
    // ---- start state, 18Py:777-815 ----
    uint32_t ld = r_ld >> 1, rd = r_rd >> 1;
    uint32_t col = ((r_col >> 1) | ~small) & mask;
    if (start > N || N1 - j < 0 || N1 - l < 0) { D.status = kErrShift; return; }
    const uint32_t LD = (1u << (N1 - j)) | (1u << (N1 - l));
    ld |= (N - start) >= 32 ? 0u : (LD >> (N - start));
    if (start > k) {
        int sh = N1 - (start - k + 1);
        if (sh < 0) { D.status = kErrShift; return; }
        rd |= 1u << sh;
    }
    // j >= 2N-33-start at 32 bits: (N1-j) + (N2-start) < gate_bits
    const bool j_gate = j >= 2 * N - 1 - gate_bits - start;
    if (j_gate) {
        int sh = (N1 - j) + (N2 - start);
        if (N2 - start < 0 || sh > 31) { D.status = kErrShift; return; }
        rd |= 1u << sh;
    }
    uint32_t fr = mask & ~(ld | rd | col);
    D.free0 = fr;
    // ---- 28-leaf target tree, 18Py:820-1031 ----
    int jmark = 0, mark1 = 0, mark2 = 0, endmark = 0, target = 0;
    const bool k_lt_l = k < l, s_lt_k = start < k, s_lt_l = start < l;
    const bool l_eq_kp1 = l == k + 1, k_eq_lp1 = k == l + 1;
    if (j < N - 3) {
        jmark = j + 1; endmark = N2;
        if (j_gate) {
            if (k_lt_l) {
                mark1 = k - 1; mark2 = l - 1;
                target = s_lt_l ? (s_lt_k ? (l_eq_kp1 ? 4 : 0) : 1) : 2;
            } else {
                mark1 = l - 1; mark2 = k - 1;
                target = s_lt_k ? (s_lt_l ? (k_eq_lp1 ? 7 : 5) : 6) : 2;
            }
        } else {
            if (k_lt_l) { mark1 = k - 1; mark2 = l - 1; target = l_eq_kp1 ? 9 : 8; }
            else { mark1 = l - 1; mark2 = k - 1; target = k_eq_lp1 ? 11 : 10; }
        }
    } else if (j == N - 3) {
        endmark = N2;
        if (k_lt_l) {
            mark1 = k - 1; mark2 = l - 1;
            target = s_lt_l ? (s_lt_k ? (l_eq_kp1 ? 15 : 12) : 13) : 14;
        } else {
            mark1 = l - 1; mark2 = k - 1;
            target = s_lt_k ? (s_lt_l ? (k_eq_lp1 ? 18 : 16) : 17) : 14;
        }
    } else if (j == N - 2) {
        if (k_lt_l) {
            endmark = N2;
            if (s_lt_l) {
                if (s_lt_k) {
                    mark1 = k - 1;
                    if (!l_eq_kp1) { mark2 = l - 1; target = 19; }
                    else target = 22;
                } else { mark2 = l - 1; target = 20; }
            } else target = 21;
        } else {
            if (s_lt_k) {
                if (s_lt_l) {
                    if (k < N2) {
                        mark1 = l - 1; endmark = N2;
                        if (!k_eq_lp1) { mark2 = k - 1; target = 23; }
                        else target = 24;
                    } else if (l != N - 3) { mark2 = l - 1; endmark = N - 3; target = 20; }
                    else { endmark = N - 4; target = 21; }
                } else if (k != N2) { mark2 = k - 1; endmark = N2; target = 25; }
                else { endmark = N - 3; target = 21; }
            } else { endmark = N2; target = 21; }
        }
    } else {
        endmark = N2;
        if (start > k) target = 26;
        else { mark1 = k - 1; target = 27; }
    }
    D.fid = target; D.jmark = jmark; D.endmark = endmark; D.mark1 = mark1; D.mark2 = mark2;
    D.wshift = weight_shift(ijkl, N);
    // ---- per-depth schedule: the functionid/row walk of dfs_iter ----
    // P5 (functionids 8-11, the "SQBjl*" of
    // 13Bit_codon/13Py_constellations_codon.py:620-678): records whose j
    // diagonal failed the gate above start before that diagonal reaches the
    // board.  At row N-1-jmark (= N-2-j) the frame takes it -- rd |= 1 << (N-1),
    // column N-1 no longer free -- and continues as next_funcid on the same
    // row; the rows above are placed without the one-row lookahead, as the
    // SQBjl loops do.  18Py's dfs_iter has no such branch (its P5 frames never
    // switch, 18Py:368-562), which is why its N=24 total is wrong (18Py:49).
    // Here the bit comes from the byte of the switch frame's parent (JD), or
    // is applied to the root when the root is the switch frame.
    int f = target, row = start, T = -1;
    uint8_t qb[kMaxDepth] = {};
    bool root_jm = false, root_jd = false;
    const int jd_row = N1 - jmark;
    for (int d = 0; d < kMaxDepth;) {
        const int ptn = meta_ptn(f);
        if (ptn == 5 && row == endmark) { T = d; break; }
        if (ptn == 4 && row == jd_row) {
            if (d == 0) root_jd = true;
            else qb[d - 1] |= 0x80u;
            f = meta_next(f);
            continue;   // the same frame, as next_funcid
        }
        uint32_t s = 1, AL = 0, ARc = 0, LA = 0;
        if (ptn == 0 || ptn == 1 || ptn == 2) {
            const bool at = (ptn == 1) ? (row == mark2) : (row == mark1);
            if (at) {
                s = (ptn == 2) ? 3u : 2u;
                AL = (uint32_t)block_l(f) | ((ptn == 1 && f == 20) ? 1u : 0u);
                ARc = (uint32_t)block_k_code(f);
                f = meta_next(f);
            }
            // mark functionids have availptn 0: no lookahead on their plain rows
        } else if (ptn == 3 && row == jmark) {
            // this frame's ld |= 1: from its parent's AL bit 0, or on the root
            if (d == 0) root_jm = true;
            else qb[d - 1] |= 0x04u;
            LA = (meta_avail(f) && row + 1 < endmark) ? 1u : 0u;
            f = meta_next(f);
        } else if (ptn == 4) {
            if (row > jd_row) { D.status = kErrP5; return; }   // past its switch row
        } else {
            LA = (meta_avail(f) && row + 1 < endmark) ? 1u : 0u;
        }
        qb[d] = (uint8_t)(s | (AL << 2) | (ARc << 4) | ((LA ^ 1u) << 6));
        row += (int)s;
        d++;
        if (row > N + 3) break;
    }
    if (T < 0) { D.status = kErrNoTerminal; return; }
    if (T >= kMaxDepth - 1) { D.status = kErrDepth; return; }
    D.T = (uint32_t)T;
    D.tm = (f == 14) ? (mask & ~1u) : mask;
    // functionid 14 counts a terminal child only for columns other than 0:
    // the same as a jmark on the terminal row (which the row walk never
    // produces itself), so the walk needs no separate terminal mask
    if (f == 14 && T > 0) qb[T - 1] |= 0x04u;
    for (int d = 0; d < kMaxDepth; d++) D.tab[d >> 2] |= (uint32_t)qb[d] << ((d & 3) * 8);
    // the root as the switch frame (its j diagonal) and as the jmark frame
    if (root_jd) { fr &= ~(1u << N1); rd |= 1u << N1; }
    if (root_jm) { fr &= ~1u; ld |= 1u; }
    D.ld = ld; D.rd = rd; D.col = col; D.avail = fr;
}

// One more preset queen (set_pre_queens, 18Py:1535-1613, one level): a
// record of preset p stands on its start row, which is not one of the
// pre-placed rows k and l; every free column of that row gives a record of
// preset p + 1, whose start row is the next row below that is not k or l
// (the rows between move ld/rd by one each, 18Py:1591-1593).  The free set
// is ~(ld | rd | col | LD >> (N-1-row) | RD << (N-1-row)) on the board
// (18Py:1606-1611), LD and RD the j/l and j/k diagonals of the start
// constellation (18Py:1749-1750).  emit(ld, rd, col, startijkl) receives the
// children with columns ascending.  Returns kOk, kErrShift when the start row
// is a pre-placed row or lies below the board, or kErrOverflow when a
// child's col or start row does not fit the u32 record.
constexpr int32_t kErrOverflow = 6;
template <class Emit>
NQ_HD inline int32_t preset_children(int N, uint32_t r_ld, uint32_t r_rd, uint32_t r_col, uint32_t r_sij,
                                     Emit &&emit)
{
    if (N < kMinN || N > kMaxN) return kErrN;
    const int row = (int)(r_sij >> 20);
    const uint32_t ijkl = r_sij & 0xFFFFFu;
    const int j = (int)((ijkl >> 10) & 31u), k = (int)((ijkl >> 5) & 31u), l = (int)(ijkl & 31u);
    const int sh = N - 1 - row;
    if (row == k || row == l || sh < 0 || j > N - 1 || l > N - 1 || k > N - 1) return kErrShift;
    const uint64_t L = 1ull << (N - 1), mask = (1ull << N) - 1u;
    const uint64_t LD = (L >> j) | (L >> l), RD = (L >> j) | (1ull << k);
    uint64_t fr = ~((uint64_t)r_ld | r_rd | r_col | (LD >> sh) | (RD << sh)) & mask;
    int32_t status = kOk;
    while (fr) {
        const uint64_t b = fr & (0ull - fr);
        fr ^= b;
        uint64_t nld = ((uint64_t)r_ld | b) << 1, nrd = ((uint64_t)r_rd | b) >> 1;
        const uint64_t ncol = (uint64_t)r_col | b;
        int nrow = row + 1;
        while (nrow == k || nrow == l) { nld <<= 1; nrd >>= 1; nrow++; }
        // ld bits at 32 and above are off the board (N <= 28 after the
        // decode's ld >> 1) and only move further up: the u32 record drops them
        if (ncol >> 32 || nrow > 31) { status = kErrOverflow; continue; }
        emit((uint32_t)nld, (uint32_t)nrd, (uint32_t)ncol, ((uint32_t)nrow << 20) | ijkl);
    }
    return status;
}

// The per-record schedule the kernels read (2 x 16 bytes):
//   words 0-3   byte r (r = 0..15) = parameter byte of depth T-1-r: the byte
//               of a frame with r = T-1-depth (T <= 16, so depth 0 is byte
//               T-1); bytes r >= T are zero
//   word 4      T | wshift << 8 | N << 16
//   word 5      terminal free mask (used for a root on the terminal row)
//   word 6      decode status
NQ_HD inline void sched_words(const Desc &D, int N, uint32_t w[8])
{
    for (int a = 0; a < 8; a++) w[a] = 0;
    for (uint32_t r = 0; r < 16u && r < D.T; r++)
        w[r >> 2] |= param_byte(D.tab, D.T - 1u - r) << ((r & 3u) * 8u);
    w[4] = D.T | (D.wshift << 8) | ((uint32_t)N << 16);
    w[5] = D.tm;
    w[6] = (uint32_t)D.status;
}

}  // namespace nqt
