/*
 * nq_oracle.c -- CPU restatement of the N-Queens constellation count path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: it is loaded by the
 * tests, by the package smoke check and by the CPU-baseline leg of bench.py.
 * The product (the HIP library under n-queens_amd/) never links, loads or
 * calls it.
 *
 * Every function restates one step of the reference solver
 * 13Bit_codon/18Py_constellations_GPU_cuda_codon.py (Codon, 64-bit ints);
 * the citations are file:line in that file unless stated otherwise.  All
 * masks are uint64_t, matching Codon's 64-bit `int`, so that the u32
 * arithmetic of the GPU kernels is checked against the wider semantics
 * rather than assumed.
 *
 * Parity pins (see tests/test_oracle.py):
 *   gen_constellations / set_pre_queens / jasmin / check_rotations
 *       -> constellations_N{6..21}_5.bin record multisets (exact)
 *   build_soa + dfs_iter + symmetry weight
 *       -> solutions_N{6..19}_5_v2.bin per-record weighted counts (exact)
 *       -> A000170 totals; brute force for N <= 12
 *   symmetric-solution counts (F90, F180)
 *       -> brute force for N <= 12 and the published Unique table for
 *          N = 4..21 through unique = (T + F180 + 2*F90) / 8
 *   node counts: no reference log exists per record; pinned by the
 *       identity nodes == sum over frames of popcount(avail) computed by an
 *       independent recursive restatement (nqo_count_records_recursive).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define NQO_EXPORT __attribute__((visibility("default")))

/* ------------------------------------------------------------------ */
/* A1: ijkl packing and board symmetries (18Py:1396-1412)             */
/* ------------------------------------------------------------------ */
static inline int geti(int x) { return (x >> 15) & 31; }
static inline int getj(int x) { return (x >> 10) & 31; }
static inline int getk(int x) { return (x >> 5) & 31; }
static inline int getl(int x) { return x & 31; }
static inline int to_ijkl(int i, int j, int k, int l) { return (i << 15) + (j << 10) + (k << 5) + l; }
/* 18Py:1405 */
static inline int rot90(int x, int N)
{
    return ((N - 1 - getk(x)) << 15) + ((N - 1 - getl(x)) << 10) + (getj(x) << 5) + geti(x);
}
/* 18Py:1397 */
static inline int mirvert(int x, int N)
{
    return to_ijkl(N - 1 - geti(x), N - 1 - getj(x), getl(x), getk(x));
}
static inline int imin(int a, int b) { return a < b ? a : b; }
/* 18Py:1408 */
static inline int symmetry90(int x, int N) { return x == rot90(x, N); }
/* A3, 18Py:1407: weight 2 (90-degree fixed), 4 (180-degree fixed), else 8 */
static inline int symmetry_w(int x, int N)
{
    if (symmetry90(x, N)) return 2;
    if (geti(x) == N - 1 - getj(x) && getk(x) == N - 1 - getl(x)) return 4;
    return 8;
}
/* 18Py:1429-1453: rotate the border queen nearest an edge onto the j side,
 * then mirror so that j sits in the right half. */
static int jasmin(int x, int N)
{
    int arg = 0;
    int mn = imin(getj(x), N - 1 - getj(x));
    if (imin(geti(x), N - 1 - geti(x)) < mn) { arg = 2; mn = imin(geti(x), N - 1 - geti(x)); }
    if (imin(getk(x), N - 1 - getk(x)) < mn) { arg = 3; mn = imin(getk(x), N - 1 - getk(x)); }
    if (imin(getl(x), N - 1 - getl(x)) < mn) { arg = 1; mn = imin(getl(x), N - 1 - getl(x)); }
    for (int t = 0; t < arg; t++) x = rot90(x, N);
    if (getj(x) < N - 1 - getj(x)) x = mirvert(x, N);
    return x;
}

NQO_EXPORT int nqo_jasmin(int ijkl, int N) { return jasmin(ijkl, N); }
NQO_EXPORT int nqo_symmetry(int ijkl, int N) { return symmetry_w(ijkl, N); }

/* ------------------------------------------------------------------ */
/* insertion-ordered set of ijkl values (ijkl < 2^20)                 */
/* ------------------------------------------------------------------ */
typedef struct {
    uint8_t *in;   /* membership, 2^20 bytes */
    int *v;        /* insertion order */
    int n, cap;
} ijset;

static void ijset_init(ijset *s)
{
    s->in = (uint8_t *)calloc(1u << 20, 1);
    s->cap = 1024; s->n = 0;
    s->v = (int *)malloc(sizeof(int) * s->cap);
}
static void ijset_free(ijset *s) { free(s->in); free(s->v); }
static inline int ijset_has(const ijset *s, int x) { return s->in[x & 0xFFFFF]; }
static void ijset_add(ijset *s, int x)
{
    if (s->in[x]) return;
    s->in[x] = 1;
    if (s->n == s->cap) { s->cap *= 2; s->v = (int *)realloc(s->v, sizeof(int) * s->cap); }
    s->v[s->n++] = x;
}
/* 18Py:1411-1412: any of the 90/180/270-degree images already present */
static int check_rotations(const ijset *s, int i, int j, int k, int l, int N)
{
    int r90 = ((N - 1 - k) << 15) + ((N - 1 - l) << 10) + (j << 5) + i;
    int r180 = ((N - 1 - j) << 15) + ((N - 1 - i) << 10) + ((N - 1 - l) << 5) + (N - 1 - k);
    int r270 = (l << 15) + (k << 10) + ((N - 1 - i) << 5) + (N - 1 - j);
    return ijset_has(s, r90) || ijset_has(s, r180) || ijset_has(s, r270);
}

/* ------------------------------------------------------------------ */
/* hash set of set_pre_queens states (the subconst cache, 18Py:1462)  */
/* ------------------------------------------------------------------ */
typedef struct { uint64_t ld, rd, col, LD, RD, small; int owner; } skey;
typedef struct {
    skey *t; uint8_t *used; size_t cap, n;
} sset;

static uint64_t mix64(uint64_t x)
{
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
    return x ^ (x >> 31);
}
static uint64_t skey_hash(const skey *k)
{
    uint64_t h = mix64(k->ld + 0x9E3779B97F4A7C15ULL);
    h = mix64(h ^ k->rd); h = mix64(h ^ k->col); h = mix64(h ^ k->LD);
    h = mix64(h ^ k->RD); h = mix64(h ^ k->small);
    return h;
}
static int skey_eq(const skey *a, const skey *b)
{
    return a->ld == b->ld && a->rd == b->rd && a->col == b->col && a->LD == b->LD &&
           a->RD == b->RD && a->small == b->small;
}
static void sset_init(sset *s, size_t cap)
{
    s->cap = cap; s->n = 0;
    s->t = (skey *)malloc(sizeof(skey) * cap);
    s->used = (uint8_t *)calloc(cap, 1);
}
static void sset_free(sset *s) { free(s->t); free(s->used); s->t = NULL; s->used = NULL; }
static void sset_clear(sset *s) { memset(s->used, 0, s->cap); s->n = 0; }
/* returns pointer to an existing entry, or NULL after inserting */
static skey *sset_find_or_add(sset *s, const skey *k)
{
    if (2 * (s->n + 1) > s->cap) {
        sset o = *s;
        sset_init(s, o.cap * 2);
        for (size_t a = 0; a < o.cap; a++)
            if (o.used[a]) {
                size_t p = skey_hash(&o.t[a]) & (s->cap - 1);
                while (s->used[p]) p = (p + 1) & (s->cap - 1);
                s->used[p] = 1; s->t[p] = o.t[a]; s->n++;
            }
        sset_free(&o);
    }
    size_t p = skey_hash(k) & (s->cap - 1);
    while (s->used[p]) {
        if (skey_eq(&s->t[p], k)) return &s->t[p];
        p = (p + 1) & (s->cap - 1);
    }
    s->used[p] = 1; s->t[p] = *k; s->n++;
    return NULL;
}

/* ------------------------------------------------------------------ */
/* A2: gen_constellations / set_pre_queens (18Py:1462-1785)           */
/* ------------------------------------------------------------------ */
typedef struct {
    int N, preset, k, l, sc_index, scope;
    uint64_t LD, RD, mask;
    sset *cache;      /* 18Py keeps one cache for the whole call (cleared at :1681) */
    sset *sigs;       /* per-sc signature set, preset <= 5 (:1595-1604) */
    int nosig;        /* record expansion: every leaf is kept */
    uint64_t *rec;    /* 4 x u64 per record: ld, rd, col, startijkl */
    size_t nrec, caprec;
    long cache_hits, cache_hits_cross_sc, sig_hits;
} genctx;

static void emit(genctx *g, uint64_t ld, uint64_t rd, uint64_t col, uint64_t startijkl)
{
    if (g->nrec == g->caprec) {
        g->caprec = g->caprec ? g->caprec * 2 : 4096;
        g->rec = (uint64_t *)realloc(g->rec, sizeof(uint64_t) * 4 * g->caprec);
    }
    uint64_t *r = g->rec + 4 * g->nrec++;
    r[0] = ld; r[1] = rd; r[2] = col; r[3] = startijkl;
}

static void set_pre_queens(genctx *g, uint64_t ld, uint64_t rd, uint64_t col, int row, int queens);

/* 18Py:1462-1531: the cache key is the whole state (ld,rd,col,k,l,row,queens,LD,RD,N,preset) */
static void set_pre_queens_cached(genctx *g, uint64_t ld, uint64_t rd, uint64_t col, int row, int queens)
{
    if (g->scope == 2) { set_pre_queens(g, ld, rd, col, row, queens); return; }
    skey key;
    key.ld = ld; key.rd = rd; key.col = col; key.LD = g->LD; key.RD = g->RD;
    key.small = (uint64_t)g->k | ((uint64_t)g->l << 8) | ((uint64_t)row << 16) |
                ((uint64_t)queens << 24) | ((uint64_t)g->N << 32) | ((uint64_t)g->preset << 40);
    key.owner = g->sc_index;
    skey *hit = sset_find_or_add(g->cache, &key);
    if (hit) {
        g->cache_hits++;
        if (hit->owner != g->sc_index) g->cache_hits_cross_sc++;
        return;
    }
    set_pre_queens(g, ld, rd, col, row, queens);
}

/* 18Py:1535-1613 (use_visited_prune is False at :1533, so the zobrist
 * visited set never prunes and is not restated). */
static void set_pre_queens(genctx *g, uint64_t ld, uint64_t rd, uint64_t col, int row, int queens)
{
    const int N = g->N;
    if (row == g->k || row == g->l) {                      /* :1591 skip rows k and l */
        set_pre_queens_cached(g, ld << 1, rd >> 1, col, row + 1, queens);
        return;
    }
    if (queens == g->preset) {                             /* :1594 leaf */
        if (g->preset <= 5 && !g->nosig) {                 /* :1596 signature dedup */
            skey sig;
            sig.ld = ld; sig.rd = rd; sig.col = col; sig.LD = 0; sig.RD = 0;
            sig.small = (uint64_t)g->k | ((uint64_t)g->l << 8) | ((uint64_t)row << 16);
            sig.owner = g->sc_index;
            if (sset_find_or_add(g->sigs, &sig)) { g->sig_hits++; return; }
        }
        emit(g, ld, rd, col, (uint64_t)row << 20);
        return;
    }
    /* :1606 free = ~(ld|rd|col|(LD>>(N-1-row))|(RD<<(N-1-row))) & board_mask */
    uint64_t fr = ~(ld | rd | col | (g->LD >> (N - 1 - row)) | (g->RD << (N - 1 - row))) & g->mask;
    while (fr) {
        uint64_t bit = fr & (0 - fr);
        fr &= fr - 1;
        set_pre_queens_cached(g, (ld | bit) << 1, (rd | bit) >> 1, col | bit, row + 1, queens + 1);
    }
}

/*
 * nqo_gen_constellations: 18Py:1622-1785 (gen_constellations).
 * Writes u32 records {ld, rd, col, startijkl} in generation order.
 * Returns the record count, or -1 when a value does not fit 32 bits.
 * stats (may be NULL): [0] canonical start constellations, [1] cache hits,
 * [2] cache hits whose first visitor was another start constellation,
 * [3] signature hits.  Records are malloc'd into *out (caller frees with
 * nqo_free).
 * scope selects the lifetime of the subconst cache: 0 = the whole call, as
 * 18Py:1681 (clear once); 1 = one start constellation, as
 * 14Bit_codon_GPU/359Py_pushbranch_partial.py:2919 (clear per sc).  They
 * differ from preset 6 on, where states collide across start
 * constellations and scope 0 drops records (18Py's totals go wrong there).
 * scope 2 disables the cache: every path is kept, which is the exact
 * enumeration (equal to scope 1 whenever no state repeats within one start
 * constellation, as for every N <= 21 at presets 4..6).
 */
NQO_EXPORT long nqo_gen_constellations_scoped(int N, int preset, int scope, uint32_t **out, long *stats)
{
    const int halfN = (N + 1) / 2, N1 = N - 1, N2 = N - 2;
    ijset S; ijset_init(&S);
    if (N % 2 == 1) {                                      /* :1687-1699 odd-N centre block */
        int center = N / 2;
        for (int l = center + 1; l < N1; l++)
            for (int i = center + 1; i < N1; i++) {
                if (i == N1 - l) continue;
                for (int j = N - center - 2; j > 0; j--) {
                    if (j == i || j == l) continue;
                    if (!check_rotations(&S, i, j, center, l, N)) ijset_add(&S, to_ijkl(i, j, center, l));
                }
            }
    }
    for (int k = 1; k < halfN; k++)                        /* :1703-1712 no queen in a corner */
        for (int l = k + 1; l < N1; l++)
            for (int i = k + 1; i < N1; i++) {
                if (i == N1 - l) continue;
                for (int j = N - k - 2; j > 0; j--) {
                    if (j == i || j == l) continue;
                    if (!check_rotations(&S, i, j, k, l, N)) ijset_add(&S, to_ijkl(i, j, k, l));
                }
            }
    for (int j = 1; j < N2; j++)                           /* :1716 queen in the corner */
        for (int l = j + 1; l < N1; l++) ijset_add(&S, to_ijkl(0, j, 0, l));
    ijset J; ijset_init(&J);                               /* :1719 Jasmin normal forms */
    for (int a = 0; a < S.n; a++) ijset_add(&J, jasmin(S.v[a], N));
    ijset_free(&S);

    genctx g; memset(&g, 0, sizeof g);
    sset cache, sigs;
    sset_init(&cache, 1u << 16); sset_init(&sigs, 1u << 10);
    g.N = N; g.preset = preset; g.cache = &cache; g.sigs = &sigs; g.scope = scope;
    g.mask = (N >= 64) ? ~0ULL : ((1ULL << N) - 1);
    const uint64_t L = 1ULL << N1;
    for (int s = 0; s < J.n; s++) {
        int sc = J.v[s];
        int i = geti(sc), j = getj(sc), k = getk(sc), l = getl(sc);
        sset_clear(&sigs);                                 /* :1729 signatures per sc */
        if (scope == 1) sset_clear(&cache);                /* 359Py:2919 cache per sc */
        uint64_t ld = ((i > 0) ? (L >> (i - 1)) : 0) | (1ULL << (N - k));   /* :1737-1750 */
        uint64_t rd = (L >> (i + 1)) | (1ULL << (l - 1));
        uint64_t col = 1 | L | (L >> i) | (L >> j);
        g.LD = (L >> j) | (L >> l);
        g.RD = (L >> j) | (1ULL << k);
        g.k = k; g.l = l; g.sc_index = s;
        size_t before = g.nrec;
        set_pre_queens_cached(&g, ld, rd, col, 1, (j == N1) ? 3 : 4);
        uint64_t base = (uint64_t)to_ijkl(i, j, k, l);     /* :1777-1781 */
        for (size_t a = before; a < g.nrec; a++) g.rec[4 * a + 3] |= base;
    }
    if (stats) {
        stats[0] = J.n; stats[1] = g.cache_hits; stats[2] = g.cache_hits_cross_sc; stats[3] = g.sig_hits;
    }
    ijset_free(&J); sset_free(&cache); sset_free(&sigs);
    uint32_t *o = (uint32_t *)malloc(sizeof(uint32_t) * 4 * (g.nrec ? g.nrec : 1));
    long rc = (long)g.nrec;
    for (size_t a = 0; a < 4 * g.nrec; a++) {
        if (g.rec[a] >> 32) rc = -1;
        o[a] = (uint32_t)g.rec[a];
    }
    free(g.rec);
    *out = o;
    return rc;
}

/*
 * nqo_expand_records: the records of preset p + levels that set_pre_queens
 * (18Py:1535-1613) reaches from each given record of preset p, continuing
 * its recursion from the record's state (start row, masks) with the
 * diagonals LD/RD of its border queens (:1749-1750) and no cache or
 * signature dedup; records in input order, columns ascending within one.
 * Writes at most cap records to out (u32 {ld, rd, col, startijkl}); ld
 * bits at 32 and above are dropped (after build_soa's ld >> 1 they lie at
 * bit 31 or above, off every board up to N = 31, and ld only shifts up).
 * Returns the record count, -1 when another field does not fit 32 bits, -2
 * when cap is too small.
 */
NQO_EXPORT long nqo_expand_records(int N, const uint32_t *rec, long m, int levels, uint32_t *out, long cap)
{
    genctx g; memset(&g, 0, sizeof g);
    g.N = N; g.preset = levels; g.scope = 2; g.nosig = 1;
    g.mask = (1ULL << N) - 1;
    const uint64_t L = 1ULL << (N - 1);
    for (long a = 0; a < m; a++) {
        const uint32_t *r = rec + 4 * a;
        const int ijkl = (int)(r[3] & 0xFFFFF), row = (int)(r[3] >> 20);
        const int j = getj(ijkl), k = getk(ijkl), l = getl(ijkl);
        g.LD = (L >> j) | (L >> l);
        g.RD = (L >> j) | (1ULL << k);
        g.k = k; g.l = l;
        size_t before = g.nrec;
        set_pre_queens(&g, r[0], r[1], r[2], row, 0);
        for (size_t b = before; b < g.nrec; b++) g.rec[4 * b + 3] |= (uint64_t)ijkl;
    }
    long rc = (long)g.nrec;
    if (rc > cap) rc = -2;
    for (size_t b = 0; rc >= 0 && b < 4 * g.nrec; b++) {
        if ((b & 3) != 0 && (g.rec[b] >> 32)) rc = -1;
        else out[b] = (uint32_t)g.rec[b];
    }
    free(g.rec);
    return rc;
}

NQO_EXPORT long nqo_gen_constellations(int N, int preset, uint32_t **out, long *stats)
{
    return nqo_gen_constellations_scoped(N, preset, 0, out, stats);
}

NQO_EXPORT void nqo_free(void *p) { free(p); }

/* ------------------------------------------------------------------ */
/* A5: per-function-id behaviour (18Py:1080-1156)                     */
/* ------------------------------------------------------------------ */
/* func_meta = (next_funcid, funcptn, availptn), 18Py:1103-1132 */
static const int META_NEXT[28] = {1, 2, 3, 3, 2, 6, 2, 2, 0, 4, 5, 7, 13, 14,
                                  14, 14, 17, 14, 14, 20, 21, 21, 21, 25, 21, 21, 26, 26};
static const int META_PTN[28] = {0, 1, 3, 5, 2, 0, 1, 2, 4, 4, 4, 4, 0, 1,
                                 5, 2, 0, 1, 2, 0, 1, 5, 2, 0, 2, 1, 5, 0};
static const int META_AVAIL[28] = {0, 0, 1, 1, 0, 0, 0, 0, 1, 1, 1, 1, 0, 0,
                                   1, 0, 0, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0};
/* blockL_by_funcid, 18Py:1227 */
static const int BLOCK_L[28] = {0, 1, 0, 0, 1, 1, 0, 2, 0, 0, 0, 0, 0, 1,
                                0, 1, 1, 0, 2, 0, 0, 0, 1, 1, 2, 0, 0, 0};
/* FUNC_CATEGORY x FID (18Py:1080-1100, 1228-1230): 3 -> 1<<(N-3), 4 -> 1<<(N-4), else 0 */
static const int BLOCK_K_CAT[28] = {3, 0, 0, 0, 4, 0, 3, 3, 0, 0, 0, 0, 3, 0,
                                    0, 4, 0, 3, 3, 3, 0, 0, 4, 0, 3, 3, 0, 3};

/* ------------------------------------------------------------------ */
/* A4: build_soa_for_range (18Py:698-1054)                            */
/* ------------------------------------------------------------------ */
typedef struct {
    uint64_t ld, rd, col, free;
    int start, fid, jmark, endmark, mark1, mark2, w, ijkl;
} task_t;

/* Width of the rd word build_soa's gate assumes: the j queen's rd bit is set
 * at the start row when its position (N-1-j) + (N-2-start) is below it
 * (18Py:814-815, j >= 2N-33-start for 32 bits), and the target tree sends
 * the other records to the P5 functions (j > 2N-34-start, 18Py:849), which
 * set it when it reaches the board.  32 is the reference; the tests lower it
 * (any width >= N-1 must give the same counts) to put P5 records on small
 * boards. */
static int g_gate_bits = 32;
NQO_EXPORT void nqo_set_gate_bits(int w) { g_gate_bits = w; }

static int build_task(int N, const uint32_t *rec, task_t *t)
{
    const uint64_t board_mask = (1ULL << N) - 1;
    const uint64_t small_mask = (N >= 2) ? ((1ULL << (N - 2)) - 1) : 0;
    const int N1 = N - 1, N2 = N - 2;
    int jmark = 0, mark1 = 0, mark2 = 0, endmark = 0, target = 0;
    uint64_t start_ijkl = rec[3];
    int start = (int)(start_ijkl >> 20);
    int ijkl = (int)(start_ijkl & ((1u << 20) - 1));
    int j = getj(ijkl), k = getk(ijkl), l = getl(ijkl);
    uint64_t ld = (uint64_t)rec[0] >> 1;
    uint64_t rd = (uint64_t)rec[1] >> 1;
    uint64_t col = (((uint64_t)rec[2] >> 1) | ~small_mask) & board_mask;
    uint64_t LD = (1ULL << (N1 - j)) | (1ULL << (N1 - l));
    if (N - start < 0 || N - start > 63) return -1;
    ld |= LD >> (N - start);
    if (start > k) {
        int sh = N1 - (start - k + 1);
        if (sh < 0) return -1;
        rd |= 1ULL << sh;
    }
    /* (N1-j) + (N2-start) <= gate_bits - 1, i.e. j >= 2N-33-start at 32 bits */
    const int j_gate = j >= 2 * N - 1 - g_gate_bits - start;
    if (j_gate) {
        if (N2 - start < 0) return -1;
        rd |= (1ULL << (N1 - j)) << (N2 - start);
    }
    uint64_t fr = board_mask & ~(ld | rd | col);
    int j_lt_N3 = j < N - 3, j_eq_N3 = j == N - 3, j_eq_N2 = j == N - 2;
    int k_lt_l = k < l, start_lt_k = start < k, start_lt_l = start < l;
    int l_eq_kp1 = l == k + 1, k_eq_lp1 = k == l + 1;
    if (j_lt_N3) {                                         /* 18Py:839-906 */
        jmark = j + 1; endmark = N2;
        if (j_gate) {
            if (k_lt_l) {
                mark1 = k - 1; mark2 = l - 1;
                if (start_lt_l) target = start_lt_k ? (!l_eq_kp1 ? 0 : 4) : 1;
                else target = 2;
            } else {
                mark1 = l - 1; mark2 = k - 1;
                if (start_lt_k) target = start_lt_l ? (!k_eq_lp1 ? 5 : 7) : 6;
                else target = 2;
            }
        } else {
            if (k_lt_l) { mark1 = k - 1; mark2 = l - 1; target = !l_eq_kp1 ? 8 : 9; }
            else { mark1 = l - 1; mark2 = k - 1; target = !k_eq_lp1 ? 10 : 11; }
        }
    } else if (j_eq_N3) {                                  /* 18Py:914-950 */
        endmark = N2;
        if (k_lt_l) {
            mark1 = k - 1; mark2 = l - 1;
            if (start_lt_l) {
                if (start_lt_k) target = !l_eq_kp1 ? 12 : 15;
                else { mark2 = l - 1; target = 13; }
            } else target = 14;
        } else {
            mark1 = l - 1; mark2 = k - 1;
            if (start_lt_k) {
                if (start_lt_l) target = !k_eq_lp1 ? 16 : 18;
                else { mark2 = k - 1; target = 17; }
            } else target = 14;
        }
    } else if (j_eq_N2) {                                  /* 18Py:957-1017 */
        if (k_lt_l) {
            endmark = N2;
            if (start_lt_l) {
                if (start_lt_k) {
                    mark1 = k - 1;
                    if (!l_eq_kp1) { mark2 = l - 1; target = 19; }
                    else target = 22;
                } else { mark2 = l - 1; target = 20; }
            } else target = 21;
        } else {
            if (start_lt_k) {
                if (start_lt_l) {
                    if (k < N2) {
                        mark1 = l - 1; endmark = N2;
                        if (!k_eq_lp1) { mark2 = k - 1; target = 23; }
                        else target = 24;
                    } else {
                        if (l != N - 3) { mark2 = l - 1; endmark = N - 3; target = 20; }
                        else { endmark = N - 4; target = 21; }
                    }
                } else {
                    if (k != N2) { mark2 = k - 1; endmark = N2; target = 25; }
                    else { endmark = N - 3; target = 21; }
                }
            } else { endmark = N2; target = 21; }
        }
    } else {                                               /* 18Py:1023-1031 */
        endmark = N2;
        if (start > k) target = 26;
        else { mark1 = k - 1; target = 27; }
    }
    t->ld = ld; t->rd = rd; t->col = col; t->free = fr;
    t->start = start; t->fid = target; t->jmark = jmark; t->endmark = endmark;
    t->mark1 = mark1; t->mark2 = mark2; t->w = symmetry_w(ijkl, N); t->ijkl = ijkl;
    return 0;
}

/* Exported for host-logic tests: per record {fid, start, jmark, endmark,
 * mark1, mark2, w, free}.  Returns 0 or -1 (shift out of range). */
NQO_EXPORT int nqo_build_soa(int N, const uint32_t *rec, long m, int32_t *out8)
{
    for (long a = 0; a < m; a++) {
        task_t t;
        if (build_task(N, rec + 4 * a, &t)) return -1;
        int32_t *o = out8 + 8 * a;
        o[0] = t.fid; o[1] = t.start; o[2] = t.jmark; o[3] = t.endmark;
        o[4] = t.mark1; o[5] = t.mark2; o[6] = t.w; o[7] = (int32_t)t.free;
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* A6: dfs_iter (18Py:368-562), with node counting                    */
/* ------------------------------------------------------------------ */
/* A node is one extraction bit = avail & -avail in any `while avail` loop
 * (after the jmark mask); base frames at endmark contribute none. */
typedef struct { uint64_t ld, rd, col, fr; int fid, row; } frame_t;
#define NQO_STACK 2048

/* prune = 0: the func_meta availptn table of 18Py (:1103-1132), the node
 * definition of the product.  prune = 1: the pruning of the reference's
 * recursive SQ* functions (13Bit_codon/13Py_constellations_codon.py:183-678),
 * which differs in three places only: SQBjrB (fid 2) and the SQBjl* (fids
 * 8-11) place without the one-row lookahead, and SQd1BlB (fid 20) applies it
 * on its mark2 step (:306-314).  Counts are the same (every lookahead is a
 * sound prune); node counts differ. */
#define NQO_PRUNE_13PY 1
#define NQO_PLANT_BASE 2     /* planted mistake (tests only): a base frame counts as a node */
#define NQO_PLANT_JMARK 4    /* planted mistake (tests only): column 0 counted before the jmark mask */
static int dfs_iter_mode(int N, const task_t *t, uint64_t *count, uint64_t *nodes, int mode)
{
    const int prune = mode & NQO_PRUNE_13PY;
    const uint64_t board_mask = (1ULL << N) - 1;
    const uint64_t n3 = 1ULL << (N > 3 ? N - 3 : 0), n4 = 1ULL << (N > 4 ? N - 4 : 0);
    frame_t st[NQO_STACK];
    int sp = 0;
    uint64_t total = 0, nn = 0;
    st[sp++] = (frame_t){t->ld, t->rd, t->col, t->free, t->fid, t->start};
    while (sp > 0) {
        frame_t f = st[--sp];
        if (!f.fr) continue;
        const int fid = f.fid, ptn = META_PTN[fid], nxt = META_NEXT[fid];
        uint64_t avail = f.fr, ld = f.ld;
        if (ptn == 5 && f.row == t->endmark) {             /* base */
            if (fid == 14) total += (avail >> 1) ? 1 : 0;
            else total += 1;
            if (mode & NQO_PLANT_BASE) nn++;
            continue;
        }
        int step = 1, add1 = 0, row_step = f.row + 1, use_blocks = 0;
        int use_future = META_AVAIL[fid] == 1 && !(prune == 1 && fid == 2), lnext = fid;
        uint64_t bK = 0, bL = 0;
        if (ptn == 0 || ptn == 1 || ptn == 2) {             /* mark rows */
            int at_mark = (ptn == 0 || ptn == 2) ? (f.row == t->mark1) : (f.row == t->mark2);
            if (at_mark && avail) {
                step = (ptn == 0 || ptn == 1) ? 2 : 3;
                add1 = (ptn == 1 && fid == 20) ? 1 : 0;
                row_step = f.row + step;
                bK = BLOCK_K_CAT[fid] == 3 ? n3 : (BLOCK_K_CAT[fid] == 4 ? n4 : 0);
                bL = (uint64_t)BLOCK_L[fid];
                use_blocks = 1; use_future = 0; lnext = nxt;
            }
        } else if (ptn == 3 && f.row == t->jmark) {         /* jmark row */
            if ((mode & NQO_PLANT_JMARK) && (avail & 1)) nn++;
            avail &= ~1ULL;
            ld |= 1;
            lnext = nxt;
            if (!avail) continue;
        } else if (ptn == 4) {
            /* P5 (fids 8-11, "SQBjl*"), 13Bit_codon/13Py_constellations_codon.py:620-678
             * (= 11Bit_Python/17Py_constellations.py:406-465): these records start above
             * the row where the j queen's down-right diagonal enters the board, the bit
             * build_soa leaves out (its 32-bit gate, 18Py:814-815).  At row N-1-jmark
             * the frame takes that bit (rd |= 1 << (N-1), column N-1 no longer free) and
             * continues as next_funcid on the same row; rows above it are placed without
             * the one-row lookahead, as the SQBjl loops do. */
            if (f.row == N - 1 - t->jmark) {
                const uint64_t bj = 1ULL << (N - 1);
                if (sp >= NQO_STACK) return -1;
                st[sp++] = (frame_t){ld, f.rd | bj, f.col, avail & ~bj, nxt, f.row};
                continue;
            }
            use_future = 0;
        }
        if (use_blocks) {
            const int mlook = prune == 1 && fid == 20 && row_step < t->endmark;
            while (avail) {
                uint64_t bit = avail & (0 - avail); avail &= avail - 1; nn++;
                uint64_t nld = ((ld | bit) << step) | (uint64_t)add1 | bL;
                uint64_t nrd = ((f.rd | bit) >> step) | bK;
                uint64_t ncol = f.col | bit;
                uint64_t nf = board_mask & ~(nld | nrd | ncol);
                if (nf) {
                    if (mlook && !(board_mask & ~((nld << 1) | (nrd >> 1) | ncol))) continue;
                    if (sp >= NQO_STACK) return -1;
                    st[sp++] = (frame_t){nld, nrd, ncol, nf, lnext, row_step};
                }
            }
            continue;
        }
        const int look = use_future && row_step < t->endmark;
        while (avail) {
            uint64_t bit = avail & (0 - avail); avail &= avail - 1; nn++;
            uint64_t nld = (ld | bit) << 1, nrd = (f.rd | bit) >> 1, ncol = f.col | bit;
            uint64_t nf = board_mask & ~(nld | nrd | ncol);
            if (!nf) continue;
            if (look && !(board_mask & ~((nld << 1) | (nrd >> 1) | ncol))) continue;
            if (sp >= NQO_STACK) return -1;
            st[sp++] = (frame_t){nld, nrd, ncol, nf, lnext, row_step};
        }
    }
    *count = total;
    if (nodes) *nodes = nn;
    return 0;
}

static int dfs_iter(int N, const task_t *t, uint64_t *count, uint64_t *nodes)
{
    return dfs_iter_mode(N, t, count, nodes, 0);
}

/* Recursive form of the same count, 18Py:564-696 (`dfs`).  Kept as an
 * independent restatement: the tests require it to agree with dfs_iter on
 * counts and node counts for every record. */
static uint64_t dfs_rec(int N, const task_t *t, int fid, uint64_t ld, uint64_t rd, uint64_t col,
                        int row, uint64_t fr, uint64_t *nodes)
{
    const uint64_t board_mask = (1ULL << N) - 1;
    const int nxt = META_NEXT[fid], ptn = META_PTN[fid];
    uint64_t avail = fr, total = 0;
    if (!avail) return 0;
    if (ptn == 5 && row == t->endmark) return (fid == 14) ? ((avail >> 1) ? 1 : 0) : 1;
    int step = 1, add1 = 0, row_step = row + 1, use_blocks = 0;
    int use_future = META_AVAIL[fid] == 1, lnext = fid;
    uint64_t bK = 0, bL = 0;
    if (ptn == 0 || ptn == 1 || ptn == 2) {
        int at_mark = (ptn == 0 || ptn == 2) ? (row == t->mark1) : (row == t->mark2);
        if (at_mark && avail) {
            step = (ptn == 0 || ptn == 1) ? 2 : 3;
            add1 = (ptn == 1 && fid == 20) ? 1 : 0;
            row_step = row + step;
            int cat = BLOCK_K_CAT[fid];
            bK = cat == 3 ? (1ULL << (N - 3)) : (cat == 4 ? (1ULL << (N - 4)) : 0);
            bL = (uint64_t)BLOCK_L[fid];
            use_blocks = 1; use_future = 0; lnext = nxt;
        }
    } else if (ptn == 3 && row == t->jmark) {
        avail &= ~1ULL; ld |= 1; lnext = nxt;
        if (!avail) return 0;
    } else if (ptn == 4) {
        /* P5, 13Py:620-678: the j diagonal enters at row N-1-jmark */
        if (row == N - 1 - t->jmark) {
            const uint64_t bj = 1ULL << (N - 1);
            return dfs_rec(N, t, nxt, ld, rd | bj, col, row, avail & ~bj, nodes);
        }
        use_future = 0;
    }
    while (avail) {
        uint64_t bit = avail & (0 - avail);
        avail &= avail - 1;
        (*nodes)++;
        uint64_t nld, nrd, ncol = col | bit;
        if (use_blocks) {
            nld = ((ld | bit) << step) | (uint64_t)add1 | bL;
            nrd = ((rd | bit) >> step) | bK;
        } else {
            nld = (ld | bit) << 1;
            nrd = (rd | bit) >> 1;
        }
        uint64_t nf = board_mask & ~(nld | nrd | ncol);
        if (!nf) continue;
        if (!use_blocks && use_future && row_step < t->endmark &&
            !(board_mask & ~((nld << 1) | (nrd >> 1) | ncol)))
            continue;
        total += dfs_rec(N, t, lnext, nld, nrd, ncol, row_step, nf, nodes);
    }
    return total;
}

/*
 * nqo_count_records: exec_solutions(use_gpu=False), 18Py:1057-1275.
 * For each u32 record {ld, rd, col, startijkl}: build_soa (A4), dfs_iter
 * (A6), weighted count = count * w (A3, 18Py:1266).
 * weighted/raw/nodes may be NULL.  threads <= 0 uses the OpenMP default.
 * schedule: 0 = static (Codon @par), 1 = dynamic,64.
 * Returns 0, -1 (stack capacity), -3 (shift out of range).
 */
NQO_EXPORT int nqo_count_records(int N, const uint32_t *rec, long m, uint64_t *weighted,
                                 uint64_t *raw, uint64_t *nodes, int threads, int schedule)
{
    int err = 0;
#ifdef _OPENMP
    if (threads <= 0) threads = omp_get_max_threads();
#else
    (void)threads;
#endif
    (void)schedule;
#define NQO_BODY                                                              \
    {                                                                         \
        task_t t;                                                             \
        uint64_t c = 0, nn = 0;                                               \
        int rc = build_task(N, rec + 4 * a, &t) ? -3 : dfs_iter(N, &t, &c, &nn); \
        if (rc) { _Pragma("omp atomic write") err = rc; }                     \
        if (weighted) weighted[a] = c * (uint64_t)t.w;                        \
        if (raw) raw[a] = c;                                                  \
        if (nodes) nodes[a] = nn;                                             \
    }
    if (schedule == 1) {
#pragma omp parallel for schedule(dynamic, 64) num_threads(threads)
        for (long a = 0; a < m; a++) NQO_BODY
    } else {
#pragma omp parallel for schedule(static) num_threads(threads)
        for (long a = 0; a < m; a++) NQO_BODY
    }
#undef NQO_BODY
    return err;
}

/* Weighted counts and node counts under the pruning of the reference's
 * recursive SQ* functions (dfs_iter_mode with NQO_PRUNE_13PY).  plant: 0, or
 * NQO_PLANT_BASE / NQO_PLANT_JMARK to plant a node-accounting mistake (the
 * tests show that the reference's node census catches it). */
NQO_EXPORT int nqo_count_records_13py(int N, const uint32_t *rec, long m, uint64_t *weighted, uint64_t *nodes,
                                      int plant)
{
    int err = 0;
#pragma omp parallel for schedule(dynamic, 16)
    for (long a = 0; a < m; a++) {
        task_t t;
        uint64_t c = 0, nn = 0;
        int rc = build_task(N, rec + 4 * a, &t) ? -3 : dfs_iter_mode(N, &t, &c, &nn, NQO_PRUNE_13PY | plant);
        if (rc) {
#pragma omp atomic write
            err = rc;
        }
        weighted[a] = c * (uint64_t)t.w;
        nodes[a] = nn;
    }
    return err;
}

/* Same result through the recursive restatement (single thread per record). */
NQO_EXPORT int nqo_count_records_recursive(int N, const uint32_t *rec, long m, uint64_t *weighted,
                                           uint64_t *nodes)
{
    int err = 0;
#pragma omp parallel for schedule(dynamic, 16)
    for (long a = 0; a < m; a++) {
        task_t t;
        uint64_t nn = 0, c = 0;
        if (build_task(N, rec + 4 * a, &t)) {
#pragma omp atomic write
            err = -3;
        } else {
            c = dfs_rec(N, &t, t.fid, t.ld, t.rd, t.col, t.start, t.free, &nn);
        }
        weighted[a] = c * (uint64_t)t.w;
        if (nodes) nodes[a] = nn;
    }
    return err;
}

/* ------------------------------------------------------------------ */
/* A9: solutions fixed by the 180- and 90-degree rotations            */
/* ------------------------------------------------------------------ */
/* The reference obtains Unique from the bit-symmetry backtracker
 * (13Bit_codon/07Py_bit_symmetry_codon.py:118-411: unique = c2+c4+c8,
 * total = 2c2+4c4+8c8, :410-411).  With F90 = 2*c2 solutions fixed by the
 * quarter turn and F180 = 2*c2 + 4*c4 fixed by the half turn, Burnside's
 * lemma over the 8 board symmetries (no solution is fixed by a reflection
 * when N > 1) gives unique = (T + F180 + 2*F90) / 8.  These enumerators
 * count F180 and F90 directly. */

/* rows r (top) and N-1-r (bottom) are filled together; a queen at (r,c)
 * implies one at (N-1-r, N-1-c).  diag index c-r+N-1, anti index r+c. */
static uint64_t f180_rec(int N, int r, uint64_t cols, uint64_t dg, uint64_t an)
{
    const int h = N / 2;
    if (r == h) return 1;
    const int rb = N - 1 - r;
    uint64_t cnt = 0;
    for (int c = 0; c < N; c++) {
        int cb = N - 1 - c;
        if (c == cb) continue;                     /* same column */
        if (r == c) continue;                      /* c-r == cb-rb: same diagonal */
        if (r + c == N - 1) continue;              /* r+c == rb+cb: same anti-diagonal */
        uint64_t bc = 1ULL << c, bcb = 1ULL << cb;
        uint64_t d1 = 1ULL << (c - r + N - 1), d2 = 1ULL << (cb - rb + N - 1);
        uint64_t a1 = 1ULL << (r + c), a2 = 1ULL << (rb + cb);
        if ((cols & (bc | bcb)) || (dg & (d1 | d2)) || (an & (a1 | a2))) continue;
        cnt += f180_rec(N, r + 1, cols | bc | bcb, dg | d1 | d2, an | a1 | a2);
    }
    return cnt;
}

/* quarter-turn orbit of (r,c): (r,c) (c,N-1-r) (N-1-r,N-1-c) (N-1-c,r) */
static uint64_t f90_rec(int N, int *q, int placed)
{
    if (placed == N) return 1;
    int r = 0;
    while (q[r] >= 0) r++;
    uint64_t cnt = 0;
    for (int c = 0; c < N; c++) {
        int pr[4] = {r, c, N - 1 - r, N - 1 - c};
        int pc[4] = {c, N - 1 - r, N - 1 - c, r};
        int ok = 1, n = 0;
        int ur[4], uc[4];
        for (int a = 0; a < 4 && ok; a++) {       /* distinct squares of the orbit */
            int dup = 0;
            for (int b = 0; b < n; b++) if (ur[b] == pr[a] && uc[b] == pc[a]) dup = 1;
            if (!dup) { ur[n] = pr[a]; uc[n] = pc[a]; n++; }
        }
        for (int a = 0; a < n && ok; a++) {
            if (q[ur[a]] >= 0) { ok = 0; break; }
            for (int b = 0; b < N && ok; b++) {
                if (q[b] < 0) continue;
                int dr = b - ur[a], dc = q[b] - uc[a];
                if (dc == 0 || dr == dc || dr == -dc) ok = 0;
            }
            for (int b = 0; b < a && ok; b++) {
                int dr = ur[b] - ur[a], dc = uc[b] - uc[a];
                if (dr == 0 || dc == 0 || dr == dc || dr == -dc) ok = 0;
            }
        }
        if (!ok) continue;
        for (int a = 0; a < n; a++) q[ur[a]] = uc[a];
        cnt += f90_rec(N, q, placed + n);
        for (int a = 0; a < n; a++) q[ur[a]] = -1;
    }
    return cnt;
}

NQO_EXPORT int nqo_symmetric_counts(int N, uint64_t *f90, uint64_t *f180)
{
    if (N < 1 || N > 31) return -1;
    uint64_t c180;
    if (N % 2 == 1) {
        /* the middle row maps onto itself: its queen must be the centre */
        int m = N / 2;
        c180 = 0;
        if (N == 1) c180 = 1;
        else {
            /* start with the centre queen placed, then fill row pairs */
            uint64_t cols = 1ULL << m, dg = 1ULL << (N - 1), an = 1ULL << (2 * m);
            c180 = f180_rec(N, 0, cols, dg, an);
        }
    } else {
        c180 = f180_rec(N, 0, 0, 0, 0);
    }
    *f180 = c180;
    int q[32];
    for (int a = 0; a < 32; a++) q[a] = -1;
    *f90 = f90_rec(N, q, 0);
    return 0;
}

/* Brute force for small N: every solution as a permutation, classified by
 * the size of its orbit under the 8 board symmetries (the c2/c4/c8 of
 * 07Py_bit_symmetry_codon.py:176-257).  out = {total, unique, F90, F180}. */
static void bf_rec(int N, int r, int *q, uint64_t cols, uint64_t dg, uint64_t an, uint64_t *out)
{
    if (r == N) {
        int rot[32] = {0}, img[32];
        memcpy(img, q, sizeof(int) * N);
        int fix90 = 0, fix180 = 0;
        /* rot90: (r,c) -> (c, N-1-r) */
        for (int a = 0; a < N; a++) rot[q[a]] = N - 1 - a;
        fix90 = memcmp(rot, q, sizeof(int) * N) == 0;
        for (int a = 0; a < N; a++) img[a] = N - 1 - q[N - 1 - a];
        fix180 = memcmp(img, q, sizeof(int) * N) == 0;
        /* canonical iff lexicographically minimal among its 8 images */
        int best = 1;
        int cur[32], t[32];
        memcpy(cur, q, sizeof(int) * N);
        for (int s = 0; s < 8 && best; s++) {
            if (s == 4) { for (int a = 0; a < N; a++) t[a] = q[N - 1 - a]; memcpy(cur, t, sizeof(int) * N); }
            else if (s > 0) { for (int a = 0; a < N; a++) t[cur[a]] = N - 1 - a; memcpy(cur, t, sizeof(int) * N); }
            if (s > 0) {
                for (int a = 0; a < N; a++) {
                    if (cur[a] < q[a]) { best = 0; break; }
                    if (cur[a] > q[a]) break;
                }
            }
        }
        out[0]++;
        if (best) out[1]++;
        if (fix90) out[2]++;
        if (fix180) out[3]++;
        return;
    }
    for (int c = 0; c < N; c++) {
        uint64_t bc = 1ULL << c, bd = 1ULL << (c - r + N - 1), ba = 1ULL << (r + c);
        if ((cols & bc) || (dg & bd) || (an & ba)) continue;
        q[r] = c;
        bf_rec(N, r + 1, q, cols | bc, dg | bd, an | ba, out);
    }
}

NQO_EXPORT int nqo_bruteforce(int N, uint64_t *out4)
{
    if (N < 1 || N > 14) return -1;
    int q[32];
    out4[0] = out4[1] = out4[2] = out4[3] = 0;
    bf_rec(N, 0, q, 0, 0, 0, out4);
    return 0;
}

NQO_EXPORT int nqo_abi_version(void) { return 1; }
