/*
 * nq.h -- C ABI of libnq, the MI355X (gfx950) constellation counter.
 *
 * Drop-in for the constellation solver's counting step:
 *   exec_solutions(N, constellations, use_gpu)   13Bit_codon/18Py_constellations_GPU_cuda_codon.py:1057-1275
 *   kernel_dfs_iter_gpu launch (Codon @gpu.kernel, gpu.raw pointers)      18Py:1195-1215, kernel 18Py:149-365
 * The reference binds its kernel through Codon's gpu.raw(list) + grid=/block=;
 * this library is bound through ctypes with plain pointers and sizes.
 *
 * Records are the reference's 16-byte constellation record, u32 little
 * endian {ld, rd, col, startijkl} (writer 18Py:2009-2016).  Per-record
 * outputs are the weighted count (count * w, w in {2,4,8}, 18Py:1266) and,
 * optionally, the DFS node count (one node = one extraction
 * bit = avail & -avail of the reference dfs_iter, 18Py:368-562).
 *
 * All pointers passed in are caller-owned HOST memory unless the name says
 * "dev"; nothing is retained after a call returns.  A context is bound to
 * one device and one HIP stream and is not thread-safe.  Every entry point
 * returns 0 or a negative nq_status; it never returns a partial count
 * (the reference's stack "safety valve" at 18Py:354-356 returns a partial
 * total silently -- here that case is NQ_ECAPACITY).
 */
#ifndef NQ_H
#define NQ_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NQ_ABI_VERSION 1

enum nq_status {
    NQ_OK = 0,
    NQ_EINVAL = -1,        /* bad argument (N out of 6..28, NULL pointer, size) */
    NQ_EUNSUPPORTED = -2,  /* record outside the u32 path: shift out of range,
                              no terminal row, terminal depth beyond 16 */
    NQ_ECAPACITY = -3,     /* DFS stack / work-item buffer capacity exceeded */
    NQ_EHIP = -4,          /* HIP runtime error (message in nq_last_error) */
    NQ_ENOMEM = -5,
    NQ_ESTATE = -6         /* call order: e.g. run before load */
};

typedef struct nq_ctx nq_ctx;

/* Batched-kernel schedule (no reference counterpart: the reference launches
 * one thread per constellation, 18Py:1195-1215, or a grid-stride batch,
 * 14Bit_codon_GPU/359Py_pushbranch_partial.py:1781-1867). */
typedef struct {
    int32_t split_depth;   /* frames at this DFS depth become work items (0 = no split,
                              -1 = the per-record depths of nq_ctx_set_splits) */
    int32_t block;         /* workgroup size: 64, 128 or 256 */
    int32_t lds_stack;     /* 1: DFS stack in LDS, 0: in private (scratch) memory */
    int32_t use_graph;     /* 1: replay the launch sequence as a captured hipGraph */
    int32_t chunk;         /* work items a wave claims per queue fetch (>= 1) */
    int32_t waves_per_eu;  /* persistent occupancy target; 0 = from the occupancy API */
} nq_sched;

typedef struct {
    double kernel_ms;      /* device time of the DFS kernel, HIP events on the context stream
                              (-1 for a graph replay, which is timed as a whole) */
    double total_ms;       /* device time of the whole run (expand + DFS + finalize) */
    uint64_t items;        /* work items after splitting */
    uint64_t launches;     /* DFS kernel launches in the run */
} nq_timing;

/* ---- library / context -------------------------------------------------- */
int nq_abi_version(void);
const char *nq_last_error(void);
int nq_ctx_create(int device, nq_ctx **out);
int nq_ctx_destroy(nq_ctx *ctx);
/* device properties the scheduler uses: {compute units, max LDS bytes per
 * workgroup, wavefront size, clock kHz} */
int nq_device_info(nq_ctx *ctx, int32_t *out4);
/* No counterpart in the reference (a measurement for the roofline): the VALU
 * issue rate of independent v_or_b32 / v_and_b32 chains with waves_per_simd
 * (1..8) waves resident on every SIMD, `iters` (1..2^24) iterations of 32
 * ops per lane.  *ms_out: device time; *clk_out: clocks per wave instruction
 * per SIMD at the reported clock; *xor_out: XOR of every lane's result (a
 * check that the chains ran).  Any out pointer may be NULL. */
int nq_issue_probe(nq_ctx *ctx, int waves_per_simd, uint32_t iters, double *ms_out, double *clk_out,
                   uint32_t *xor_out);

/* ---- resident task set ---------------------------------------------------
 * exec_solutions' input list (18Py:1057): copies m records to HBM and decodes
 * them there (build_soa_for_range, 18Py:698-1054, and the func_meta schedule,
 * 18Py:1103-1156) into per-record DFS descriptors.  `order` (may be NULL)
 * is the processing order for the batched kernel (a permutation of 0..m-1);
 * results are always reported in record order. */
int nq_ctx_load(nq_ctx *ctx, int N, const uint32_t *rec4, uint64_t m, const uint32_t *order);
/* The same for a set mixing board sizes (N[i] for record i): several count
 * requests share one resident set and one persistent queue.  No reference
 * counterpart (the reference counts one N per exec_solutions call,
 * 18Py:2116-2160). */
int nq_ctx_load_mixed(nq_ctx *ctx, const int32_t *N, const uint32_t *rec4, uint64_t m, const uint32_t *order);

/* Preset expansion on the device: each of the m records of preset p (the
 * reference's 16-byte constellation records) becomes one record per free
 * column of its start row, and this repeats `levels` times (1..8): the
 * records of preset p + levels, in input order, columns ascending.  Replaces
 * the recursion of set_pre_queens beyond preset p (18Py:1535-1613) for the
 * records gen_constellations produced at preset p (18Py:1752-1781).
 * With out4 == NULL only *m_out (the record count) is set; otherwise cap must
 * be at least that count (NQ_ECAPACITY, with *m_out set, when it is not).
 * A record whose start row is pre-placed or below the board, or a child field
 * beyond 32 bits, is NQ_EUNSUPPORTED. */
int nq_expand_records(nq_ctx *ctx, int N, const uint32_t *rec4, uint64_t m, int levels, uint32_t *out4,
                      uint64_t cap, uint64_t *m_out);
/* nq_ctx_load of the records nq_expand_records would produce, expanded and
 * decoded on the device without a round trip through the host; *m_out gets
 * the number of records loaded (may be NULL). */
int nq_ctx_load_expanded(nq_ctx *ctx, int N, const uint32_t *rec4, uint64_t m, int levels, uint64_t *m_out);

/* Per-record split depths for the batched kernel (used when
 * sched->split_depth == -1): the cost model bins constellations by predicted
 * work and splits the heavy ones deeper.  No reference counterpart (the
 * reference splits every task at one preset depth, 359Py:5776-5785).
 * Reset by the next load. */
int nq_ctx_set_splits(nq_ctx *ctx, const uint8_t *depth, uint64_t m);

/* Plain kernel: one lane per constellation, iterative bitboard DFS with the
 * stack in LDS, u64 wave reduction of the weighted counts.  Replaces
 * kernel_dfs_iter_gpu (18Py:149-365) and its launch loop (18Py:1195-1215). */
int nq_ctx_run_plain(nq_ctx *ctx, int count_nodes, uint64_t *total_out, nq_timing *t);

/* Batched kernel: heavy subtrees split at sched->split_depth into work items,
 * one persistent launch pulling items from a device queue. */
int nq_ctx_run_batched(nq_ctx *ctx, const nq_sched *sched, int count_nodes,
                       uint64_t *total_out, nq_timing *t);

/* Per-record results of the last run (weighted count; node count when the
 * run counted nodes, else zeros).  Either pointer may be NULL. */
int nq_ctx_results(nq_ctx *ctx, uint64_t *weighted_out, uint64_t *nodes_out);

/* One-shot form of exec_solutions(N, list, use_gpu=True) (18Py:1057-1275):
 * load + run + results.  sched == NULL selects the plain kernel. */
int nq_count(nq_ctx *ctx, int N, const uint32_t *rec4, uint64_t m, const nq_sched *sched,
             uint64_t *weighted_out, uint64_t *nodes_out, uint64_t *total_out);

/* The same count as nq_count, allocating nothing: no hipMalloc, hipFree or
 * other allocating call.  No counterpart in the reference; for callers that
 * cannot allocate (code under graph capture, tight serving loops, harnesses
 * that forbid allocation).  Every device buffer of the call (records,
 * descriptors, work items, counts) is carved out of one static 128 MiB
 * device array of the library, which each device holds from its first
 * context until the process ends; the context only supplies its stream.
 * Limits: at most 2^18 records (NQ_ECAPACITY beyond), and with a schedule at
 * most the work items that fit the rest of the array (about 2.2 M; NQ_ECAPACITY
 * with the count in nq_last_error: use a shallower split depth).  The
 * schedule's use_graph is ignored (the launches run eagerly) and per-record
 * split depths (split_depth -1) are refused.  Calls on one device share the
 * array: calls that overlap on one device race.  The context's resident set
 * (nq_ctx_load) is left as it was. */
int nq_count_noalloc(nq_ctx *ctx, int N, const uint32_t *rec4, uint64_t m, const nq_sched *sched,
                     uint64_t *weighted_out, uint64_t *nodes_out, uint64_t *total_out);

/* ---- Unique ---------------------------------------------------------------
 * The constellation solvers print Unique = 0 (18Py:2160).  Unique is the
 * orbit count of 13Bit_codon/07Py_bit_symmetry_codon.py:410-411
 * (unique = c2+c4+c8).  This returns the number of solutions fixed by the
 * quarter turn (F90) and by the half turn (F180); the host forms
 * unique = (total + F180 + 2*F90) / 8.  2 <= N <= 27. */
int nq_symmetric_counts(nq_ctx *ctx, int N, uint64_t *f90_out, uint64_t *f180_out);

/* ---- small boards ----------------------------------------------------------
 * Replaces _bit_total (18Py:2071-2084), which main (18Py:2118-2126) uses for
 * N <= 5 instead of constellations: every solution of the N x N board by a
 * plain DFS over all rows, on the device.  1 <= N <= 12. */
int nq_small_total(nq_ctx *ctx, int N, uint64_t *total_out);

/* ---- cost model (no reference counterpart; the reference orders work by a
 * hand-made key, 14Bit_codon_GPU/359Py_pushbranch_partial.py:4342-4404) ----
 * 3-layer FP32 MLP 24 -> 64 -> 64 -> 1 (ReLU), parameters packed as
 * W1[64x24] b1[64] W2[64x64] b2[64] W3[64] b3[1] = NQ_MLP_NPARAM floats. */
#define NQ_MLP_IN 24
#define NQ_MLP_H 64
#define NQ_MLP_NPARAM (64 * 24 + 64 + 64 * 64 + 64 + 64 + 1)

/* Features of the resident task set (row-major m x 24 floats, record order). */
int nq_ctx_features(nq_ctx *ctx, float *feat_out);
/* Processing order of the resident task set from the model, on the device:
 * features, forward pass and a stable radix sort, predicted heaviest first
 * (ties in record order), written to the order the batched kernel reads
 * (and to order_out[m] unless NULL).  The order only schedules work. */
int nq_ctx_order_by_cost(nq_ctx *ctx, const float *params, uint32_t *order_out);
/* Forward pass: x[m x 24] -> y[m]. */
int nq_mlp_forward(nq_ctx *ctx, const float *params, const float *x, uint64_t m, float *y_out);
/* Mean-squared-error loss and its gradient w.r.t. the packed parameters. */
int nq_mlp_loss_grad(nq_ctx *ctx, const float *params, const float *x, const float *y,
                     uint64_t m, float *loss_out, float *grad_out);
/* Adam training on device: `epochs` passes of minibatches of `batch` rows in
 * the order `perm` (m entries per epoch, epochs*m total), learning rate lr.
 * params is updated in place; loss_out[e] receives epoch e's mean loss. */
int nq_mlp_train(nq_ctx *ctx, float *params, const float *x, const float *y, uint64_t m,
                 const uint32_t *perm, int epochs, int batch, float lr, float *loss_out);

/* ---- host-side descriptor check (no GPU needed) ---------------------------
 * The same decode the device runs, executed on the host for tests: per record
 * {fid, start, jmark, endmark, mark1, mark2, w, free, T, status}. */
int nq_decode_host(int N, const uint32_t *rec4, uint64_t m, int32_t *out10);

#ifdef __cplusplus
}
#endif
#endif /* NQ_H */
