/*
 * bsx.h -- C ABI of libbsx.so, the MI355X (gfx950) seed-extend engine for
 * BSBolt's bisulfite aligner (forked BWA-MEM 0.7.17 in bsbolt/External/BWA).
 *
 * The library replaces two scalar routines of the reference with batched,
 * bit-exact HIP kernels:
 *   ksw_extend2  ksw.h:108, ksw.c:380-479   -> bsx_extend_batch / bsx_extend_batch_device
 *   ksw_global2  ksw.h:84,  ksw.c:504-606   -> bsx_global_batch / bsx_global_batch_device
 * and the NM/MD/XB strings bwa_gen_cigar2 appends to a CIGAR (bwa.c:250-343)
 *                                           -> bsx_md_batch_device
 * Their callers in the reference are mem_chain2aln (bwamem.c:735,763) and
 * bwa_gen_cigar2 (bwa.c:248).  Each task of a batch carries exactly the
 * arguments one reference call takes; each result carries exactly its outputs.
 *
 * Conventions
 *  - Sequences are 1 byte per base, codes 0..4 (A C G T N, bntseq.c:48-65),
 *    concatenated; task t uses query[q_off[t] .. q_off[t]+qlen[t]) and
 *    target[t_off[t] .. t_off[t]+tlen[t]).
 *  - The scoring matrix is 5x5, indexed mat[target*5 + query] (ksw.c:390-393),
 *    as built by bwa_fill_scmat (bwa.c:169-178).
 *  - Every entry point returns BSX_OK or a negative BSX_E* code and never
 *    aborts.  A task that violates a precondition is skipped (its result slots
 *    are left untouched), the others are computed, and the call returns
 *    BSX_EINVAL with the first offending task named by bsx_last_error().
 *  - A context is bound to one device and one stream; calls on distinct
 *    contexts are thread-safe, calls on one context are not.
 */
#ifndef BSX_H
#define BSX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BSX_OK        0
#define BSX_EINVAL   (-1) /* a task or argument violates a precondition        */
#define BSX_EDEVICE  (-2) /* HIP runtime error (text in bsx_last_error)         */
#define BSX_ENOMEM   (-3) /* device or host allocation failed                  */
#define BSX_ERANGE   (-4) /* a CIGAR did not fit its caller-sized slot          */

/* Largest query (flank) length the extension kernels take; longer -> BSX_EINVAL. */
#define BSX_MAX_QLEN 2048

typedef struct bsx_ctx bsx_ctx;

/* Scalar parameters shared by every task of a batch.  Mirrors the arguments
 * m=5, mat, o_del, e_del, o_ins, e_ins, zdrop of ksw_extend2 (ksw.h:108). */
typedef struct {
	int8_t  mat[25];
	int32_t o_del, e_del, o_ins, e_ins; /* e_del, e_ins >= 1; o_del, o_ins >= 0 */
	int32_t zdrop;                      /* <= 0 disables z-drop (ksw.c:458)    */
} bsx_scoring;

/* One ksw_extend2 call per task.  Preconditions per task (ksw.c:385,388,395):
 * 1 <= qlen <= BSX_MAX_QLEN, tlen >= 0, h0 > 0, w >= 0,
 * h0 + qlen*max(mat) < 2^20 (scores are tracked exactly in 32-bit lanes). */
typedef struct {
	int64_t n;
	const uint8_t  *query;
	const uint64_t *q_off;
	const int32_t  *qlen;
	const uint8_t  *target;
	const uint64_t *t_off;
	const int32_t  *tlen;
	const int32_t  *h0;
	const int32_t  *w;
	const int32_t  *end_bonus; /* pen_clip5 for left flanks, pen_clip3 for right */
} bsx_ext_tasks;

/* Outputs of ksw_extend2 (ksw.c:473-478); any pointer may be NULL. */
typedef struct {
	int32_t *score;   /* return value: best local score (>= h0)           */
	int32_t *qle;     /* *_qle                                            */
	int32_t *tle;     /* *_tle                                            */
	int32_t *gtle;    /* *_gtle                                           */
	int32_t *gscore;  /* *_gscore (-1 if the query end was never reached) */
	int32_t *max_off; /* *_max_off                                        */
	int64_t *cells;   /* DP cells evaluated: sum over rows of (end - beg) */
} bsx_ext_results;

/* One ksw_global2 call per task.  Preconditions: 1 <= qlen, tlen <= 65536,
 * min(qlen, 2w+1) <= BSX_MAX_QLEN (the band width, which sets the lanes per
 * task), |tlen - qlen| <= w (the reference reads uninitialised memory
 * otherwise); a task outside them fails the call with BSX_EINVAL. */
typedef struct {
	int64_t n;
	const uint8_t  *query;
	const uint64_t *q_off;
	const int32_t  *qlen;
	const uint8_t  *target;
	const uint64_t *t_off;
	const int32_t  *tlen;
	const int32_t  *w;
} bsx_glb_tasks;

/* Outputs of ksw_global2: score, *n_cigar and the BAM CIGAR (len<<4|op,
 * op 0=M 1=I 2=D), written to cigar[t*cigar_slot ...].  A CIGAR longer than
 * cigar_slot sets n_cigar[t] = -1 and the call returns BSX_ERANGE.
 * cigar == NULL asks for scores only (ksw.c:510-513). */
typedef struct {
	int32_t  *score;
	int32_t  *n_cigar;
	uint32_t *cigar;
	int64_t   cigar_slot;
} bsx_glb_results;

/* ---- context ---- */
bsx_ctx    *bsx_create(int device);               /* NULL on failure             */
void        bsx_destroy(bsx_ctx *ctx);
const char *bsx_last_error(const bsx_ctx *ctx);   /* never NULL                  */
int         bsx_set_stream(bsx_ctx *ctx, void *hip_stream); /* NULL = ctx's own  */
int         bsx_sync(bsx_ctx *ctx);               /* wait; report deferred errors */
const char *bsx_version(void);

/* Default scoring of BSBolt Align (Parser.py:47-81: -A1 -B4 -O6,6 -E1,1 -d100). */
void bsx_default_scoring(bsx_scoring *sc);

/* ---- extension (ksw_extend2) ----
 * Host-memory form: copies tasks in, runs, copies results out; synchronous. */
int bsx_extend_batch(bsx_ctx *ctx, const bsx_scoring *sc,
                     const bsx_ext_tasks *tasks, bsx_ext_results *res);
/* Device-memory form: every pointer in tasks/res is device memory; the work is
 * queued on the context's stream and the call returns without waiting.
 * Precondition violations are reported by the next bsx_sync(). */
int bsx_extend_batch_device(bsx_ctx *ctx, const bsx_scoring *sc,
                            const bsx_ext_tasks *tasks, bsx_ext_results *res);

/* Allocation-free form of bsx_extend_batch_device.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory during the
 * call (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Same arguments, results and policies (plain, learned,
 * fixed frame) as bsx_extend_batch_device, but the workspace is a static
 * device array of about 84 bytes per task for BSX_NOALLOC_MAX_TASKS tasks
 * (about 22 MB), which each GPU the library runs on holds until the process
 * ends; the call itself allocates nothing.  More than BSX_NOALLOC_MAX_TASKS
 * tasks -> BSX_EINVAL.  Every call on one GPU shares that GPU's array: calls
 * that overlap on one GPU (two contexts, or two streams) race and must be
 * ordered by the caller. */
#define BSX_NOALLOC_MAX_TASKS 262144
int bsx_extend_batch_device_noalloc(bsx_ctx *ctx, const bsx_scoring *sc,
                                    const bsx_ext_tasks *tasks, bsx_ext_results *res);

/* ---- global alignment + CIGAR (ksw_global2) ---- */
int bsx_global_batch(bsx_ctx *ctx, const bsx_scoring *sc,
                     const bsx_glb_tasks *tasks, bsx_glb_results *res);
int bsx_global_batch_device(bsx_ctx *ctx, const bsx_scoring *sc,
                            const bsx_glb_tasks *tasks, bsx_glb_results *res);
/* Bytes of direction scratch per in-flight task (default 65536): a CIGAR task
 * needs min(qlen, 2w+1) * tlen bytes; larger ones are rejected (BSX_EINVAL).
 * The context holds #CU * 192 such slots of device memory (#CU * 128 with
 * side streams off, bsx_set_side_streams).  The host-memory call
 * bsx_global_batch (and so bsx_ksw_global2) raises the slot for its own call
 * to the batch's largest need, up to 256 KiB; the device-memory call cannot see
 * the sizes and uses the slot as set. */
void bsx_set_global_zslot(bsx_ctx *ctx, int64_t bytes);

/* ---- NM, MD:Z and XB:Z of CIGARs (the tail of bwa_gen_cigar2, bwa.c:250-343) ----
 * For each task: the region [rb, re) of the doubled reference a CIGAR was made
 * for (rb < re, not bridging l_pac), the read as sequenced (before the
 * in-silico conversion; codes 0..4, read order) and that CIGAR as
 * bsx_global_batch wrote it (computed on the reversed sequences when
 * rb >= l_pac, bwa.c:216-221).  opac is the unconverted reference in the .pac
 * 2-bit layout (both copies, l_pac bases, bwa.c:430-431); crick_l is the first
 * position of the Crick copy (bntseq.c:139).  Outputs per task: the MD:Z value
 * (md[t*md_slot ..], md_len[t] bytes, no terminator), the XB:Z value (xb, likewise),
 * NM and meth[4*t ..] = CpG methylated, CpG unmethylated, CH methylated, CH
 * unmethylated (bwa_gen_cigar2's cg_meth, cg_unmeth, ch_meth, ch_unmeth).
 * status[t]: 0 ok; 1 the context window opac[rb-2, re+2) bridges the strands
 * (the reference appends nothing; nm = -1); -1 an output slot is too small;
 * -2 the walk would leave the window (within two bases of the reference ends,
 * where the reference reads past its buffer); -3 the region is not one the
 * reference makes a CIGAR for (rb < 0, rb >= re, re > 2*l_pac) or the CIGAR has
 * more than cigar_slot ops, an op other than M/I/D, or consumes more than
 * l_query read bases.  Tasks with status != 0 have
 * nm = -1 and empty strings; the call still returns BSX_OK.  Device memory
 * only; queued on the context's stream. */
typedef struct {
	int64_t n;
	int64_t l_pac, crick_l;
	const uint8_t  *opac;
	const int64_t  *rb, *re;
	const uint8_t  *oquery;
	const uint64_t *oq_off;
	const int32_t  *l_query;
	const int32_t  *n_cigar;
	const uint32_t *cigar;
	int64_t         cigar_slot;
} bsx_md_tasks;

typedef struct {
	char    *md;
	int64_t  md_slot;
	int32_t *md_len;
	char    *xb;
	int64_t  xb_slot;
	int32_t *xb_len;
	int32_t *nm;
	int32_t *meth;
	int32_t *status;
} bsx_md_results;

int bsx_md_batch_device(bsx_ctx *ctx, const bsx_md_tasks *tasks, bsx_md_results *res);

/* ---- SAM CIGAR of a region (the tail of mem_reg2aln, bwamem.c:1237-1263) ----
 * From the CIGAR bwa_gen_cigar2 made for region [rb, re) x [qb, qe) of a read
 * of l_seq bases: the position bns_depos gives for rb (re - 1 on the reverse
 * strand) and is_rev; a leading deletion is squeezed out and moves the
 * position, or else (only if the first op is not a deletion) a trailing one is
 * dropped; the unaligned ends become clip ops len << 4 | 3 (5' first, swapped
 * on the reverse strand), which the SAM writer prints as S or H.  pos is on
 * the doubled reference's forward copy (subtract the contig's offset for the
 * SAM POS).  n_cigar_out[t] = -1 for a task outside its preconditions
 * (n_cigar > cigar_slot, rb < 0, re <= rb, re > 2*l_pac, qb < 0, qe < qb,
 * qe > l_seq), -2 when the result does not fit cigar_slot_out.  Device memory;
 * queued on the context's stream. */
typedef struct {
	int64_t n, l_pac;
	const int64_t  *rb, *re;
	const int32_t  *qb, *qe, *l_seq, *n_cigar;
	const uint32_t *cigar;
	int64_t         cigar_slot;
} bsx_reg2aln_tasks;

typedef struct {
	uint32_t *cigar;
	int64_t   cigar_slot;
	int32_t  *n_cigar;
	int64_t  *pos;
	int32_t  *is_rev;
} bsx_reg2aln_results;

int bsx_reg2aln_cigar_device(bsx_ctx *ctx, const bsx_reg2aln_tasks *tasks, bsx_reg2aln_results *res);

/* ---- device memory and timing on the context's device/stream ----
 * Helpers for callers that hold their batches in device memory without a
 * HIP runtime of their own (e.g. Python through ctypes).  No counterpart in
 * the reference, which has no device. */
void *bsx_dev_alloc(bsx_ctx *ctx, size_t bytes);             /* NULL on failure   */
int   bsx_dev_free(bsx_ctx *ctx, void *ptr);
/* kind: 1 host->device, 2 device->host, 3 device->device; waits for completion */
int   bsx_memcpy(bsx_ctx *ctx, void *dst, const void *src, size_t bytes, int kind);
/* Records a start / stop event on the context's stream; bsx_timer_stop waits
 * for the stop event and returns the elapsed milliseconds in *ms. */
int   bsx_timer_start(bsx_ctx *ctx);
int   bsx_timer_stop(bsx_ctx *ctx, float *ms);

/* ---- learned-band form of the extension ----
 * No counterpart in the reference (which always runs ksw_extend2 over its full
 * adaptive band).  Under BSX_BAND_LEARNED the extension entry points first
 * predict a frame class per task with the band model (band-following frames of
 * 24, 40, 80 or 144 columns, or plain), run the frame kernel for the tasks whose
 * class is a frame, and keep a frame result only when the run's band never left
 * the frame with a non-zero cell (checked exactly, per row); every other task
 * runs the plain kernel.  Results are identical to BSX_BAND_PLAIN in every case.
 * BSX_BAND_FIXED runs every task whose query is <= 256 bases through frame class
 * `band_class` first (tests, labels). */
#define BSX_BAND_PLAIN   0
#define BSX_BAND_LEARNED 1
#define BSX_BAND_FIXED   2
#define BSX_BAND_CLASSES 5      /* 0..3 = frames of 24/40/80/144 columns, 4 = plain */
#define BSX_BAND_FEATURES 24    /* features per task (stored padded to 32)      */
#define BSX_BAND_FEATURE_STRIDE 32
#define BSX_BAND_PARAMS 6085    /* MLP 24-64-64-5: W1,b1,W2,b2,W3,b3 (row = out) */
int bsx_set_band_policy(bsx_ctx *ctx, int policy, int band_class);
/* Per class c of the last learned/fixed extension call: out[2c] tasks routed to
 * class c (c = 4: plain only), out[2c+1] of those whose frame run fell back. */
int bsx_band_stats(bsx_ctx *ctx, int64_t out[2 * BSX_BAND_CLASSES]);
/* FP32 master parameters of the band model (set resets the Adam state). */
int bsx_band_model_set(bsx_ctx *ctx, const float *params, int64_t n_params);
int bsx_band_model_get(bsx_ctx *ctx, float *params, int64_t n_params);
/* Device-memory forms (queued on the context's stream):
 *  features: feats[n][32] BF16 bit patterns (24 used, rest 0);
 *  labels:   labels[n] = smallest frame class whose run holds the task, 4 if
 *            none, 0xff for a task that violates a precondition;
 *  forward:  logits[n][5] (nullable) and argmax class cls[n] (nullable);
 *  grad:     grad[0..6084] = mean softmax cross-entropy gradient over the n
 *            samples, grad[6085] = mean loss;
 *  train:    `steps` Adam steps (lr, betas 0.9/0.999, eps 1e-8) on consecutive
 *            minibatches of `batch` samples (wrapping at n); losses[s] (device,
 *            nullable) receives the mean loss of step s. */
int bsx_band_features_device(bsx_ctx *ctx, const bsx_scoring *sc, const bsx_ext_tasks *tasks,
                             uint16_t *feats);
int bsx_band_labels_device(bsx_ctx *ctx, const bsx_scoring *sc, const bsx_ext_tasks *tasks,
                           uint8_t *labels);
int bsx_band_forward_device(bsx_ctx *ctx, const uint16_t *feats, int64_t n, float *logits,
                            uint8_t *cls);
/* The fused features + forward pass the learned policy runs before routing:
 * cls[n] = the predicted class of each task (device memory). */
int bsx_band_predict_device(bsx_ctx *ctx, const bsx_scoring *sc, const bsx_ext_tasks *tasks,
                            uint8_t *cls);
int bsx_band_grad_device(bsx_ctx *ctx, const uint16_t *feats, const uint8_t *labels, int64_t n,
                         float *grad);
int bsx_band_train_device(bsx_ctx *ctx, const uint16_t *feats, const uint8_t *labels, int64_t n,
                          int64_t batch, int steps, float lr, float *losses);

/* ---- graph-captured batch submission (no counterpart in the reference) ----
 * Captures one device-memory batch call (its binning, kernels and the error
 * read-back) into a HIP graph that keeps the task/result pointers and the
 * batch size, and owns its own device workspace.  To stream batches, refill
 * the same task buffers and replay with bsx_graph_launch; bsx_sync reports
 * rejected tasks as for the eager call.  The band policy, scoring and (for
 * CIGARs) slot size in force at capture are part of the graph; a band model
 * set later is used by later replays.  Replays of one graph on a context are
 * ordered on its stream. */
typedef struct bsx_graph bsx_graph;
int  bsx_graph_capture_extend(bsx_ctx *ctx, const bsx_scoring *sc, const bsx_ext_tasks *tasks,
                              bsx_ext_results *res, bsx_graph **graph);
int  bsx_graph_capture_global(bsx_ctx *ctx, const bsx_scoring *sc, const bsx_glb_tasks *tasks,
                              bsx_glb_results *res, bsx_graph **graph);
int  bsx_graph_launch(bsx_ctx *ctx, bsx_graph *graph);
void bsx_graph_destroy(bsx_graph *graph);

/* ---- per-kernel timing (measurement; no counterpart in the reference) ----
 * bsx_set_kernel_timing(ctx, k): the next k extension calls on the context
 * record HIP events around each kernel launch (0 switches it off).
 * bsx_kernel_times: waits, then writes the summed milliseconds per slot over
 * the recorded calls (slots 0..bins-1 = plain kernels by query-length bin,
 * then the 4 narrow classes, slot 15 = band predictor), the number of calls
 * in *calls, and starts a new recording.  bsx_ext_bin_limits writes the
 * largest query length of each plain bin and returns the number of bins. */
int bsx_set_kernel_timing(bsx_ctx *ctx, int max_calls);
int bsx_kernel_times(bsx_ctx *ctx, float *ms, int n_slots, int *calls);
int bsx_ext_bin_limits(int32_t *limits, int n);
/* bsx_set_side_streams(ctx, 0|1): the extension calls spread their independent
 * launches (query-length bins, frame classes) over the context's stream and
 * three side streams of its own, and the CIGAR calls run their band-width bins
 * 25-64 on a side stream beside the others (1, the default); or every launch
 * runs one after the other on the context's stream (0: each kernel then has
 * the GPU to itself, which is how per-kernel times are taken for the
 * roofline).  No counterpart in the reference. */
int bsx_set_side_streams(bsx_ctx *ctx, int enable);

/* ---- tuning table (no counterpart in the reference) ----
 * table[0] descriptors per block of the plain group kernels, table[1] of the
 * frame group kernels (both 16..4096), table[2] / table[3] the group shape of
 * the bins of queries 65-96 / 97-128 bases (0: 16 lanes x 6 / 8 columns,
 * 1: 32 lanes x 3 / 4 columns).  Every table gives the same results.
 * bsx_create loads the autotuner's cache entry for the device (key
 * "<gfx arch>/<CU count>" in $BSX_TUNE, else tune.json next to libbsx.so;
 * BSX_TUNE=off keeps the defaults).  bsx_set_tuning(ctx, NULL, 0) restores the
 * defaults.  bsx_get_tuning writes the table and the device's key and returns
 * where the table came from (BSX_TUNE_*), or BSX_EINVAL. */
#define BSX_TUNE_FIELDS  4
#define BSX_TUNE_DEFAULT 0
#define BSX_TUNE_CACHE   1
#define BSX_TUNE_CALLER  2
int bsx_set_tuning(bsx_ctx *ctx, const int32_t *table, int n);
int bsx_get_tuning(const bsx_ctx *ctx, int32_t *table, int n, char *key, int key_cap);

/* Scalar drop-in with the exact signature of ksw_extend2 (ksw.h:108), run as
 * a one-task batch on device 0.  m must be 5.  It returns INT32_MIN and leaves
 * *qle, *tle, *gtle, *gscore and *max_off untouched in these cases, which the
 * caller must check for (mem_chain2aln, bwamem.c:735,763, reads all five):
 *  - a precondition of bsx_ext_tasks fails: h0 <= 0 (the reference asserts,
 *    ksw.c:385), qlen < 1 (undefined in the reference, ksw.c:388,395),
 *    qlen > BSX_MAX_QLEN, h0 + qlen*max(mat) >= 2^20, w < 0, tlen < 0;
 *  - m != 5, query or mat NULL, or target NULL with tlen > 0;
 *  - no device, or a device error (bsx_create or the batch call fails).
 * Calls from many threads are safe: each call borrows a context from a
 * process-wide pool and returns it, so the pool holds as many contexts as the
 * peak number of concurrent callers (contexts live until the process ends). */
int bsx_ksw_extend2(int qlen, const uint8_t *query, int tlen, const uint8_t *target,
                    int m, const int8_t *mat, int o_del, int e_del, int o_ins, int e_ins,
                    int w, int end_bonus, int zdrop, int h0,
                    int *qle, int *tle, int *gtle, int *gscore, int *max_off);
/* Scalar drop-in with the exact signature of ksw_global2 (ksw.h:84), run as a
 * one-task batch on a pooled context like bsx_ksw_extend2.  With n_cigar and
 * cigar non-NULL, *cigar receives a fresh malloc'd BAM CIGAR of *n_cigar ops
 * that the caller frees (ksw.h:79; the reference builds it with realloc from
 * NULL, ksw.c:492-502,602); with either NULL only the score is computed.  It
 * returns INT32_MIN (and *n_cigar = 0, *cigar untouched) where a batch task
 * would be rejected: m != 5, qlen or tlen < 1 or > 65536, |tlen - qlen| > w
 * (the reference reads uninitialised memory there), min(qlen, 2w+1) > 2048 or
 * min(qlen, 2w+1)*tlen above the context's direction slot; and on a device
 * error.  bwa_gen_cigar2 (bwa.c:231-248) always passes w >= |tlen - qlen| + 3. */
int bsx_ksw_global2(int qlen, const uint8_t *query, int tlen, const uint8_t *target, int m,
                    const int8_t *mat, int o_del, int e_del, int o_ins, int e_ins, int w,
                    int *n_cigar, uint32_t **cigar);
/* Number of contexts the scalar drop-ins' pool has created so far. */
int64_t bsx_dropin_contexts(void);

#ifdef __cplusplus
}
#endif
#endif /* BSX_H */
