/*
 * chsga/abi.h -- C ABI of the MI355X (gfx950) ConSGA graph-DP engine.
 *
 * The reference exposes this hot path only as C++ templates called by its
 * drivers; every entry point below replaces one of them:
 *
 *   chsga_ctx_create      <- the graph set-up the drivers do before aligning:
 *                            build_in_csr (src/align.hpp:220-242), rank
 *                            inversion contracted_str_to_og[rank[i]] = i
 *                            (src/ConSGA.cpp:203-206), VertexLabelMapper
 *                            (src/align.hpp:177-202), PchQuery (ConSGA.cpp:197)
 *   chsga_score_reads     <- align_combo (src/align.hpp:470-531), which runs
 *                            align_read_combo (:399-468) once per read
 *   chsga_traceback_read  <- align_read_with_traceback (src/traceback.hpp:588-1031)
 *   chsga_score_batch     <- align_combo over a read stream, batched (no
 *                            counterpart: the reference aligns reads one by one,
 *                            src/align.hpp:488)
 *   chsga_ctx_destroy     <- the delete[] of the row buffers (src/align.hpp:526)
 *   chsga_pasgal_align    <- PaSGAL_match's alignToDAGLocal, the reference's
 *                            baseline (src/PaSGALoverrides/align.hpp:500-700)
 *
 * Rules of the boundary:
 *  - plain pointers and sizes only; no exceptions cross it; every function
 *    returns CHSGA_OK (0) or a negative status, with chsga_last_error()
 *    giving a message for the calling thread;
 *  - the library copies everything chsga_ctx_create is given into device
 *    memory; the caller may free its arrays afterwards;
 *  - the library prints nothing;
 *  - one context per host thread; calls on one context are serialised on its
 *    own HIP stream and are not re-entrant.
 *
 * Scores are exact int32 edit costs (match/subst/del/ins), identical to the
 * reference's LogScore=int arithmetic (src/ConSGA.cpp:216).
 */
#ifndef CHSGA_ABI_H
#define CHSGA_ABI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CHSGA_OK 0
#define CHSGA_EINVAL (-1)   /* bad argument or inconsistent graph */
#define CHSGA_ENOMEM (-2)   /* device or host allocation failed */
#define CHSGA_EDEVICE (-3)  /* HIP runtime error */
#define CHSGA_EIO (-4)      /* file or format error */
#define CHSGA_EINTERNAL (-5)

/* Alphabets, after src/config.hpp:3-15 and src/ParSGAoverrides/csr_align.hpp:41-50.
 * MISMATCH_N (5 symbols): sigma(a,b) = a==b ? match : subst   (N == N matches)
 * MATCH_N    (4 symbols): sigma(a,b) = (a=='N'||b=='N'||a==b) ? match : subst */
#define CHSGA_ALPHABET_MATCH_N 4
#define CHSGA_ALPHABET_MISMATCH_N 5

/* Back-pointer op codes, same values as src/traceback.hpp:168-170. */
#define CHSGA_OP_DELETE 0   /* (i-1, v) -> (i, v): read char not aligned to a graph char */
#define CHSGA_OP_MATCHSUB 1 /* (i-1, u) -> (i, v), u an in-neighbour of v             */
#define CHSGA_OP_INSERT 2   /* (i, u)   -> (i, v): graph char skipped, same read row   */

/*
 * Graph descriptor: host pointers, all owned by the caller.
 *
 * Char graph: V char vertices with labels[V] and the in-CSR in_off[V+1] /
 * in_src[E] (in-neighbours of v in ascending order, as build_in_csr produces,
 * src/align.hpp:220-242).  String graph: n_s string vertices whose chars are
 * [str_off[s], str_off[s+1]) (the -o label offsets) and out-CSR
 * str_out_off[n_s+1] / str_out_dst[m_str] (the -i graph).  Contraction
 * hierarchy (the -c file; fields of the reference's PchGraph as used at
 * src/align.hpp:294-343): CH ids 0..n_s-1, rank[og] = CH id of string og,
 * layers layer_off[num_layers+1] (CH-id ranges, one per (component, level)),
 * cc_off[num_cc+1] (layer-index ranges per connected component), up-edges
 * stored at the lower endpoint u: up_off/up_v/up_w (u -> v, v higher),
 * down-edges stored at the lower endpoint u: dn_off/dn_v/dn_w (v -> u, v
 * higher).  up_mid / dn_mid (CH id of the contracted middle vertex of a
 * shortcut, -1 for an original edge) may be NULL; they are used only to
 * expand shortcuts in traceback.
 */
typedef struct chsga_graph_desc {
  int64_t V;
  const char* labels;
  const uint64_t* in_off;
  const uint32_t* in_src;
  int64_t E;

  int64_t n_s;
  const uint64_t* str_off;
  const uint64_t* str_out_off;
  const uint32_t* str_out_dst;
  int64_t m_str;

  int64_t num_layers;
  const uint64_t* layer_off;
  int64_t num_cc;
  const uint64_t* cc_off;
  const uint32_t* rank;
  const uint64_t* up_off;
  const uint32_t* up_v;
  const int32_t* up_w;
  int64_t m_up;
  const uint64_t* dn_off;
  const uint32_t* dn_v;
  const int32_t* dn_w;
  int64_t m_dn;
  const int32_t* up_mid;
  const int32_t* dn_mid;

  int32_t alphabet;
  int32_t match_score, subst_score, del_score, ins_score;
} chsga_graph_desc;

typedef struct chsga_ctx chsga_ctx;

/* Timing and traffic figures of the last call (all optional to read). */
typedef struct chsga_stats {
  double wall_ms;          /* host wall clock of the call                  */
  double row_kernel_ms;    /* sum of row-kernel durations (profile only)    */
  int64_t row_kernel_launches;
  double sweep_ms;         /* sum of CH sweep + post-pass durations (profile)*/
  int64_t cells;           /* sum over reads of len * V                     */
  int64_t kernel_launches; /* all kernel launches of the call               */
  int64_t rows;            /* DP rows computed (forward + recompute)        */
  int64_t walk_fallbacks;  /* traceback rows that needed the staged walk    */
  int64_t peak_device_bytes;
  double forward_ms;       /* chsga_traceback_batch: wall clock of its forward passes (scores) */
  int64_t host_bytes;      /* chsga_traceback_read: checkpoint bytes in pinned host memory */
} chsga_stats;

/* Context options, chsga_ctx_set_option. */
#define CHSGA_OPT_PROFILE 1      /* 1: time row kernels with HIP events      */
#define CHSGA_OPT_USE_GRAPHS 2   /* 1: replay captured hipGraphs per rows     */
#define CHSGA_OPT_STAGED_WALK 5  /* 1: walk every traceback row through its stages (validation) */
#define CHSGA_OPT_NARROW 6       /* 1 (default): store scores as int16 when provably < 2^15; 0: int32 */
#define CHSGA_OPT_BATCHED_BACKWARD 7 /* 1 (default): chsga_traceback_batch recomputes segments of 64 reads
                                        at once when its row buffers fit; 0: read by read */
#define CHSGA_OPT_CKPT_HBM_ROWS 8 /* chsga_traceback_read: at most this many checkpoint rows in HBM, the
                                     rest in pinned host memory (-1, default: as many as fit) */
#define CHSGA_OPT_SEG_ROWS 9      /* per-read segment store: whole segments up to this many rows, else
                                     two levels (-1, default: whole when it fits in free memory) */
#define CHSGA_OPT_PACKED_ROWS 10  /* read-only check: 1 if the batched rows are packed (setting a different
                                     value is an error: the choice is made by chsga_ctx_create_ex) */
#define CHSGA_OPT_TOP_CLOSURE 11  /* 1 (default): the batched sweep and the plain path's full sweeps run the
                                     top CH levels (at most 16384 vertices batched, 262144 plain) as one launch over precomputed
                                     closure lists; 0: level by level (same results) */
#define CHSGA_OPT_RING_ROWS 12    /* chsga_traceback_batch, batched backward: at most this many rows in the
                                     ring that pass B recomputes into (-1, default: the plan with the fewest
                                     recomputed rows whose buffers fit in free memory; never fewer than
                                     ceil(sqrt(k)) rows) */
#define CHSGA_OPT_BACKWARD_TAIL 13 /* chsga_traceback_batch, batched backward: the cost of a round of 64
                                     segments in segments recomputed read by read (default 6; 0: every read
                                     in rounds).  Before each round, all reads still walking finish read by
                                     read on the plain path when their summed segments are at most this
                                     many rounds' worth, else the read with the most segments left does when
                                     the rounds it alone keeps going cost more than its segments (same
                                     paths) */
#define CHSGA_OPT_FORWARD_TAIL 14  /* batched passes (chsga_score_batch, chsga_traceback_batch's forward): once
                                     every read has started and the rows the running reads have left are at
                                     most this many times the ticks left (on average that many busy slots),
                                     with 256 ticks or more left, those reads finish one by one on the plain
                                     path (same scores, argmins and checkpoint rows; default 3; 0: never) */
#define CHSGA_OPT_BACKWARD_BAND 16 /* chsga_traceback_batch, batched backward: 1 (default) each round
                                     recomputes its segments once, its ticks stopping at the highest char
                                     its walks start from, and keeps each lane's rows as a window of 8192
                                     chars below the walk's position (> 1: a window of that many chars),
                                     walking there and redoing a segment on the plain path where the walk
                                     leaves the window; 0: whole rows (snapshots and a ring, each row
                                     recomputed about twice).  Paths are the same in every case */
#define CHSGA_OPT_CUT_CLOSURE 15   /* 1 (default on contexts created with CHSGA_CREATE_CUT_CLOSURE): the
                                     plain path closes rows on the match-cut components; 0: with the CH
                                     sweep (same values).  1 on other contexts: CHSGA_EINVAL */

const char* chsga_last_error(void);
int chsga_device_count(int* count);

/* Free and total device memory of the context's GPU, in bytes: the budget
 * for choosing a traceback's checkpoint interval.  No counterpart in the
 * reference (its checkpoints go to disk). */
int chsga_mem_info(chsga_ctx* ctx, int64_t* free_bytes, int64_t* total_bytes);

int chsga_ctx_create(const chsga_graph_desc* g, int device, chsga_ctx** out);
/* chsga_ctx_create with flags; no counterpart in the reference.
 * CHSGA_CREATE_CHAR_ROWS: the batched path stores one score per char and
 * slot even where packed rows (2-bit deltas per char, exact under unit
 * costs on compact graphs, the default there) apply.  Results are the same
 * either way. */
#define CHSGA_CREATE_CHAR_ROWS 1u
/* CHSGA_CREATE_EXACT_CLOSURE: the exact insertion closure of every row, the
 * semantics of ParSGA's serial DAG form (a topological insertion pass,
 * src/ParSGAoverrides/csr_align.hpp:316-327, row 0 without closure): the CH
 * sweep takes a multi-char source's live last value (its first value plus
 * len - 1 insertions) besides its stale one, so scores can be lower than
 * ConSGA's.  Scores are exact int32 (the reference's ParSGA drivers keep
 * uint8 scores, which wrap above 255). */
#define CHSGA_CREATE_EXACT_CLOSURE 2u
/* CHSGA_CREATE_CUT_CLOSURE: CHSGA_CREATE_EXACT_CLOSURE's values, and the
 * plain path computes each row's closure on match-cut string components
 * instead of the CH sweep: ParSGA's cut-graph insertion
 * (src/ParSGAoverrides/csr_align.hpp:422-452, cut graphs per read base as
 * read by src/ParSGAoverrides/read_csr.hpp:50-98), which drops the edges
 * into chars whose label matches the row's read char and relaxes each weakly
 * connected component on its own, in topological order.  Needs edit costs
 * with match <= subst, match <= del + ins and ins >= subst - match (where the
 * cut is exact; unit costs qualify), a char graph that expands an acyclic
 * string graph, else CHSGA_EINVAL.  Rows whose read char is not one of
 * A C G T N, or whose class has a component of more than 64 strings (N under
 * mismatchN on graphs without N labels), and the batched paths, keep the CH
 * sweep: the values are the same. */
#define CHSGA_CREATE_CUT_CLOSURE 4u
int chsga_ctx_create_ex(const chsga_graph_desc* g, int device, uint32_t flags, chsga_ctx** out);
void chsga_ctx_destroy(chsga_ctx* ctx);
int chsga_ctx_set_option(chsga_ctx* ctx, int option, int64_t value);

/* Score-only alignment of n reads, one read per launch sequence (the plain
 * path).  scores_out[n]; per-read seconds total/sub/ins may be NULL. */
int chsga_score_reads(chsga_ctx* ctx, const char* const* reads, const int64_t* lens, int64_t n,
                      int32_t* scores_out, double* total_sec, double* sub_sec, double* ins_sec,
                      chsga_stats* stats);

/* No counterpart in the reference.  chsga_score_reads for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc): the same plain path and scores,
 * with the per-call argmin buffer in a static device array of the library
 * (768 KiB per GPU, kept until the process ends) instead of a hipMalloc, so
 * the call allocates nothing.  At most CHSGA_NOALLOC_MAX_READS reads per call
 * (more: CHSGA_EINVAL).  Every context on one GPU shares that GPU's array:
 * calls that overlap on one GPU (two host threads, two contexts) race. */
#define CHSGA_NOALLOC_MAX_READS 65536
int chsga_score_reads_noalloc(chsga_ctx* ctx, const char* const* reads, const int64_t* lens, int64_t n,
                              int32_t* scores_out);

/* Last DP row of one read (V int32 values) copied to row_out; for tests. */
int chsga_last_row(chsga_ctx* ctx, const char* read, int64_t len, int32_t* row_out);

/* Batched score-only alignment: reads share slot-interleaved rows on the GPU
 * and a device work queue refills a slot when its read ends.  order[n] (may
 * be NULL) is the order in which reads enter the queue.  slots in 1..256
 * (default 64): every 64 slots form a group with its own rows and queue,
 * running on its own stream; the order is dealt to the groups by least
 * queued work. */
typedef struct chsga_batch_opts {
  int32_t slots;
  int32_t use_graphs;
  const int64_t* order;
} chsga_batch_opts;
int chsga_score_batch(chsga_ctx* ctx, const char* const* reads, const int64_t* lens, int64_t n,
                      const chsga_batch_opts* opts, int32_t* scores_out, chsga_stats* stats);

/* Schedule of the batched path: row-pass tile size and sweep launch shape.
 * No counterpart in the reference (its row loop and sweep run on CPU
 * threads); the values only change speed, never results. */
typedef struct chsga_batch_schedule {
  int32_t tile_chars; /* chars per row-pass tile, whole strings (32..4096); default 128 */
  int32_t big_level;  /* CH levels with at least this many string vertices ... */
  int32_t big_vpw;    /* ... take this many vertices per wave (1, 2, 4, 8 or 16); smaller levels take 1 */
} chsga_batch_schedule;
int chsga_get_batch_schedule(chsga_ctx* ctx, chsga_batch_schedule* out);
int chsga_set_batch_schedule(chsga_ctx* ctx, const chsga_batch_schedule* sched);

/* Autotuner of the batched schedule: times `ticks` ticks (after a few
 * untimed ones) of candidate schedules on 64 slots filled from the given
 * reads (read chars are reused cyclically when a read is shorter than the
 * run) and keeps the fastest, by coordinate descent over tile_chars, then
 * big_vpw, then big_level; a candidate displaces the schedule kept so far
 * only when it is more than 1% faster (timing noise).  results[max_results] receives every candidate
 * tried with its microseconds per tick; *n_results their number.  Scores are
 * not computed.  No counterpart in the reference. */
typedef struct chsga_tune_result {
  chsga_batch_schedule sched;
  double us_per_tick;
} chsga_tune_result;
int chsga_autotune_batch(chsga_ctx* ctx, const char* const* reads, const int64_t* lens, int64_t n, int32_t ticks,
                         chsga_tune_result* results, int32_t max_results, int32_t* n_results);

/* Asynchronous form of chsga_score_batch: the work runs on a host worker
 * thread; chsga_wait joins it.  The arrays must stay valid until chsga_wait. */
int chsga_submit_batch(chsga_ctx* ctx, const char* const* reads, const int64_t* lens, int64_t n,
                       const chsga_batch_opts* opts, int32_t* scores_out);
int chsga_wait(chsga_ctx* ctx, chsga_stats* stats);

/* One step of an alignment path: the reference's AlignOp
 * (src/traceback.hpp:185-189). */
typedef struct chsga_align_op {
  int64_t row;
  uint32_t vertex;
  uint8_t op;
} chsga_align_op;

typedef struct chsga_path {
  int64_t n_ops;
  chsga_align_op* ops;
  int32_t score;
  uint32_t end_vertex;
} chsga_path;

/* Checkpointed traceback of one read: rows 0, k, 2k, ... are kept in HBM,
 * segments are recomputed on the GPU and walked by score equality with the
 * reference's predecessor priority (src/traceback.hpp:524-569).  k >= 1. */
int chsga_traceback_read(chsga_ctx* ctx, const char* read, int64_t len, int64_t k, chsga_path** out,
                         chsga_stats* stats);
void chsga_path_free(chsga_path* p);

/* Batched checkpointed traceback: the batched forward pass (slots, device
 * work queue, graphs as in chsga_score_batch) keeps rows 0, k, 2k, ... of
 * every read and the argmin of its last row in HBM; each read's path is then
 * recovered by segment recompute and walk on the GPU, as in
 * chsga_traceback_read.  Reads enter in opts->order, in waves whose
 * checkpoint rows fit budget_bytes (0: half the free device memory after the
 * batch buffers); a read whose rows alone exceed it is an error (use a larger
 * k).  paths[n] receives one path per read (free each with chsga_path_free).
 * Replaces align_combo_with_traceback (src/traceback.hpp:1040-1053). */
int chsga_traceback_batch(chsga_ctx* ctx, const char* const* reads, const int64_t* lens, int64_t n, int64_t k,
                          const chsga_batch_opts* opts, int64_t budget_bytes, chsga_path** paths, chsga_stats* stats);

/*
 * PaSGAL_match, the reference's baseline aligner (Makefile:197-213 builds
 * PaSGAL's scalar code with src/PaSGALoverrides/align.hpp in place of its
 * own align.hpp): local alignment of each read to a char DAG, maximising
 * match - mismatch - ins - del scores with a floor of 0.
 *
 *   chsga_pasgal_align  <- alignToDAGLocal (src/PaSGALoverrides/align.hpp:500-700):
 *                          phase 1 on the read and its reverse complement
 *                          (alignToDAGLocal_Phase1_scalar, :43-132, N matches
 *                          any char), the strand with the higher score ('-'
 *                          on ties, :601-613), phase 1 in reverse for the
 *                          start (alignToDAGLocal_Phase1_rev_scalar,
 *                          :146-245, exact match) and phase 2, the DP
 *                          recomputed in the [start, end] window and walked
 *                          into a CIGAR (alignToDAGLocal_Phase2, :256-491,
 *                          exact match, cigarCompact)
 *   chsga_pasgal_free   <- the release of BestScoreInfo's cigar strings
 *
 * Each DP row is the plain path's row in negated scores, closed by the exact
 * insertion closure; fwd and rev must be created with
 * CHSGA_CREATE_EXACT_CLOSURE, fwd on the char DAG and rev on its reversal
 * (char j becomes V - 1 - j; chsga_host_reverse builds it).  Char ids must
 * be a topological order (every in-neighbour id smaller), as PaSGAL's loader
 * numbers its CSR_char_container.  rev may be NULL when score_only.
 */
typedef struct chsga_pasgal_params {
  int32_t match, mismatch, ins, del; /* PaSGAL's Parameters, each 1..63 (PaSGAL's defaults are 1; the
                                        bound keeps its int8 vertical differences exact, align.hpp:347) */
  int32_t score_only;                /* 1: phase 1 only, as PaSGAL built with PASGAL_SCORE_ONLY */
} chsga_pasgal_params;

/* result flags */
#define CHSGA_PASGAL_EMPTY_WINDOW 1 /* phase 1-R put the start after the end: no phase 2 (the reference
                                       would size a negative window) */
#define CHSGA_PASGAL_PIN_MOVED 2    /* phase 1-R: the forward end cell, set to match + 1, was raised by the
                                       closure, i.e. the reference's assertion currentMax == match
                                       (align.hpp:221) fails; only where the forward pass's N-matches
                                       disagree with the reverse pass's exact matches */

typedef struct chsga_pasgal_result {
  int32_t score;                 /* BestScoreInfo.score of the chosen strand */
  char strand;                   /* '+' or '-' */
  int32_t flags;
  int64_t qry_start, qry_end;    /* qryRowStart, qryRowEnd: rows of the chosen strand's read (start -1 when score_only) */
  int64_t ref_start, ref_end;    /* refColumnStart, refColumnEnd: char ids */
  int32_t score_fwd, score_rc;   /* phase 1's best on each strand */
  int64_t qry_end_fwd, ref_end_fwd, qry_end_rc, ref_end_rc;
  int32_t rev_best;              /* phase 1-R's best (score + 1 when the assertion above holds) */
  char* cigar;                   /* compact CIGAR of =, X, I, D (NULL when score_only); chsga_pasgal_free */
} chsga_pasgal_result;

int chsga_pasgal_align(chsga_ctx* fwd, chsga_ctx* rev, const chsga_pasgal_params* params, const char* const* reads,
                       const int64_t* lens, int64_t n, chsga_pasgal_result* out, chsga_stats* stats);
void chsga_pasgal_free(chsga_pasgal_result* res, int64_t n);

/*
 * Learned cost model (no counterpart in the reference, which schedules
 * nothing): an FP32 multilayer perceptron in -> h1 -> h2 -> out with ReLU
 * hidden layers and mean-squared-error loss, trained on the GPU by the
 * library's own forward, backward and Adam kernels.  Parameters are flat:
 * W1[h1*in], b1[h1], W2[h2*h1], b2[h2], W3[out*h2], b3[out] (row-major,
 * W[o*I + i]).  X is row-major [n][in], T and Y are [n][out].
 */
typedef struct chsga_mlp chsga_mlp;
int chsga_mlp_create(int32_t n_in, int32_t n_h1, int32_t n_h2, int32_t n_out, uint64_t seed, int device,
                     chsga_mlp** out);
void chsga_mlp_destroy(chsga_mlp* m);
int64_t chsga_mlp_num_params(const chsga_mlp* m);
int chsga_mlp_get_params(chsga_mlp* m, float* out, int64_t n);
int chsga_mlp_set_params(chsga_mlp* m, const float* in, int64_t n);
int chsga_mlp_forward(chsga_mlp* m, const float* X, int64_t n, float* Y);
/* loss = mean squared error over n*out values; grad[num_params] its gradient */
int chsga_mlp_loss_grad(chsga_mlp* m, const float* X, const float* T, int64_t n, float* loss, float* grad);
int chsga_mlp_train(chsga_mlp* m, const float* X, const float* T, int64_t n, int32_t epochs, int32_t batch, float lr,
                    uint64_t seed, float* final_loss);

#ifdef __cplusplus
}
#endif
#endif
