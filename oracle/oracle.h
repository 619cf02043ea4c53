/*
 * oracle.h -- CPU restatement of ConSGA's per-read graph DP.
 *
 * TEST INFRASTRUCTURE ONLY.  This code is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and
 * never as the thing measured or shipped.  The product path is the HIP
 * engine behind include/chsga/abi.h.
 *
 *   O2  oracle_consga_*   ConSGA exactly as src/align.hpp:177-531 computes it
 *                         (row kernel, within-vertex pre-pass, per-component
 *                         CH up/down layer sweeps with the stale last(.) of
 *                         multi-char vertices, post-pass), OpenMP-parallel in
 *                         the same places as the reference (blocked 4096-vertex
 *                         row loop, parallel string vertices, parallel layer).
 *   O1  oracle_exact_*    the exact insertion closure (topological relaxation
 *                         over the char DAG), the textbook recurrence of
 *                         src/ParSGAoverrides/csr_align.hpp:319-327 with int32.
 *   O3  oracle_traceback  walk_traceback_from_scores (src/traceback.hpp:483-586)
 *                         over the full score matrix; oracle_validate_path is a
 *                         path checker stronger than ConSGA_traceback.cpp:184-221
 *                         (it also checks every step follows a char edge).
 *   O5  oracle_pasgal_*   PaSGAL_match's local alignment, the reference's
 *                         scalar baseline (src/PaSGALoverrides/align.hpp).
 */
#ifndef CHSGA_ORACLE_H
#define CHSGA_ORACLE_H

#include <stdint.h>

#include "chsga/abi.h"

#ifdef __cplusplus
extern "C" {
#endif

int oracle_threads(int n);

int oracle_consga_scores(const chsga_graph_desc* g, const char* reads, const int64_t* off, int64_t n,
                         int32_t* scores);
int oracle_consga_rows(const chsga_graph_desc* g, const char* read, int64_t len, int32_t* rows_out /* len*V */);
int oracle_consga_last_row(const chsga_graph_desc* g, const char* read, int64_t len, int32_t* row_out);

int oracle_exact_last_row(const chsga_graph_desc* g, const char* read, int64_t len, int32_t* row_out);
/* ParSGA's match-cut insertion, serially (edges into chars labelled N or the
 * row's read char dropped); equals oracle_exact_last_row where the cut is
 * exact. */
int oracle_cut_last_row(const chsga_graph_desc* g, const char* read, int64_t len, int32_t* row_out);
int oracle_exact_scores(const chsga_graph_desc* g, const char* reads, const int64_t* off, int64_t n,
                        int32_t* scores);

/* Traceback on the full ConSGA score matrix; returns the number of ops
 * written (<= cap), or -1 when a cell has no predecessor (the reference
 * throws there), -2 on bad arguments. */
int64_t oracle_traceback(const chsga_graph_desc* g, const char* read, int64_t len, chsga_align_op* ops,
                         int64_t cap, int32_t* score_out);

/* 0 when the path is valid and rescores to `score`; otherwise a positive
 * code naming the first failed check:
 *   1 empty path, 2 first op not (row 0, MATCHSUB), 3 bad step (row/edge),
 *   4 vertex out of range, 5 last consuming row != len-1, 6 cost != score */
int oracle_validate_path(const chsga_graph_desc* g, const char* read, int64_t len, const chsga_align_op* ops,
                         int64_t n_ops, int32_t score, int64_t* bad_index);

/* O5: PaSGAL_match's alignToDAGLocal (src/PaSGALoverrides/align.hpp:500-700),
 * scalar, one read per OpenMP thread: phase 1 on both strands, phase 1-R,
 * phase 2 and the compact CIGAR (oracle/pasgal.cpp).  out[n]; free the
 * CIGARs with oracle_pasgal_free.  CHSGA_EINVAL when the char ids are not a
 * topological order or a read has a char outside ACGTN. */
int oracle_pasgal_align(const chsga_graph_desc* g, const char* reads, const int64_t* off, int64_t n,
                        const chsga_pasgal_params* params, chsga_pasgal_result* out);
void oracle_pasgal_free(chsga_pasgal_result* res, int64_t n);

/* sigma(a, b) for the given alphabet (4 = matchN, 5 = mismatchN). */
int32_t oracle_sigma(int32_t alphabet, char a, char b, int32_t match, int32_t subst);

#ifdef __cplusplus
}
#endif
#endif
