/*
 * chsga/host.h -- host-side C ABI: input formats, the seeded synthetic
 * graph/read generator and the contraction-hierarchy builder.
 *
 * Counterparts in the reference:
 *   chsga_host_load        <- the drivers' input stage: read_graph (-I, -i),
 *                             read_contracted_graph (-c), the labels reader
 *                             (src/ConSGA.cpp:159-170) and the label-offsets
 *                             reader (src/ConSGA.cpp:175-185)
 *   chsga_host_build_ch    <- build_pch (src/CHoverrides/build_pch.cpp:94-130,
 *                             build_pch.hpp:912-1304), vertex-weighted
 *   chsga_host_save        <- csr_to_adj (src/csr_to_adj.cpp:46-114) and
 *                             write_pbbs_format (build_pch.cpp:126)
 *   chsga_host_load_reads  <- load_reads (src/load_reads.hpp:7-40)
 *   chsga_host_load_gfa,
 *   chsga_host_load_string_csr,
 *   chsga_host_write_csr   <- handlegraph_to_csr (src/handlegraph_to_csr.cpp:39-150
 *                             and its char expansion, :165-247), reading GFA
 *                             in place of a libbdsg PackedGraph
 *   chsga_host_reverse     <- the reversed traversal of PaSGAL's phase 1-R
 *                             (src/PaSGALoverrides/align.hpp:181-197)
 *   chsga_host_generate,
 *   chsga_host_gen_reads   <- no counterpart (the reference ships no data)
 */
#ifndef CHSGA_HOST_H
#define CHSGA_HOST_H

#include <stdint.h>

#include "chsga/abi.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct chsga_dataset chsga_dataset;
typedef struct chsga_readset chsga_readset;

/* Seeded chain-of-bubbles DAG: uniform ACGT backbone, SNP bubbles, indel
 * bubbles (50/50 deletion allele or insertion allele, each with a skip edge)
 * and structural bubbles, at geometric gaps. */
typedef struct chsga_graph_params {
  uint64_t seed;
  int64_t backbone_bp;
  double snp_rate;   /* events per bp */
  double indel_rate; /* events per bp */
  int32_t indel_min, indel_max;
  double sv_rate; /* events per bp */
  int32_t sv_min, sv_max;
} chsga_graph_params;

typedef struct chsga_dataset_info {
  int64_t V, E, n_s, m_str, m_up, m_dn, num_layers, num_cc, max_level, max_str_len;
} chsga_dataset_info;

const char* chsga_host_last_error(void);

int chsga_host_generate(const chsga_graph_params* p, chsga_dataset** out);
/* Builds (or rebuilds) the contraction hierarchy of the string graph with
 * vertex weights len(s)-1 (shortcut weight w1 + w2 + len(mid) - 1). */
int chsga_host_build_ch(chsga_dataset* ds, uint64_t seed);
/* Drop-in loader.  ch_path may be NULL (then the CH is built). */
int chsga_host_load(const char* char_adj, const char* str_adj, const char* ch_path, const char* labels,
                    const char* offsets, chsga_dataset** out);
/* Writes <prefix>.char.adj, .str.adj, .ch, .labels, .label_offsets. */
int chsga_host_save(const chsga_dataset* ds, const char* prefix);
/* GFA 1 input (S and L lines; link overlaps * or 0M): every segment in both
 * orientations, as handlegraph_to_csr lists a handle graph
 * (src/handlegraph_to_csr.cpp:39-110): segments ordered by id (the numeric
 * name, or file order when a name is not a number), segment k's forward
 * strand as string vertex 2k and its reverse complement as 2k + 1, a link
 * a -> b also giving flip(b) -> flip(a), out-lists sorted; then the char
 * expansion (handlegraph_to_csr.cpp:165-247) and the CH. */
int chsga_host_load_gfa(const char* path, chsga_dataset** out);
/* The string-labelled CSR text handlegraph_to_csr writes (write_csr,
 * src/handlegraph_to_csr.cpp:112-150): n, m, n labels, n + 1 offsets, m
 * destinations; then the char expansion and the CH. */
int chsga_host_load_string_csr(const char* path, chsga_dataset** out);
/* The string graph and its char expansion as two CSR texts in that format
 * (handlegraph_to_csr's two outputs). */
int chsga_host_write_csr(const chsga_dataset* ds, const char* string_csr, const char* char_csr);
int chsga_host_save_ch(const chsga_dataset* ds, const char* path);
/* build_pch: CH of a string graph file; label_offsets may be NULL, which
 * gives every string vertex length 1 (no vertex weights, as build_pch without
 * -v, src/CHoverrides/build_pch.cpp:96-114). */
int chsga_host_build_ch_file(const char* str_adj, const char* label_offsets, const char* out_ch, uint64_t seed);
int chsga_host_describe(const chsga_dataset* ds, int32_t alphabet, chsga_graph_desc* out);
/* The reversed graph PaSGAL's phase 1-R aligns to (it walks chars V-1..0
 * over out-edges, src/PaSGALoverrides/align.hpp:181-197): string s becomes
 * n_s - 1 - s with its label reversed and every edge turned around, so char
 * j becomes V - 1 - j; with its own CH.  The char graph must expand the
 * string graph. */
int chsga_host_reverse(const chsga_dataset* ds, uint64_t seed, chsga_dataset** out);
int chsga_host_info(const chsga_dataset* ds, chsga_dataset_info* out);
void chsga_host_free(chsga_dataset* ds);

/* Read generation: windows of random source-to-sink paths (uniform branch
 * choice) with i.i.d. substitutions, insertions, deletions and N characters.
 * len_mode 0: fixed len_a; 1: log-normal, median len_a, sigma len_b, clipped
 * to [len_min, len_max]; 2: mixture: with probability mix_frac a uniform
 * length in [len_min, len_max] with error rates scaled by err_scale_a,
 * otherwise uniform in [len2_min, len2_max] scaled by err_scale_b. */
typedef struct chsga_read_params {
  uint64_t seed;
  int64_t n_reads;
  int32_t len_mode;
  double len_a, len_b;
  int64_t len_min, len_max;
  double sub_rate, ins_rate, del_rate, n_rate;
  double mix_frac;
  int64_t len2_min, len2_max;
  double err_scale_a, err_scale_b;
} chsga_read_params;

int chsga_host_gen_reads(const chsga_dataset* ds, const chsga_read_params* p, chsga_readset** out);
/* FASTA/FASTQ, optionally gzip-compressed, or one sequence per line.
 * Sequences are upper-cased.  cap > 0 stops after cap reads (the reference
 * stops after 5, src/load_reads.hpp:5). */
int chsga_host_load_reads(const char* path, int64_t cap, chsga_readset** out);
int chsga_host_save_reads(const chsga_readset* rs, const char* path);
int64_t chsga_reads_count(const chsga_readset* rs);
const char* chsga_reads_data(const chsga_readset* rs);
const int64_t* chsga_reads_offsets(const chsga_readset* rs);
void chsga_reads_free(chsga_readset* rs);

/* Path verification <- src/ConSGA_traceback.cpp:184-225 (rescoring of the
 * path with the edit costs, last consuming row = len - 1), with the
 * structure it assumes also checked.  Returns 0 when the path is valid and
 * rescores to `score`, else the first failed check: 1 empty path, 2 first op
 * not (row 0, MATCHSUB), 3 a step that does not follow a char edge or
 * advances the row wrongly, 4 vertex / row / op out of range, 5 last
 * consuming row != len - 1, 6 cost != score; CHSGA_EINVAL on null
 * arguments.  *bad_index (optional) names the failing op (-1 when valid). */
int chsga_host_verify_path(const chsga_graph_desc* g, const char* read, int64_t len, const chsga_align_op* ops,
                           int64_t n_ops, int32_t score, int64_t* bad_index);

/* Direct array access for tests and tools (pointers into the dataset). */
const char* chsga_ds_labels(const chsga_dataset* ds);
const uint64_t* chsga_ds_str_off(const chsga_dataset* ds);

#ifdef __cplusplus
}
#endif
#endif
