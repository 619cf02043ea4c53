// Context of the MI355X engine (private to the library).
#pragma once

#include <hip/hip_runtime.h>

#include <functional>
#include <memory>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "chsga/abi.h"
#include "kernels.hip.inc"

namespace chsga_engine {

struct HipError : std::runtime_error {
  using std::runtime_error::runtime_error;
};
void set_error(const std::string& m);
void hip_check(hipError_t e, const char* what);

// Pinned host memory (hipHostMalloc), released with the owner.
struct HostMem {
  void* p = nullptr;
  size_t n = 0;
  HostMem() = default;
  HostMem(const HostMem&) = delete;
  HostMem& operator=(const HostMem&) = delete;
  ~HostMem() { release(); }
  void alloc(size_t bytes);
  void release();
};

struct DevMem {
  void* p = nullptr;
  size_t n = 0;
  DevMem() = default;
  DevMem(const DevMem&) = delete;
  DevMem& operator=(const DevMem&) = delete;
  ~DevMem() { release(); }
  void alloc(size_t bytes);
  void release();
};

}  // namespace chsga_engine

struct chsga_batch_state;

// Top levels of the CH handled by one LDS-resident workgroup.
constexpr int64_t kTopPlain = 4096;  // string vertices: 2 x 4 B each in LDS
// entries the batched level sweep may read past a table's end (>= its vertices per wave)
constexpr int64_t kSweepPad = 16;
// largest number of pull entries the batched level sweep loads per vertex at
// once (one vertex per wave, on levels whose lists exceed two entries)
constexpr int kSweepChunk = 16;
// words of a batched sweep pull record with R pull slots: first char, R
// sources, R 16-bit weights, rounded up to a power of two
constexpr int sweep_record_words(int R) { return R == 2 ? 4 : R == 4 ? 8 : 16; }
// most vertices of the top CH levels the batched sweep runs as one closure
constexpr int64_t kTopClosureMax = 16384;
constexpr int64_t kTopClosurePlainMax = 262144;  // the plain path's closure
// most {source, weight} terms of that closure; more (long relay chains, as
// exact closure can give) and the top levels keep their level launches
constexpr int64_t kTopClosureTermsMax = int64_t(1) << 22;
// chars the batched row buffers hold past V (>= one row-pass chunk)
constexpr int64_t kRowPad = 32;
struct TopSet {
  int32_t j_top = 0;
  uint32_t K0 = 0, NT = 0;
  uint32_t n_upi = 0, n_dni = 0;  // internal pull counts
  chsga_engine::DevMem ext_off, ext, upi_off, upi, dni_off, dni;
};

// cut closure (CHSGA_CREATE_CUT_CLOSURE): read chars A, C, G, T, N
constexpr int kCutClasses = 5;
// largest component (strings) one thread of the cut closure relaxes; a class
// with a larger one keeps the CH sweep
constexpr uint32_t kCutMaxStrings = 64;
constexpr uint32_t kCutWindow = 65536;  // chars per window of the cut closure's component order
struct CutClass {
  bool ok = false;  // the class runs the cut closure
  uint32_t n_items = 0, n_comp = 0, largest = 0;
  std::vector<uint4> h_items;  // until uploaded
  std::vector<uint32_t> h_comp_off;
  chsga_engine::DevMem items, comp_off;
};

struct OgEdge {
  uint32_t v;
  int32_t w;
  int32_t mid;
};

struct chsga_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  // graph
  int64_t V = 0, E = 0, n_s = 0;
  int32_t alphabet = 5, msc = 0, ssc = 1, dsc = 1, isc = 1;
  int64_t n_tiles = 0;
  int32_t J = 0, j_top = 0;
  int64_t top_level_opt = -1;
  int64_t tile_chars = 2040;  // one 256 x 8 block round even when c0 is not 8-aligned
  int64_t batch_tile_chars = 128;   // batched schedule (chsga_batch_schedule)
  int64_t batch_big_level = 16384;
  int32_t batch_big_vpw = 8;
  std::vector<uint32_t> h_lvl_off, h_up_off, h_dn_off;
  // string of each sweep position (levels in turn, each in ascending string
  // id, so on a graph whose char ids are topological the strings below a
  // char form a prefix of every level)
  std::vector<uint32_t> h_sw_vert;
  // host copies for traceback
  std::string h_labels;
  std::vector<uint32_t> h_str_off, h_in_off, h_in_src, h_sin_off, h_sin_src;
  bool has_mid = false;
  std::vector<std::vector<OgEdge>> h_up_og, h_dn_og, h_upin_og;  // CH edges by original string id
  chsga_engine::DevMem d_sin_off, d_sin_src;
  // device graph
  chsga_engine::DevMem d_labels, d_in_off, d_in_src, d_first_bits, d_tile_char, d_tile_str, d_str_first;
  bool nb2_all = false;  // compact, and no first char has more than 2 in-neighbours
  bool batch_fast = false;  // nb2_all and every label in ACGTN: the batched compact row kernel
  bool allow_packed = true;  // chsga_ctx_create_ex without CHSGA_CREATE_CHAR_ROWS
  bool exact = false;        // CHSGA_CREATE_EXACT_CLOSURE: live last values in the sweep
  bool packed = false;       // batched rows packed as 2-bit deltas (batch_fast, unit costs)
  bool cut = false;          // CHSGA_CREATE_CUT_CLOSURE: cut tables built
  bool cut_on = false;       // CHSGA_OPT_CUT_CLOSURE: the plain path runs them
  CutClass cut_cls[kCutClasses];
  std::vector<int4> h_pbnb;        // packed: per string {string, length - 1} of its two in-neighbour strings
  std::vector<uint8_t> h_lclass;   // packed: match class of each char's label
  chsga_engine::DevMem d_cstr;  // packed: string of each char (batched traceback walk)
  chsga_engine::DevMem d_lclass;  // match class of each char's label (batched compact kernel)
  chsga_engine::DevMem d_first_nb;  // plain compact kernel: first two in-neighbours of each first char
  chsga_engine::DevMem d_bnb;     // per string: {last, first} chars of <= 2 in-neighbour strings
  chsga_engine::DevMem d_cfirst;  // first char of each char's string (general batched kernel)
  chsga_engine::DevMem d_sw_vert, d_lvl_off, d_up_off, d_up_in, d_dn_off, d_dn_in;
  // batched sweep: first char of the vertex at each sweep position; pull
  // lists with the last char of each source string
  chsga_engine::DevMem d_vfirst, d_bup_in, d_bdn_in;
  std::vector<uint32_t> h_up_maxp, h_dn_maxp;  // longest pull list of each level
  // pull records {first, src_0 .. src_R-1, packed 16-bit weights} of levels
  // with h_*_rec[j] = R in {2, 4, 8} (0: the level takes the pull lists),
  // level j's at word h_*_rbase[j]
  chsga_engine::DevMem d_bup_rec, d_bdn_rec;
  // batched sweep: levels top_l0 .. J - 1 as one launch over precomputed
  // {source, summed weight} lists by sweep position (top_l0 == J: none)
  int32_t top_l0 = 0;
  bool top_closure = true;  // CHSGA_OPT_TOP_CLOSURE
  uint32_t top_maxp = 0;
  chsga_engine::DevMem d_top_off, d_top_in;
  // plain sweep: levels top_l0_plain .. J - 1 as one launch over lists keyed
  // by string code (F / Lst), from a larger top set (kTopClosurePlainMax)
  int32_t top_l0_plain = 0;
  chsga_engine::DevMem d_top_off_plain, d_top_in_plain;
  std::vector<uint8_t> h_up_rec, h_dn_rec;
  std::vector<uint64_t> h_up_rbase, h_dn_rbase;
  TopSet top_plain;
  // working rows
  chsga_engine::DevMem d_F, d_Lst, d_row[2];
  void* d_scratch = nullptr;  // this device's copy of the no-allocation scratch array
  // options and counters
  bool profile = false, use_graphs = false, force_staged = false;
  bool batched_backward = true;  // batched traceback: lock-step segment recompute when it fits
  int64_t ckpt_hbm_rows = -1;     // per-read traceback: checkpoint rows kept in HBM (-1: all that fit)
  int64_t seg_rows_cap = -1;      // per-read traceback: segment store rows (-1: by free memory)
  int64_t ring_rows_cap = -1;     // batched backward: ring rows at most (-1: by free memory)
  int64_t backward_tail = 6;      // batched backward: tail read by read (0: never)
  int64_t backward_band = 1;      // batched backward: banded rounds of that many chars (1: the default width; 0: off)
  int64_t forward_tail = 3;       // batched forward: tail on the plain path (0: never)
  bool compact = false;  // char graph is the expansion of the string graph
  int topo_ids = -1;     // char ids are a topological order (1), not (0), unknown (-1)
  uint64_t serial = 0;   // unique per context of the process
  uint64_t pasgal_rev = 0;  // serial of the context last verified as this one's reversal (0: none)
  bool allow_narrow = true, narrow = false;  // int16 row storage (exact, see fits_narrow)
  int64_t launches = 0, rows = 0;
  std::vector<hipEvent_t> ev_pool;
  size_t ev_used = 0;
  double prof_row_ms = 0, prof_ins_ms = 0;
  int64_t prof_rows = 0;
  // asynchronous batch
  std::thread worker;
  int worker_status = 0;
  chsga_stats worker_stats{};
  std::string worker_error;
  std::shared_ptr<chsga_batch_state> batch;

  ~chsga_ctx();
  void launch_init(void* H, uint8_t c0);
  void launch_row(const void* H, void* D, uint8_t rc);
  void launch_insertion(void* D, uint8_t rc);
  void launch_argmin(const void* D, unsigned long long* key);
  bool fits_narrow(int64_t max_len) const;
  void launch_sweeps(int mode);  // bit 0: up levels, bit 1: down levels
  void ensure_traceback_tables();
  void ensure_staged_tables(const chsga_graph_desc* g);
  void staged_walk_row(const int32_t* prow_dev, const int32_t* fin_dev, const char* read, int64_t i, int64_t v0,
                       std::vector<chsga_align_op>& rev, int64_t* v_out);
  int forward(const char* read, int64_t len, int64_t k, int32_t** ckpt_rows);
  void flush_profile();
};
