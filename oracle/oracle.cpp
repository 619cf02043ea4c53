// CPU restatement of ConSGA (the oracle).  TEST INFRASTRUCTURE ONLY: loaded
// by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg, never by
// the product path.  Every function cites the reference code it follows.
#include "oracle.h"

#include <omp.h>

#include <algorithm>
#include <climits>
#include <cstring>
#include <vector>

namespace {

// sigma: src/ParSGAoverrides/csr_align.hpp:41-50 (ALPHABET_SIZE 5 / 4)
inline int32_t sigma(int32_t alphabet, char a, char b, int32_t m, int32_t s) {
  if (alphabet == CHSGA_ALPHABET_MISMATCH_N) return (a == b) ? m : s;
  return (a == 'N' || b == 'N' || a == b) ? m : s;
}

// cas_update as an atomic min: src/ParSGAoverrides/csr_align.hpp:385-402
inline void atomic_min(int32_t* t, int32_t v) {
  int32_t cur = __atomic_load_n(t, __ATOMIC_RELAXED);
  while (v < cur && !__atomic_compare_exchange_n(t, &cur, v, false, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) {
  }
}

struct Model {
  const chsga_graph_desc* g;
  std::vector<uint32_t> ch_to_og;  // src/ConSGA.cpp:203-206
  explicit Model(const chsga_graph_desc* gg) : g(gg), ch_to_og(gg->n_s) {
    for (int64_t i = 0; i < g->n_s; ++i) ch_to_og[g->rank[i]] = uint32_t(i);
  }
  // VertexLabelMapper<true>: src/align.hpp:177-202
  int64_t first(int64_t s) const { return int64_t(g->str_off[s]); }
  int64_t last(int64_t s) const { return int64_t(g->str_off[s + 1]) - 1; }

  // init_func_parallel: src/align.hpp:205-210
  void init_row(int32_t* H, char c0) const {
#pragma omp parallel for schedule(static)
    for (int64_t v = 0; v < g->V; ++v) H[v] = sigma(g->alphabet, g->labels[v], c0, g->match_score, g->subst_score);
  }

  // fused_deletion_substitution: src/align.hpp:245-277 (blocked 4096)
  void del_sub(int32_t* D, const int32_t* H, char rc) const {
    const int64_t V = g->V, nb = (V + 4095) / 4096;
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t b = 0; b < nb; ++b) {
      const int64_t lo = b * 4096, hi = std::min(V, lo + 4096);
      for (int64_t v = lo; v < hi; ++v) {
        int32_t best = H[v] + g->del_score;
        const int32_t delta = sigma(g->alphabet, g->labels[v], rc, g->match_score, g->subst_score);
        for (uint64_t e = g->in_off[v]; e < g->in_off[v + 1]; ++e) {
          const int32_t cand = H[g->in_src[e]] + delta;
          if (cand < best) best = cand;
        }
        D[v] = best;
      }
    }
  }

  // insertionQueryLayerParallelization: src/align.hpp:279-345
  void sweep(int32_t* D) const {
    const int32_t ins = g->ins_score;
    for (int64_t cc = 0; cc < g->num_cc; ++cc) {
      // forward: layers low -> high, push last(u) + w*ins into first(v)
      for (uint64_t layer = g->cc_off[cc]; layer < g->cc_off[cc + 1]; ++layer) {
        const int64_t lo = int64_t(g->layer_off[layer]), hi = int64_t(g->layer_off[layer + 1]);
#pragma omp parallel for schedule(dynamic, 256)
        for (int64_t u = lo; u < hi; ++u) {
          const int64_t u_ix = last(ch_to_og[u]);
          for (uint64_t j = g->up_off[u]; j < g->up_off[u + 1]; ++j) {
            const int64_t v_ix = first(ch_to_og[g->up_v[j]]);
            atomic_min(&D[v_ix], D[u_ix] + g->up_w[j] * ins);
          }
        }
      }
      // backward: layers high -> low, first(u) pulls last(v) + w*ins
      for (uint64_t layer = g->cc_off[cc + 1]; layer-- > g->cc_off[cc];) {
        const int64_t lo = int64_t(g->layer_off[layer]), hi = int64_t(g->layer_off[layer + 1]);
#pragma omp parallel for schedule(dynamic, 256)
        for (int64_t u = lo; u < hi; ++u) {
          const int64_t u_ix = first(ch_to_og[u]);
          int32_t t = D[u_ix];
          for (uint64_t j = g->dn_off[u]; j < g->dn_off[u + 1]; ++j) {
            const int32_t cand = D[last(ch_to_og[g->dn_v[j]])] + g->dn_w[j] * ins;
            if (cand < t) t = cand;
          }
          D[u_ix] = t;
        }
      }
    }
  }

  // insertion_func_parallel: src/align.hpp:347-397
  void insertion(int32_t* D) const {
    const int32_t ins = g->ins_score;
#pragma omp parallel for schedule(dynamic, 1024)
    for (int64_t s = 0; s < g->n_s; ++s)  // pre-pass :362-373
      for (int64_t v = first(s); v < last(s); ++v)
        if (D[v] + ins < D[v + 1]) D[v + 1] = D[v] + ins;
    sweep(D);
#pragma omp parallel for schedule(dynamic, 1024)
    for (int64_t s = 0; s < g->n_s; ++s) {  // post-pass :381-396
      int64_t v = first(s);
      const int64_t end = last(s);
      while (v < end) {
        if (D[v] + ins < D[v + 1]) {
          D[v + 1] = D[v] + ins;
          ++v;
        } else {
          break;
        }
      }
    }
  }

  // align_read_combo: src/align.hpp:399-468; rows_out (optional) keeps all rows
  // as the score-matrix fast path does (src/traceback.hpp:670-693).
  void run(const char* read, int64_t len, int32_t* last_row, int32_t* rows_out) const {
    const int64_t V = g->V;
    std::vector<int32_t> buf(rows_out ? 0 : 2 * V);
    int32_t* prev = rows_out ? rows_out : buf.data();
    init_row(prev, read[0]);
    for (int64_t i = 1; i < len; ++i) {
      int32_t* cur = rows_out ? rows_out + i * V : (prev == buf.data() ? buf.data() + V : buf.data());
      del_sub(cur, prev, read[i]);
      insertion(cur);
      prev = cur;
    }
    if (last_row) std::memcpy(last_row, prev, sizeof(int32_t) * V);
  }

  // O1: exact closure D'[v] = min(D[v], min_{u in In(v)} D'[u] + ins) in a
  // topological order of the char DAG.
  std::vector<uint32_t> topo() const {
    const int64_t V = g->V;
    std::vector<uint32_t> indeg(V), order;
    std::vector<std::vector<uint32_t>> outl(V);
    for (int64_t v = 0; v < V; ++v)
      for (uint64_t e = g->in_off[v]; e < g->in_off[v + 1]; ++e) {
        outl[g->in_src[e]].push_back(uint32_t(v));
        ++indeg[v];
      }
    order.reserve(V);
    for (int64_t v = 0; v < V; ++v)
      if (!indeg[v]) order.push_back(uint32_t(v));
    for (size_t k = 0; k < order.size(); ++k)
      for (uint32_t w : outl[order[k]])
        if (--indeg[w] == 0) order.push_back(w);
    return order;
  }
  void run_exact(const char* read, int64_t len, int32_t* last_row, const std::vector<uint32_t>& order) const {
    const int64_t V = g->V;
    std::vector<int32_t> a(V), b(V);
    init_row(a.data(), read[0]);
    for (int64_t i = 1; i < len; ++i) {
      del_sub(b.data(), a.data(), read[i]);
      for (uint32_t v : order)
        for (uint64_t e = g->in_off[v]; e < g->in_off[v + 1]; ++e) {
          const int32_t c = b[g->in_src[e]] + g->ins_score;
          if (c < b[v]) b[v] = c;
        }
      std::swap(a, b);
    }
    std::memcpy(last_row, a.data(), sizeof(int32_t) * V);
  }

  // ParSGA's cut-graph insertion, serially: row i relaxes only the edges of
  // the match-cut graph of r[i], whose edge u -> v survives iff L[v] is
  // neither 'N' nor r[i] (src/ParSGAoverrides/preprocessing.hpp:122-126),
  // in a topological order (the DAG branch of insertion,
  // src/ParSGAoverrides/csr_align.hpp:436-447, per component; the serial
  // order relaxes the same edges with every dependency before it).  Under
  // matchN with read char N nothing is relaxed (csr_align.hpp:584-590).
  void run_cut(const char* read, int64_t len, int32_t* last_row, const std::vector<uint32_t>& order) const {
    const int64_t V = g->V;
    std::vector<int32_t> a(V), b(V);
    init_row(a.data(), read[0]);
    for (int64_t i = 1; i < len; ++i) {
      del_sub(b.data(), a.data(), read[i]);
      const bool skip = g->alphabet == CHSGA_ALPHABET_MATCH_N && read[i] == 'N';
      for (uint32_t v : order) {
        if (skip || g->labels[v] == 'N' || g->labels[v] == read[i]) continue;
        for (uint64_t e = g->in_off[v]; e < g->in_off[v + 1]; ++e) {
          const int32_t c = b[g->in_src[e]] + g->ins_score;
          if (c < b[v]) b[v] = c;
        }
      }
      std::swap(a, b);
    }
    std::memcpy(last_row, a.data(), sizeof(int32_t) * V);
  }
};

bool check_desc(const chsga_graph_desc* g) {
  return g && g->V > 0 && g->labels && g->in_off && g->str_off && g->rank && g->layer_off && g->cc_off &&
         g->up_off && g->dn_off &&
         (g->alphabet == CHSGA_ALPHABET_MATCH_N || g->alphabet == CHSGA_ALPHABET_MISMATCH_N);
}

int32_t row_min(const int32_t* r, int64_t V) {
  int32_t m = INT32_MAX;
  for (int64_t v = 0; v < V; ++v) m = std::min(m, r[v]);
  return m;
}

}  // namespace

extern "C" {

int oracle_threads(int n) {
  if (n > 0) omp_set_num_threads(n);
  return omp_get_max_threads();
}

int32_t oracle_sigma(int32_t alphabet, char a, char b, int32_t match, int32_t subst) {
  return sigma(alphabet, a, b, match, subst);
}

// align_combo: src/align.hpp:470-531 (per read: rows, then min over the last row)
int oracle_consga_scores(const chsga_graph_desc* g, const char* reads, const int64_t* off, int64_t n,
                         int32_t* scores) {
  if (!check_desc(g) || !reads || !off || !scores) return CHSGA_EINVAL;
  Model m(g);
  std::vector<int32_t> row(g->V);
  for (int64_t r = 0; r < n; ++r) {
    const int64_t len = off[r + 1] - off[r];
    if (len <= 0) return CHSGA_EINVAL;
    m.run(reads + off[r], len, row.data(), nullptr);
    scores[r] = row_min(row.data(), g->V);
  }
  return CHSGA_OK;
}

int oracle_consga_rows(const chsga_graph_desc* g, const char* read, int64_t len, int32_t* rows_out) {
  if (!check_desc(g) || !read || len <= 0 || !rows_out) return CHSGA_EINVAL;
  Model m(g);
  m.run(read, len, nullptr, rows_out);
  return CHSGA_OK;
}

int oracle_consga_last_row(const chsga_graph_desc* g, const char* read, int64_t len, int32_t* row_out) {
  if (!check_desc(g) || !read || len <= 0 || !row_out) return CHSGA_EINVAL;
  Model m(g);
  m.run(read, len, row_out, nullptr);
  return CHSGA_OK;
}

int oracle_exact_last_row(const chsga_graph_desc* g, const char* read, int64_t len, int32_t* row_out) {
  if (!check_desc(g) || !read || len <= 0 || !row_out) return CHSGA_EINVAL;
  Model m(g);
  const auto order = m.topo();
  if (int64_t(order.size()) != g->V) return CHSGA_EINVAL;  // not a DAG
  m.run_exact(read, len, row_out, order);
  return CHSGA_OK;
}

int oracle_cut_last_row(const chsga_graph_desc* g, const char* read, int64_t len, int32_t* row_out) {
  if (!check_desc(g) || !read || len <= 0 || !row_out) return CHSGA_EINVAL;
  Model m(g);
  const auto order = m.topo();
  if (int64_t(order.size()) != g->V) return CHSGA_EINVAL;
  m.run_cut(read, len, row_out, order);
  return CHSGA_OK;
}

int oracle_exact_scores(const chsga_graph_desc* g, const char* reads, const int64_t* off, int64_t n,
                        int32_t* scores) {
  if (!check_desc(g) || !reads || !off || !scores) return CHSGA_EINVAL;
  Model m(g);
  const auto order = m.topo();
  if (int64_t(order.size()) != g->V) return CHSGA_EINVAL;
  std::vector<int32_t> row(g->V);
  for (int64_t r = 0; r < n; ++r) {
    const int64_t len = off[r + 1] - off[r];
    if (len <= 0) return CHSGA_EINVAL;
    m.run_exact(reads + off[r], len, row.data(), order);
    scores[r] = row_min(row.data(), g->V);
  }
  return CHSGA_OK;
}

// walk_traceback_from_scores: src/traceback.hpp:483-586, with the best end
// vertex taken as the smallest index among the minima of the last row
// (src/traceback.hpp:695-714 breaks ties by worker order).
int64_t oracle_traceback(const chsga_graph_desc* g, const char* read, int64_t len, chsga_align_op* ops, int64_t cap,
                         int32_t* score_out) {
  if (!check_desc(g) || !read || len <= 0 || !ops || !g->str_out_off) return -2;
  const int64_t V = g->V;
  std::vector<int32_t> all(size_t(len) * size_t(V));
  Model m(g);
  m.run(read, len, nullptr, all.data());
  const int32_t* lastr = all.data() + (len - 1) * V;
  int64_t best = 0;
  for (int64_t v = 1; v < V; ++v)
    if (lastr[v] < lastr[best]) best = v;
  if (score_out) *score_out = lastr[best];
  // string in-CSR (build_in_csr_from, src/traceback.hpp:102-118) and vertex_of
  std::vector<uint64_t> sin_off(g->n_s + 1, 0);
  std::vector<uint32_t> sin_src(g->m_str);
  for (int64_t u = 0; u < g->n_s; ++u)
    for (uint64_t e = g->str_out_off[u]; e < g->str_out_off[u + 1]; ++e) ++sin_off[g->str_out_dst[e] + 1];
  for (int64_t s = 0; s < g->n_s; ++s) sin_off[s + 1] += sin_off[s];
  {
    std::vector<uint64_t> cur(sin_off.begin(), sin_off.end() - 1);
    for (int64_t u = 0; u < g->n_s; ++u)
      for (uint64_t e = g->str_out_off[u]; e < g->str_out_off[u + 1]; ++e)
        sin_src[cur[g->str_out_dst[e]]++] = uint32_t(u);
  }
  auto vertex_of = [&](int64_t p) -> int64_t {
    return int64_t(std::upper_bound(g->str_off, g->str_off + g->n_s, uint64_t(p)) - g->str_off) - 1;
  };
  std::vector<chsga_align_op> path;
  int64_t cur_v = best, cur_row = len - 1;
  while (cur_row > 0) {
    const int32_t* row = all.data() + cur_row * V;
    const int32_t* prow = row - V;
    const int32_t s = row[cur_v];
    int op = -1;
    int64_t src = cur_v;
    const int64_t sv = vertex_of(cur_v);
    if (cur_v > int64_t(g->str_off[sv]) && row[cur_v - 1] + g->ins_score == s) {
      op = CHSGA_OP_INSERT;
      src = cur_v - 1;
    }
    if (op < 0 && cur_v == int64_t(g->str_off[sv])) {
      for (uint64_t ie = sin_off[sv]; ie < sin_off[sv + 1]; ++ie) {
        const int64_t u_last = int64_t(g->str_off[sin_src[ie] + 1]) - 1;
        if (row[u_last] + g->ins_score == s) {
          op = CHSGA_OP_INSERT;
          src = u_last;
          break;
        }
      }
    }
    if (op < 0 && prow[cur_v] + g->del_score == s) {
      op = CHSGA_OP_DELETE;
      src = cur_v;
    }
    if (op < 0) {
      const int32_t delta = sigma(g->alphabet, g->labels[cur_v], read[cur_row], g->match_score, g->subst_score);
      for (uint64_t ie = g->in_off[cur_v]; ie < g->in_off[cur_v + 1]; ++ie)
        if (prow[g->in_src[ie]] + delta == s) {
          op = CHSGA_OP_MATCHSUB;
          src = g->in_src[ie];
          break;
        }
    }
    if (op < 0) return -1;
    path.push_back({cur_row, uint32_t(cur_v), uint8_t(op)});
    cur_v = src;
    if (op != CHSGA_OP_INSERT) --cur_row;
  }
  path.push_back({0, uint32_t(cur_v), CHSGA_OP_MATCHSUB});
  std::reverse(path.begin(), path.end());
  if (int64_t(path.size()) > cap) return -2;
  std::copy(path.begin(), path.end(), ops);
  return int64_t(path.size());
}

int oracle_validate_path(const chsga_graph_desc* g, const char* read, int64_t len, const chsga_align_op* ops,
                         int64_t n_ops, int32_t score, int64_t* bad_index) {
  auto fail = [&](int code, int64_t k) {
    if (bad_index) *bad_index = k;
    return code;
  };
  if (!check_desc(g) || n_ops <= 0 || !ops) return fail(1, 0);
  if (ops[0].row != 0 || ops[0].op != CHSGA_OP_MATCHSUB) return fail(2, 0);
  auto has_edge = [&](int64_t u, int64_t v) {
    for (uint64_t e = g->in_off[v]; e < g->in_off[v + 1]; ++e)
      if (int64_t(g->in_src[e]) == u) return true;
    return false;
  };
  int64_t cost = 0, last_consuming = 0;
  for (int64_t k = 0; k < n_ops; ++k) {
    const chsga_align_op& o = ops[k];
    if (int64_t(o.vertex) >= g->V || o.row < 0 || o.row >= len) return fail(4, k);
    if (k > 0) {
      const chsga_align_op& p = ops[k - 1];
      bool ok = false;
      if (o.op == CHSGA_OP_DELETE) ok = (o.row == p.row + 1) && (o.vertex == p.vertex);
      if (o.op == CHSGA_OP_MATCHSUB) ok = (o.row == p.row + 1) && has_edge(p.vertex, o.vertex);
      if (o.op == CHSGA_OP_INSERT) ok = (o.row == p.row) && has_edge(p.vertex, o.vertex);
      if (!ok) return fail(3, k);
    }
    // rescoring as in src/ConSGA_traceback.cpp:190-206
    if (o.op == CHSGA_OP_DELETE) {
      cost += g->del_score;
      last_consuming = o.row;
    } else if (o.op == CHSGA_OP_MATCHSUB) {
      cost += sigma(g->alphabet, g->labels[o.vertex], read[o.row], g->match_score, g->subst_score);
      last_consuming = o.row;
    } else {
      cost += g->ins_score;
    }
  }
  if (last_consuming != len - 1) return fail(5, n_ops - 1);
  if (cost != score) return fail(6, n_ops - 1);
  return 0;
}

}  // extern "C"
