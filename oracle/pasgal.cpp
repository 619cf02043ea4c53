// CPU restatement of PaSGAL_match's local alignment (the oracle, O5).  TEST
// INFRASTRUCTURE ONLY: loaded by tests/ and never by the product path.
//
// Follows the scalar build the reference's Makefile makes (SIMD_SUPPORT=none,
// Makefile:197-213), i.e. src/PaSGALoverrides/align.hpp with the scalar phase-1
// routines: alignToDAGLocal (:500-700) = phase 1 on both strands (:43-132,
// match among N), strand choice (:601-613), phase 1 in reverse (:146-245,
// exact match), phase 2 (:256-491, exact match: recompute in the window, walk,
// compact CIGAR).  The graph is the char DAG of the descriptor: char ids must
// be a topological order (every in-neighbour id smaller), as PaSGAL's loader
// guarantees for its CSR_char_container.
//
// Helpers PaSGAL keeps outside the override (its submodule is empty here),
// restated from their published behaviour: reverse complement (A<->T, C<->G,
// N->N, then reversed), reverse, and cigarCompact (runs of one op written as
// <count><op>).
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "oracle.h"

namespace {

struct PGraph {
  const chsga_graph_desc* g;
  std::vector<uint64_t> out_off;
  std::vector<uint32_t> out_dst;
  explicit PGraph(const chsga_graph_desc* gg) : g(gg), out_off(gg->V + 1, 0), out_dst(gg->E) {
    for (int64_t v = 0; v < g->V; ++v)
      for (uint64_t e = g->in_off[v]; e < g->in_off[v + 1]; ++e) ++out_off[g->in_src[e] + 1];
    for (int64_t v = 0; v < g->V; ++v) out_off[v + 1] += out_off[v];
    std::vector<uint64_t> cur(out_off.begin(), out_off.end() - 1);
    for (int64_t v = 0; v < g->V; ++v)
      for (uint64_t e = g->in_off[v]; e < g->in_off[v + 1]; ++e) out_dst[cur[g->in_src[e]]++] = uint32_t(v);
  }
};

struct Best {
  int32_t score = 0;
  int64_t row = 0, col = 0;
};

// alignToDAGLocal_Phase1_scalar (align.hpp:43-132): match among N
Best phase1(const PGraph& G, const std::string& r, const chsga_pasgal_params& P) {
  const chsga_graph_desc* g = G.g;
  const int64_t V = g->V, L = int64_t(r.size());
  std::vector<int32_t> m[2] = {std::vector<int32_t>(V, 0), std::vector<int32_t>(V, 0)};
  Best b;
  for (int64_t i = 0; i < L; ++i) {
    std::vector<int32_t>& cur = m[i & 1];
    const std::vector<int32_t>& up = m[(i - 1) & 1];
    for (int64_t j = 0; j < V; ++j) {
      const char c = g->labels[j];
      int32_t best = 0;
      const int32_t ms = (c == 'N' || r[i] == 'N' || c == r[i]) ? P.match : -P.mismatch;
      best = std::max(best, ms);
      for (uint64_t k = g->in_off[j]; k < g->in_off[j + 1]; ++k) {
        best = std::max(best, up[g->in_src[k]] + ms);
        best = std::max(best, cur[g->in_src[k]] - P.del);
      }
      best = std::max(best, up[j] - P.ins);
      cur[j] = best;
      b.score = std::max(b.score, best);
      if (b.score == best) {
        b.col = j;
        b.row = i;
      }
    }
  }
  return b;
}

// alignToDAGLocal_Phase1_rev_scalar (align.hpp:146-245): the reversed read
// right to left over the out-edges, exact match, the forward end cell set to
// match + 1 once computed.  Returns the start (row, col) and the best value.
Best phase1_rev(const PGraph& G, const std::string& rrev, const chsga_pasgal_params& P, int64_t row_end,
                int64_t col_end) {
  const chsga_graph_desc* g = G.g;
  const int64_t V = g->V, L = int64_t(rrev.size());
  std::vector<int32_t> m[2] = {std::vector<int32_t>(V, 0), std::vector<int32_t>(V, 0)};
  Best b;
  for (int64_t i = 0; i < L; ++i) {
    std::vector<int32_t>& cur = m[i & 1];
    const std::vector<int32_t>& up = m[(i - 1) & 1];
    for (int64_t j = V - 1; j >= 0; --j) {
      const char c = g->labels[j];
      int32_t best = 0;
      const int32_t ms = c == rrev[i] ? P.match : -P.mismatch;
      best = std::max(best, ms);
      for (uint64_t k = G.out_off[j]; k < G.out_off[j + 1]; ++k) {
        best = std::max(best, up[G.out_dst[k]] + ms);
        best = std::max(best, cur[G.out_dst[k]] - P.del);
      }
      best = std::max(best, up[j] - P.ins);
      cur[j] = best;
      b.score = std::max(b.score, best);
      if (b.score == best) {
        b.col = j;
        b.row = L - 1 - i;
      }
      if (j == col_end && (L - 1 - i) == row_end) cur[j] = P.match + 1;
    }
  }
  return b;
}

void cigar_compact(std::string& cigar) {
  if (cigar.empty()) return;
  std::string out;
  char prev = cigar[0];
  int64_t count = 1;
  for (size_t k = 1; k <= cigar.size(); ++k) {
    if (k < cigar.size() && cigar[k] == prev) {
      ++count;
      continue;
    }
    out += std::to_string(count);
    out.push_back(prev);
    if (k < cigar.size()) {
      prev = cigar[k];
      count = 1;
    }
  }
  cigar.swap(out);
}

// alignToDAGLocal_Phase2 (align.hpp:256-491) for one read: the DP recomputed
// in the window [i0, i0 + h) x [j0, j0 + w) with in-edges from below j0
// ignored, vertical differences kept as int8 (as the reference keeps them),
// then the walk from the end cell.  Returns false when the window is empty.
bool phase2(const PGraph& G, const std::string& r, const chsga_pasgal_params& P, int64_t i0, int64_t i1, int64_t j0,
            int64_t j1, std::string& cigar_out) {
  const chsga_graph_desc* g = G.g;
  if (j1 < j0 || i1 < i0) return false;
  const int64_t w = j1 - j0 + 1, h = i1 - i0 + 1;
  std::vector<std::vector<int8_t>> log(h, std::vector<int8_t>(w, 0));
  std::vector<int32_t> m[2] = {std::vector<int32_t>(w, 0), std::vector<int32_t>(w, 0)};
  std::vector<int32_t> finalRow(w, 0);
  for (int64_t i = 0; i < h; ++i) {
    std::vector<int32_t>& cur = m[i & 1];
    const std::vector<int32_t>& up = m[(i - 1) & 1];
    for (int64_t j = 0; j < w; ++j) {
      const char c = g->labels[j + j0];
      const int32_t fromIns = up[j] - P.ins;
      const int32_t ms = c == r[i + i0] ? P.match : -P.mismatch;
      int32_t fromMatch = ms, fromDel = -1;
      for (uint64_t k = g->in_off[j + j0]; k < g->in_off[j + j0 + 1]; ++k) {
        if (int64_t(g->in_src[k]) >= j0) {
          fromMatch = std::max(fromMatch, up[g->in_src[k] - j0] + ms);
          fromDel = std::max(fromDel, cur[g->in_src[k] - j0] - P.del);
        }
      }
      cur[j] = std::max(std::max(fromIns, fromMatch), std::max(fromDel, 0));
      log[i][j] = int8_t(cur[j] - up[j]);
    }
    if (i == h - 1) finalRow = cur;
  }
  std::string cigar;
  std::vector<int32_t> curRow = finalRow, above(w);
  int64_t col = w - 1, row = h - 1;
  while (col >= 0 && row >= 0) {
    if (curRow[col] <= 0) break;
    for (int64_t k = 0; k < w; ++k) above[k] = curRow[k] - log[row][k];
    const char c = g->labels[col + j0];
    const int32_t fromIns = above[col] - P.ins;
    const int32_t ms = c == r[row + i0] ? P.match : -P.mismatch;
    int32_t fromMatch = ms;
    int64_t fromMatchPos = col;
    int32_t fromDel = -1;
    int64_t fromDelPos = -1;
    for (uint64_t k = g->in_off[col + j0]; k < g->in_off[col + j0 + 1]; ++k) {
      if (int64_t(g->in_src[k]) >= j0) {
        const int64_t fc = int64_t(g->in_src[k]) - j0;
        if (fromMatch < above[fc] + ms) {
          fromMatch = above[fc] + ms;
          fromMatchPos = fc;
        }
        if (fromDel < curRow[fc] - P.del) {
          fromDel = curRow[fc] - P.del;
          fromDelPos = fc;
        }
      }
    }
    if (curRow[col] == fromMatch) {
      cigar.push_back(ms == P.match ? '=' : 'X');
      if (fromMatchPos == col) break;
      col = fromMatchPos;
      --row;
      curRow = above;
    } else if (curRow[col] == fromDel) {
      cigar.push_back('D');
      col = fromDelPos;
    } else {
      cigar.push_back('I');  // curRow[col] == fromIns (align.hpp:453)
      --row;
      curRow = above;
    }
    (void)fromIns;
  }
  std::reverse(cigar.begin(), cigar.end());
  cigar_compact(cigar);
  cigar_out = cigar;
  return true;
}

bool revcomp(const std::string& in, std::string& out) {
  out.resize(in.size());
  for (size_t k = 0; k < in.size(); ++k) {
    char c;
    switch (in[k]) {
      case 'A': c = 'T'; break;
      case 'C': c = 'G'; break;
      case 'G': c = 'C'; break;
      case 'T': c = 'A'; break;
      case 'N': c = 'N'; break;
      default: return false;
    }
    out[in.size() - 1 - k] = c;
  }
  return true;
}

}  // namespace

extern "C" {

int oracle_pasgal_align(const chsga_graph_desc* g, const char* reads, const int64_t* off, int64_t n,
                        const chsga_pasgal_params* P, chsga_pasgal_result* out) {
  if (!g || !reads || !off || n < 0 || !P || !out || g->V <= 0) return CHSGA_EINVAL;
  for (int64_t v = 0; v < g->V; ++v)
    for (uint64_t e = g->in_off[v]; e < g->in_off[v + 1]; ++e)
      if (int64_t(g->in_src[e]) >= v) return CHSGA_EINVAL;  // not a topological numbering
  const PGraph G(g);
  int bad = 0;
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t q = 0; q < n; ++q) {
    chsga_pasgal_result& res = out[q];
    std::memset(&res, 0, sizeof(res));
    const std::string fwd(reads + off[q], size_t(off[q + 1] - off[q]));
    std::string rc;
    if (fwd.empty() || !revcomp(fwd, rc)) {
#pragma omp atomic write
      bad = 1;
      continue;
    }
    const Best bf = phase1(G, fwd, *P), br = phase1(G, rc, *P);
    res.score_fwd = bf.score;
    res.qry_end_fwd = bf.row;
    res.ref_end_fwd = bf.col;
    res.score_rc = br.score;
    res.qry_end_rc = br.row;
    res.ref_end_rc = br.col;
    const bool plus = bf.score > br.score;  // align.hpp:601: ties give '-'
    const Best& b = plus ? bf : br;
    const std::string& r = plus ? fwd : rc;
    res.strand = plus ? '+' : '-';
    res.score = b.score;
    res.qry_end = b.row;
    res.ref_end = b.col;
    res.qry_start = -1;
    res.ref_start = -1;
    if (P->score_only) continue;
    const std::string rrev(r.rbegin(), r.rend());
    const Best bs = phase1_rev(G, rrev, *P, b.row, b.col);
    res.qry_start = bs.row;
    res.ref_start = bs.col;
    res.rev_best = bs.score;
    std::string cigar;
    if (!phase2(G, r, *P, bs.row, b.row, bs.col, b.col, cigar)) {
      res.flags |= CHSGA_PASGAL_EMPTY_WINDOW;
      continue;
    }
    res.cigar = static_cast<char*>(std::malloc(cigar.size() + 1));
    std::memcpy(res.cigar, cigar.c_str(), cigar.size() + 1);
  }
  return bad ? CHSGA_EINVAL : CHSGA_OK;
}

void oracle_pasgal_free(chsga_pasgal_result* res, int64_t n) {
  for (int64_t q = 0; q < n; ++q) {
    std::free(res[q].cigar);
    res[q].cigar = nullptr;
  }
}

}  // extern "C"
