// Vertex-weighted contraction hierarchy of the string graph.
//
// Follows the semantics ConSGA relies on (src/CHoverrides/build_pch.hpp):
//  - contraction in rounds; each round contracts an independent set of local
//    minima of an edge-difference score (build_pch.hpp:396-478, 536-571, 953-955);
//  - contracting u adds, for every in-neighbour x and out-neighbour y (x != y),
//    the shortcut x -> y of weight w(x,u) + w(u,y) + len(u) - 1 unless a
//    strictly shorter witness path exists (build_pch.hpp:1022-1037, 590-620);
//  - u keeps its remaining out-edges as up-edges and its remaining in-edges as
//    down-edges, and every such neighbour gets level >= level(u) + 1
//    (build_pch.hpp:1039-1072); levels start at 1 (build_pch.hpp:141);
//  - vertices are then ordered by (connected component, level, id), which
//    gives rank, the layer offsets and the component offsets
//    (build_pch.hpp:1132-1220), and edges are renumbered and sorted by target
//    (build_pch.hpp:1168-1197).
// The score, the witness-search limits and the priority tie-break are this
// builder's own (the reference draws random priorities, build_pch.hpp:154).
// Each shortcut also records its middle vertex so that traceback can expand it.
#include <algorithm>
#include <cstring>
#include <numeric>

#include "host_internal.hpp"

namespace chsga {
namespace {

struct Dyn {
  uint32_t v;
  int32_t w;
  int32_t mid;
};

constexpr int kWitnessSettle = 48;
constexpr int kTable = 512;

struct SmallMap {
  uint32_t key[kTable];
  int32_t val[kTable];
  int used = 0;
  void clear() {
    std::memset(key, 0xff, sizeof(key));
    used = 0;
  }
  static uint32_t h(uint32_t k) { return (k * 2654435761u) & (kTable - 1); }
  // returns pointer to value slot, or nullptr when full
  int32_t* find_or_insert(uint32_t k, bool& inserted) {
    uint32_t i = h(k);
    for (int probe = 0; probe < kTable; ++probe) {
      if (key[i] == k) {
        inserted = false;
        return &val[i];
      }
      if (key[i] == 0xffffffffu) {
        if (used >= kTable * 3 / 4) return nullptr;
        key[i] = k;
        ++used;
        inserted = true;
        return &val[i];
      }
      i = (i + 1) & (kTable - 1);
    }
    return nullptr;
  }
  const int32_t* find(uint32_t k) const {
    uint32_t i = h(k);
    for (int probe = 0; probe < kTable; ++probe) {
      if (key[i] == k) return &val[i];
      if (key[i] == 0xffffffffu) return nullptr;
      i = (i + 1) & (kTable - 1);
    }
    return nullptr;
  }
};

struct Builder {
  int64_t n;
  std::vector<int32_t> len;
  std::vector<std::vector<Dyn>> out, in;
  std::vector<uint8_t> dead, sel, dirty;
  std::vector<uint32_t> level;
  std::vector<int64_t> score;
  std::vector<uint32_t> prio;
  std::vector<uint64_t> up_start, dn_start;
  std::vector<uint32_t> up_cnt, dn_cnt;
  std::vector<ChEdge> rec_up, rec_dn;

  // Bounded Dijkstra from x avoiding `skip` and every vertex selected in this
  // round.  Fills dist[j] for the targets tg[j] (INT32_MAX when not reached).
  void witness(uint32_t x, uint32_t skip, const std::vector<Dyn>& tg, int32_t bound, SmallMap& m,
               std::vector<std::pair<int32_t, uint32_t>>& heap, int32_t* dist) const {
    m.clear();
    heap.clear();
    bool ins;
    int32_t* d0 = m.find_or_insert(x, ins);
    *d0 = 0;
    heap.push_back({0, x});
    int settled = 0;
    auto cmp = [](const std::pair<int32_t, uint32_t>& a, const std::pair<int32_t, uint32_t>& b) {
      return a.first > b.first;
    };
    while (!heap.empty() && settled < kWitnessSettle) {
      std::pop_heap(heap.begin(), heap.end(), cmp);
      auto [d, a] = heap.back();
      heap.pop_back();
      const int32_t* da = m.find(a);
      if (!da || *da < d) continue;
      if (d > bound) break;
      ++settled;
      const int32_t through = (a == x) ? 0 : (len[a] - 1);
      for (const Dyn& e : out[a]) {
        const uint32_t b = e.v;
        if (b == skip || sel[b] || b == x) continue;
        const int32_t nd = d + through + e.w;
        if (nd > bound) continue;
        int32_t* db = m.find_or_insert(b, ins);
        if (!db) {
          settled = kWitnessSettle;  // table full: stop, conservatively
          break;
        }
        if (ins || nd < *db) {
          *db = nd;
          heap.push_back({nd, b});
          std::push_heap(heap.begin(), heap.end(), cmp);
        }
      }
    }
    for (size_t j = 0; j < tg.size(); ++j) {
      const int32_t* dv = m.find(tg[j].v);
      dist[j] = dv ? *dv : INT32_MAX;
    }
  }

  // Number of shortcuts contracting u would add.
  int64_t count_shortcuts(uint32_t u, SmallMap& m, std::vector<std::pair<int32_t, uint32_t>>& heap,
                          std::vector<int32_t>& dist) const {
    int64_t added = 0;
    dist.resize(out[u].size());
    for (const Dyn& x : in[u]) {
      int32_t bound = 0;
      for (const Dyn& y : out[u]) bound = std::max(bound, x.w + len[u] - 1 + y.w);
      witness(x.v, u, out[u], bound, m, heap, dist.data());
      for (size_t j = 0; j < out[u].size(); ++j) {
        const Dyn& y = out[u][j];
        if (y.v == x.v) continue;
        const int32_t w = x.w + len[u] - 1 + y.w;
        if (dist[j] >= w) ++added;
      }
    }
    return added;
  }

  int64_t compute_score(uint32_t u, SmallMap& m, std::vector<std::pair<int32_t, uint32_t>>& heap,
                        std::vector<int32_t>& dist) const {
    const int64_t din = int64_t(in[u].size()), dout = int64_t(out[u].size());
    if (din == 0 && dout == 0) return 0;
    const int64_t added = count_shortcuts(u, m, heap, dist);
    const int64_t removed = 1 + din + dout;
    return 1 + 500 * int64_t(level[u]) + (1000 * added) / removed + 10 * removed;
  }

  bool less_key(uint32_t a, uint32_t b) const {
    if (score[a] != score[b]) return score[a] < score[b];
    const size_t da = in[a].size() + out[a].size(), db = in[b].size() + out[b].size();
    if (da != db) return da < db;
    if (prio[a] != prio[b]) return prio[a] < prio[b];
    return a < b;
  }

  static void merge_edge(std::vector<Dyn>& lst, uint32_t v, int32_t w, int32_t mid) {
    for (Dyn& e : lst)
      if (e.v == v) {
        if (w < e.w) {
          e.w = w;
          e.mid = mid;
        }
        return;
      }
    lst.push_back({v, w, mid});
  }
  static void erase_edge(std::vector<Dyn>& lst, uint32_t v) {
    for (size_t i = 0; i < lst.size(); ++i)
      if (lst[i].v == v) {
        lst[i] = lst.back();
        lst.pop_back();
        return;
      }
  }

  void run(uint64_t seed) {
    dead.assign(n, 0);
    sel.assign(n, 0);
    dirty.assign(n, 1);
    level.assign(n, 1);
    score.assign(n, 0);
    prio.resize(n);
    Rng rng(seed ^ 0x5eedc4a11ull);
    for (int64_t i = 0; i < n; ++i) prio[i] = uint32_t(rng.next() >> 32);
    up_start.assign(n, 0);
    dn_start.assign(n, 0);
    up_cnt.assign(n, 0);
    dn_cnt.assign(n, 0);

    std::vector<uint32_t> alive(n);
    std::iota(alive.begin(), alive.end(), 0u);
    std::vector<uint32_t> chosen;
    struct Sc {
      uint32_t x, y;
      int32_t w;
      uint32_t mid;
    };
    std::vector<std::vector<Sc>> shortcuts;

    while (!alive.empty()) {
      // 1. scores of vertices whose neighbourhood changed
#pragma omp parallel
      {
        SmallMap m;
        std::vector<std::pair<int32_t, uint32_t>> heap;
        std::vector<int32_t> dist;
#pragma omp for schedule(dynamic, 1024)
        for (int64_t i = 0; i < int64_t(alive.size()); ++i) {
          const uint32_t u = alive[i];
          if (dirty[u]) {
            score[u] = compute_score(u, m, heap, dist);
            dirty[u] = 0;
          }
        }
      }
      // 2. independent set of local minima
      chosen.clear();
      for (uint32_t u : alive) {
        bool ok = true;
        for (const Dyn& e : out[u])
          if (!less_key(u, e.v)) {
            ok = false;
            break;
          }
        if (ok)
          for (const Dyn& e : in[u])
            if (!less_key(u, e.v)) {
              ok = false;
              break;
            }
        if (ok) chosen.push_back(u);
      }
      for (uint32_t u : chosen) sel[u] = 1;
      // 3a. shortcuts of every chosen vertex, against the graph before this round
      shortcuts.assign(chosen.size(), {});
#pragma omp parallel
      {
        SmallMap m;
        std::vector<std::pair<int32_t, uint32_t>> heap;
        std::vector<int32_t> dist;
#pragma omp for schedule(dynamic, 256)
        for (int64_t i = 0; i < int64_t(chosen.size()); ++i) {
          const uint32_t u = chosen[i];
          dist.resize(out[u].size());
          for (const Dyn& x : in[u]) {
            int32_t bound = 0;
            for (const Dyn& y : out[u]) bound = std::max(bound, x.w + len[u] - 1 + y.w);
            witness(x.v, u, out[u], bound, m, heap, dist.data());
            for (size_t j = 0; j < out[u].size(); ++j) {
              const Dyn& y = out[u][j];
              if (y.v == x.v) continue;
              const int32_t w = x.w + len[u] - 1 + y.w;
              if (dist[j] >= w) shortcuts[i].push_back({x.v, y.v, w, u});
            }
          }
        }
      }
      // 3b. contract in ascending id order
      for (size_t i = 0; i < chosen.size(); ++i) {
        const uint32_t u = chosen[i];
        up_start[u] = rec_up.size();
        up_cnt[u] = uint32_t(out[u].size());
        for (const Dyn& e : out[u]) rec_up.push_back({e.v, e.w, e.mid});
        dn_start[u] = rec_dn.size();
        dn_cnt[u] = uint32_t(in[u].size());
        for (const Dyn& e : in[u]) rec_dn.push_back({e.v, e.w, e.mid});
        for (const Dyn& e : out[u]) {
          level[e.v] = std::max(level[e.v], level[u] + 1);
          erase_edge(in[e.v], u);
          dirty[e.v] = 1;
        }
        for (const Dyn& e : in[u]) {
          level[e.v] = std::max(level[e.v], level[u] + 1);
          erase_edge(out[e.v], u);
          dirty[e.v] = 1;
        }
        dead[u] = 1;
      }
      for (size_t i = 0; i < chosen.size(); ++i)
        for (const Sc& s : shortcuts[i]) {
          merge_edge(out[s.x], s.y, s.w, int32_t(s.mid));
          merge_edge(in[s.y], s.x, s.w, int32_t(s.mid));
        }
      for (size_t i = 0; i < chosen.size(); ++i) {
        const uint32_t u = chosen[i];
        sel[u] = 0;
        std::vector<Dyn>().swap(out[u]);
        std::vector<Dyn>().swap(in[u]);
      }
      size_t k = 0;
      for (uint32_t u : alive)
        if (!dead[u]) alive[k++] = u;
      alive.resize(k);
    }
  }
};

struct Dsu {
  std::vector<uint32_t> p;
  explicit Dsu(int64_t n) : p(n) { std::iota(p.begin(), p.end(), 0u); }
  uint32_t find(uint32_t x) {
    while (p[x] != x) {
      p[x] = p[p[x]];
      x = p[x];
    }
    return x;
  }
  void unite(uint32_t a, uint32_t b) {
    a = find(a);
    b = find(b);
    if (a == b) return;
    if (a < b)
      p[b] = a;
    else
      p[a] = b;
  }
};

}  // namespace

void build_ch(chsga_dataset& ds, uint64_t seed) {
  const int64_t n = ds.n_s;
  Builder b;
  b.n = n;
  b.len.resize(n);
  for (int64_t s = 0; s < n; ++s) b.len[s] = int32_t(ds.str_off[s + 1] - ds.str_off[s]);
  b.out.assign(n, {});
  b.in.assign(n, {});
  for (int64_t u = 0; u < n; ++u)
    for (uint64_t e = ds.str_out_off[u]; e < ds.str_out_off[u + 1]; ++e) {
      const uint32_t v = ds.str_out_dst[e];
      if (int64_t(v) == u) continue;  // self loops are dropped (build_pch.hpp:193-195)
      Builder::merge_edge(b.out[u], v, 1, -1);
      Builder::merge_edge(b.in[v], uint32_t(u), 1, -1);
    }
  b.run(seed);

  // components of the (undirected) string graph
  Dsu dsu(n);
  for (int64_t u = 0; u < n; ++u)
    for (uint64_t e = ds.str_out_off[u]; e < ds.str_out_off[u + 1]; ++e) dsu.unite(uint32_t(u), ds.str_out_dst[e]);
  std::vector<uint32_t> cc(n);
  std::vector<uint32_t> cc_of_root(n, UINT32_MAX);
  uint32_t ncc = 0;
  for (int64_t u = 0; u < n; ++u) {
    const uint32_t r = dsu.find(uint32_t(u));
    if (cc_of_root[r] == UINT32_MAX) cc_of_root[r] = ncc++;
    cc[u] = cc_of_root[r];
  }
  std::vector<uint32_t> order(n);
  std::iota(order.begin(), order.end(), 0u);
  std::sort(order.begin(), order.end(), [&](uint32_t a, uint32_t c) {
    if (cc[a] != cc[c]) return cc[a] < cc[c];
    if (b.level[a] != b.level[c]) return b.level[a] < b.level[c];
    return a < c;
  });
  ContractionHierarchy& ch = ds.ch;
  ch = ContractionHierarchy();
  ch.n = n;
  ch.rank.resize(n);
  ch.level.resize(n);
  for (int64_t i = 0; i < n; ++i) {
    ch.rank[order[i]] = uint32_t(i);
    ch.level[i] = b.level[order[i]];
  }
  ch.layer_off.push_back(0);
  ch.cc_off.push_back(0);
  for (int64_t i = 1; i <= n; ++i) {
    const bool end = (i == n) || cc[order[i]] != cc[order[i - 1]] || b.level[order[i]] != b.level[order[i - 1]];
    if (end) {
      ch.layer_off.push_back(uint64_t(i));
      if (i == n || cc[order[i]] != cc[order[i - 1]]) ch.cc_off.push_back(ch.layer_off.size() - 1);
    }
  }
  if (n == 0) ch.layer_off.assign(1, 0);
  auto emit = [&](const std::vector<uint64_t>& start, const std::vector<uint32_t>& cnt, const std::vector<ChEdge>& rec,
                  std::vector<uint64_t>& off, std::vector<uint32_t>& vv, std::vector<int32_t>& ww,
                  std::vector<int32_t>& mm) {
    off.assign(n + 1, 0);
    std::vector<ChEdge> tmp;
    for (int64_t c = 0; c < n; ++c) {
      const uint32_t og = order[c];
      tmp.clear();
      for (uint32_t j = 0; j < cnt[og]; ++j) {
        const ChEdge& e = rec[start[og] + j];
        tmp.push_back({ch.rank[e.v], e.w, e.mid >= 0 ? int32_t(ch.rank[uint32_t(e.mid)]) : -1});
      }
      std::sort(tmp.begin(), tmp.end(), [](const ChEdge& a, const ChEdge& c2) { return a.v < c2.v; });
      for (const ChEdge& e : tmp) {
        vv.push_back(e.v);
        ww.push_back(e.w);
        mm.push_back(e.mid);
      }
      off[c + 1] = vv.size();
    }
  };
  emit(b.up_start, b.up_cnt, b.rec_up, ch.up_off, ch.up_v, ch.up_w, ch.up_mid);
  emit(b.dn_start, b.dn_cnt, b.rec_dn, ch.dn_off, ch.dn_v, ch.dn_w, ch.dn_mid);
  ch.has_mid = true;
}

}  // namespace chsga
