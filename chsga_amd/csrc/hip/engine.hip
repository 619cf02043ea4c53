// MI355X engine behind include/chsga/abi.h: context set-up, the plain
// one-read-per-launch-sequence path, the batched path and traceback.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstring>
#include <exception>
#include <mutex>
#include <numeric>
#include <unordered_map>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "chsga/abi.h"
#include "engine_internal.hpp"

namespace chsga_engine {

thread_local std::string g_err;
void set_error(const std::string& m) { g_err = m; }

void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw HipError(std::string(what) + ": " + hipGetErrorString(e));
}

void HostMem::alloc(size_t bytes) {
  release();
  if (bytes == 0) return;
  hip_check(hipHostMalloc(&p, bytes, hipHostMallocDefault), "hipHostMalloc");
  n = bytes;
}
void HostMem::release() {
  if (p) (void)hipHostFree(p);
  p = nullptr;
  n = 0;
}
void DevMem::alloc(size_t bytes) {
  release();
  if (bytes == 0) return;
  hip_check(hipMalloc(&p, bytes), "hipMalloc");
  n = bytes;
}
void DevMem::release() {
  if (p) (void)hipFree(p);
  p = nullptr;
  n = 0;
}

template <class T>
static void upload(DevMem& d, const std::vector<T>& h, hipStream_t s) {
  const size_t bytes = std::max<size_t>(1, h.size()) * sizeof(T);
  if (d.n != bytes) d.alloc(bytes);  // same size: reuse the allocation
  if (h.empty()) return;
  hip_check(hipMemcpyAsync(d.p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, s), "upload");
  // the source is pageable and may be a temporary: wait until it has been read
  hip_check(hipStreamSynchronize(s), "upload");
}

static int64_t now_ns() {
  return std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

}  // namespace chsga_engine

using namespace chsga_engine;
using namespace chsga_dev;

#include "cut.hip.inc"

// ---------------------------------------------------------------------------
// context set-up
// ---------------------------------------------------------------------------
// Pull lists of the top set (sweep positions K0 .. n_s-1): up-edges from
// below the set are external (sources are final when the top kernel starts),
// the rest are internal and addressed by local index.
static void build_top(chsga_ctx* c, TopSet& T, int32_t j_top, const std::vector<uint32_t>& up_off,
                      const std::vector<int2>& up_in, const std::vector<uint32_t>& dn_off,
                      const std::vector<int2>& dn_in, const std::vector<uint32_t>& pos_of_og) {
  T.j_top = j_top;
  T.K0 = j_top < c->J ? c->h_lvl_off[j_top] : uint32_t(c->n_s);
  T.NT = uint32_t(c->n_s) - T.K0;
  std::vector<uint32_t> ext_off{0}, upi_off{0}, dni_off{0};
  std::vector<int2> ext, upi, dni;
  auto local = [&](int2 r) {
    const uint32_t og = uint32_t(r.x) & 0x7fffffffu;
    const uint32_t pos = pos_of_og[og];
    return std::make_pair(pos >= T.K0, make_int2(int32_t((pos - T.K0) | (uint32_t(r.x) & 0x80000000u)), r.y));
  };
  for (int64_t k = T.K0; k < c->n_s; ++k) {
    for (uint32_t e = up_off[k]; e < up_off[k + 1]; ++e) {
      const auto lr = local(up_in[e]);
      if (lr.first)
        upi.push_back(lr.second);
      else
        ext.push_back(up_in[e]);
    }
    for (uint32_t e = dn_off[k]; e < dn_off[k + 1]; ++e) {
      const auto lr = local(dn_in[e]);
      if (!lr.first) throw std::invalid_argument("CH down-edge source below the top set");
      dni.push_back(lr.second);
    }
    ext_off.push_back(uint32_t(ext.size()));
    upi_off.push_back(uint32_t(upi.size()));
    dni_off.push_back(uint32_t(dni.size()));
  }
  upload(T.ext_off, ext_off, c->stream);
  upload(T.ext, ext, c->stream);
  upload(T.upi_off, upi_off, c->stream);
  upload(T.upi, upi, c->stream);
  upload(T.dni_off, dni_off, c->stream);
  upload(T.dni, dni, c->stream);
  T.n_upi = uint32_t(upi.size());
  T.n_dni = uint32_t(dni.size());
}

// Scratch of chsga_score_reads_noalloc: CHSGA_NOALLOC_MAX_READS argmin keys
// followed by as many int32 scores (1.5 * 65536 * 8 bytes = 768 KiB), held by
// every GPU the library runs on until the process ends.  Host code never
// passes the array's name to a kernel: each device's own copy is reached
// through the address hipGetSymbolAddress gives on that device.
static __device__ __align__(256) unsigned long long scratch[CHSGA_NOALLOC_MAX_READS + CHSGA_NOALLOC_MAX_READS / 2];

// Address of `scratch` on the current device, looked up once per device.
static void* noalloc_scratch(int device) {
  static std::mutex mu;
  static std::unordered_map<int, void*> addr;
  const std::lock_guard<std::mutex> lock(mu);
  auto it = addr.find(device);
  if (it != addr.end()) return it->second;
  void* p = nullptr;
  hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
  addr[device] = p;
  return p;
}

static void build_context(chsga_ctx* c, const chsga_graph_desc* g) {
  if (!g || g->V <= 0 || !g->labels || !g->in_off || !g->in_src || g->n_s <= 0 || !g->str_off || !g->rank ||
      !g->layer_off || !g->cc_off || !g->up_off || !g->dn_off)
    throw std::invalid_argument("incomplete graph descriptor");
  if (g->alphabet != CHSGA_ALPHABET_MATCH_N && g->alphabet != CHSGA_ALPHABET_MISMATCH_N)
    throw std::invalid_argument("alphabet must be 4 (matchN) or 5 (mismatchN)");
  if (g->match_score < 0 || g->subst_score < 0 || g->del_score < 0 || g->ins_score < 0)
    throw std::invalid_argument("edit costs must be non-negative");
  if (g->V >= (int64_t(1) << 31) || g->E >= (int64_t(1) << 32) || g->n_s >= (int64_t(1) << 31))
    throw std::invalid_argument("graph too large for 32-bit indices");
  if (int64_t(g->str_off[g->n_s]) != g->V || g->str_off[0] != 0)
    throw std::invalid_argument("label offsets do not cover the char vertices");
  if (int64_t(g->in_off[g->V]) != g->E) throw std::invalid_argument("char in-CSR size mismatch");
  c->V = g->V;
  c->E = g->E;
  c->n_s = g->n_s;
  c->alphabet = g->alphabet;
  c->msc = g->match_score;
  c->ssc = g->subst_score;
  c->dsc = g->del_score;
  c->isc = g->ins_score;
  const int64_t V = c->V, n_s = c->n_s;

  // host copies used by traceback
  c->h_labels.assign(g->labels, g->labels + V);
  c->h_str_off.resize(n_s + 1);
  for (int64_t s = 0; s <= n_s; ++s) c->h_str_off[s] = uint32_t(g->str_off[s]);
  c->h_in_off.resize(V + 1);
  for (int64_t v = 0; v <= V; ++v) c->h_in_off[v] = uint32_t(g->in_off[v]);
  c->h_in_src.assign(g->in_src, g->in_src + c->E);
  if (g->str_out_off && g->str_out_dst) {
    c->h_sin_off.assign(n_s + 1, 0);
    c->h_sin_src.resize(g->m_str);
    for (int64_t u = 0; u < n_s; ++u)
      for (uint64_t e = g->str_out_off[u]; e < g->str_out_off[u + 1]; ++e) ++c->h_sin_off[g->str_out_dst[e] + 1];
    for (int64_t s = 0; s < n_s; ++s) c->h_sin_off[s + 1] += c->h_sin_off[s];
    std::vector<uint32_t> cur(c->h_sin_off.begin(), c->h_sin_off.end() - 1);
    for (int64_t u = 0; u < n_s; ++u)
      for (uint64_t e = g->str_out_off[u]; e < g->str_out_off[u + 1]; ++e)
        c->h_sin_src[cur[g->str_out_dst[e]]++] = uint32_t(u);
  }

  // compact mode: every char that is not the first of its string has exactly
  // the in-neighbour p - 1 (the expansion rule of the char graph)
  c->compact = true;
  {
    size_t si = 0;
    for (int64_t p = 0; p < V && c->compact; ++p) {
      while (si + 1 < c->h_str_off.size() && c->h_str_off[si + 1] <= uint32_t(p)) ++si;
      if (int64_t(c->h_str_off[si]) == p) continue;
      c->compact = (c->h_in_off[p + 1] - c->h_in_off[p] == 1) && (c->h_in_src[c->h_in_off[p]] == uint32_t(p - 1));
    }
  }
  // two words of padding: kernels read the word after the one holding bit p
  std::vector<uint32_t> first_bits((V + 32) / 32 + 2, 0);
  for (int64_t s = 0; s < n_s; ++s) first_bits[c->h_str_off[s] >> 5] |= 1u << (c->h_str_off[s] & 31);

  // tiles of whole string vertices
  const int64_t T = c->tile_chars;
  std::vector<uint32_t> tile_char{0}, tile_str{0};
  int64_t acc = 0;
  for (int64_t s = 0; s < n_s; ++s) {
    const int64_t len = c->h_str_off[s + 1] - c->h_str_off[s];
    if (acc > 0 && acc + len > T) {
      tile_char.push_back(c->h_str_off[s]);
      tile_str.push_back(uint32_t(s));
      acc = 0;
    }
    acc += len;
  }
  tile_char.push_back(uint32_t(V));
  tile_str.push_back(uint32_t(n_s));
  c->n_tiles = int64_t(tile_char.size()) - 1;

  // CH sweep in string-id space.  step(c) = layer index of c inside its
  // component; component layers are processed side by side, which is exact
  // because components share no edges.
  std::vector<uint32_t> ch_to_og(n_s);
  for (int64_t i = 0; i < n_s; ++i) ch_to_og[g->rank[i]] = uint32_t(i);
  std::vector<int32_t> step(n_s, -1);
  int32_t J = 0;
  for (int64_t cc = 0; cc < g->num_cc; ++cc)
    for (uint64_t l = g->cc_off[cc]; l < g->cc_off[cc + 1]; ++l) {
      const int32_t j = int32_t(l - g->cc_off[cc]);
      J = std::max(J, j + 1);
      for (uint64_t x = g->layer_off[l]; x < g->layer_off[l + 1]; ++x) step[x] = j;
    }
  for (int64_t x = 0; x < n_s; ++x)
    if (step[x] < 0) throw std::invalid_argument("CH layers do not cover every vertex");
  std::vector<uint32_t> order(n_s);  // CH ids in sweep order
  std::iota(order.begin(), order.end(), 0u);
  std::sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
    if (step[a] != step[b]) return step[a] < step[b];
    return ch_to_og[a] < ch_to_og[b];
  });
  std::vector<uint32_t> kpos(n_s);
  for (int64_t k = 0; k < n_s; ++k) kpos[order[k]] = uint32_t(k);
  c->h_lvl_off.assign(J + 1, 0);
  for (int64_t x = 0; x < n_s; ++x) ++c->h_lvl_off[step[x] + 1];
  for (int32_t j = 0; j < J; ++j) c->h_lvl_off[j + 1] += c->h_lvl_off[j];
  c->J = J;
  std::vector<uint32_t> sw_vert(n_s);
  for (int64_t k = 0; k < n_s; ++k) sw_vert[k] = ch_to_og[order[k]];
  auto code = [&](uint32_t og) -> int32_t {
    const bool multi = (c->h_str_off[og + 1] - c->h_str_off[og]) > 1;
    return int32_t(og | (multi ? 0x80000000u : 0u));
  };
  // up pulls: transpose of the up-edges (x -> y stored at x)
  std::vector<uint32_t> up_off(n_s + 1, 0);
  for (int64_t x = 0; x < n_s; ++x)
    for (uint64_t e = g->up_off[x]; e < g->up_off[x + 1]; ++e) {
      const uint32_t y = g->up_v[e];
      if (step[y] <= step[x]) throw std::invalid_argument("CH up-edge does not go to a higher layer");
      ++up_off[kpos[y] + 1];
    }
  for (int64_t k = 0; k < n_s; ++k) up_off[k + 1] += up_off[k];
  std::vector<int2> up_in(up_off[n_s]);
  {
    std::vector<uint32_t> cur(up_off.begin(), up_off.end() - 1);
    for (int64_t x = 0; x < n_s; ++x)
      for (uint64_t e = g->up_off[x]; e < g->up_off[x + 1]; ++e) {
        const uint32_t y = g->up_v[e];
        up_in[cur[kpos[y]]++] = make_int2(code(ch_to_og[x]), g->up_w[e]);
      }
  }
  std::vector<uint32_t> dn_off(n_s + 1, 0);
  std::vector<int2> dn_in;
  dn_in.reserve(g->m_dn);
  for (int64_t k = 0; k < n_s; ++k) {
    const uint32_t x = order[k];
    for (uint64_t e = g->dn_off[x]; e < g->dn_off[x + 1]; ++e) {
      const uint32_t y = g->dn_v[e];
      if (step[y] <= step[x]) throw std::invalid_argument("CH down-edge does not come from a higher layer");
      dn_in.push_back(make_int2(code(ch_to_og[y]), g->dn_w[e]));
    }
    dn_off[k + 1] = uint32_t(dn_in.size());
  }
  // exact insertion closure (chsga_ctx_create_ex, CHSGA_CREATE_EXACT_CLOSURE:
  // the semantics of ParSGA's serial DAG form, a topological insertion pass,
  // src/ParSGAoverrides/csr_align.hpp:316-327): a multi-char source also
  // gives its live last value, its first value plus len - 1 insertions, so
  // every pull from one gets a twin on its first value at weight w + len - 1
  if (c->exact) {
    auto twin = [&](std::vector<uint32_t>& off, std::vector<int2>& in) {
      std::vector<uint32_t> o2(off.size(), 0u);
      std::vector<int2> i2;
      i2.reserve(in.size() * 2);
      for (int64_t k = 0; k < n_s; ++k) {
        for (uint32_t e = off[k]; e < off[k + 1]; ++e) {
          i2.push_back(in[e]);
          if (in[e].x < 0) {
            const uint32_t x = uint32_t(in[e].x) & 0x7fffffffu;
            i2.push_back(make_int2(int32_t(x), in[e].y + int32_t(c->h_str_off[x + 1] - c->h_str_off[x] - 1)));
          }
        }
        o2[k + 1] = uint32_t(i2.size());
      }
      off.swap(o2);
      in.swap(i2);
    };
    twin(up_off, up_in);
    twin(dn_off, dn_in);
  }
  // first step of the plain path's LDS-resident top kernel: as many top
  // levels as fit kTopPlain string vertices
  auto top_step = [&](int64_t max_nt) {
    int32_t jt = J;
    while (jt > 0 && n_s - int64_t(c->h_lvl_off[jt - 1]) <= max_nt) --jt;
    return jt;
  };
  c->j_top = top_step(kTopPlain);
  if (c->top_level_opt >= 0) c->j_top = std::max(c->j_top, std::min<int32_t>(int32_t(c->top_level_opt), J));
  std::vector<uint32_t> pos_of_og(n_s);
  for (int64_t og = 0; og < n_s; ++og) pos_of_og[og] = kpos[g->rank[og]];

  if (c->cut) build_cut_tables(c);  // host only: it may still refuse the graph or the costs

  // every check above is host-only; the device is touched from here on
  hip_check(hipSetDevice(c->device), "hipSetDevice");
  c->d_scratch = noalloc_scratch(c->device);  // with the context's device current
  hip_check(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), "hipStreamCreate");
  hipStream_t s = c->stream;
  std::vector<uint8_t> labels(g->labels, g->labels + V);
  upload(c->d_labels, labels, s);
  upload(c->d_in_off, c->h_in_off, s);
  upload(c->d_in_src, c->h_in_src, s);
  upload(c->d_first_bits, first_bits, s);
  if (c->compact) {
    // plain compact kernel: at each first char its first two in-neighbours
    // (0xffffffff: none), bit 31 of .x set when it has more (read from the
    // in-CSR), so its gathers wait for one load instead of two
    std::vector<int2> nb(size_t(V) + 8, make_int2(-1, -1));
    for (int64_t st = 0; st < n_s; ++st) {
      const uint32_t p = c->h_str_off[st];
      const uint32_t e0 = c->h_in_off[p], e1 = c->h_in_off[p + 1];
      uint32_t x = e1 > e0 ? c->h_in_src[e0] : 0xffffffffu;
      const uint32_t y = e1 > e0 + 1 ? c->h_in_src[e0 + 1] : 0xffffffffu;
      if (e1 > e0 + 2) x |= 0x80000000u;
      nb[p] = make_int2(int32_t(x), int32_t(y));
    }
    upload(c->d_first_nb, nb, s);
  }
  if (c->compact) {
    // batched compact kernel: for every string, the last char (and, for a
    // string of more than one char, the first char) of the at most two
    // strings whose last char feeds its first char
    std::vector<int4> bnb(n_s + 32, make_int4(-1, -1, -1, -1));
    c->nb2_all = true;
    for (int64_t st = 0; st < n_s && c->nb2_all; ++st) {
      const uint32_t p = c->h_str_off[st];
      const uint32_t e0 = c->h_in_off[p], e1 = c->h_in_off[p + 1];
      if (e1 - e0 > 2) {
        c->nb2_all = false;
        break;
      }
      int32_t ent[4] = {-1, -1, -1, -1};
      for (uint32_t e = e0; e < e1; ++e) {
        const uint32_t u = c->h_in_src[e];
        const int64_t t = int64_t(std::upper_bound(c->h_str_off.begin(), c->h_str_off.end(), u) -
                                  c->h_str_off.begin()) - 1;  // string holding u
        if (u + 1 != c->h_str_off[t + 1]) {  // not a last char: not an expansion after all
          c->nb2_all = false;
          break;
        }
        ent[2 * (e - e0)] = int32_t(u);
        ent[2 * (e - e0) + 1] = c->h_str_off[t] == u ? -1 : int32_t(c->h_str_off[t]);
      }
      bnb[st] = make_int4(ent[0], ent[1], ent[2], ent[3]);
    }
    // label classes for the batched compact kernel (see read_class)
    std::vector<uint8_t> lclass(V);
    bool acgtn = true;
    for (int64_t p = 0; p < V; ++p) {
      switch (g->labels[p]) {
        case 'A': lclass[p] = 1; break;
        case 'C': lclass[p] = 2; break;
        case 'G': lclass[p] = 4; break;
        case 'T': lclass[p] = 8; break;
        case 'N': lclass[p] = c->alphabet == CHSGA_ALPHABET_MATCH_N ? 63 : 16; break;
        default: acgtn = false;
      }
    }
    c->batch_fast = c->nb2_all && acgtn;
    if (c->batch_fast) {
      upload(c->d_lclass, lclass, s);
      upload(c->d_bnb, bnb, s);
    }
    // packed rows (2-bit deltas, batch_packed.hip.inc): exact for unit costs;
    // their scan does more per char, so longer tiles (fewer chunk starts) pay
    c->packed = c->batch_fast && c->allow_packed && c->msc == 0 && c->ssc == 1 && c->dsc == 1 && c->isc == 1;
    if (c->packed) c->batch_tile_chars = 512;
    if (c->packed) {
      // per string: {string, length - 1} of its at most two in-neighbour strings
      // (both -1: none)
      std::vector<int4> pbnb(n_s + 32, make_int4(-1, -1, -1, -1));
      for (int64_t st = 0; st < n_s; ++st) {
        const int4 e = bnb[st];
        int32_t v[4] = {-1, 0, -1, 0};
        for (int q = 0; q < 2; ++q) {
          const int32_t l = q ? e.z : e.x;
          if (l < 0) continue;
          const int64_t x = int64_t(std::upper_bound(c->h_str_off.begin(), c->h_str_off.end(), uint32_t(l)) -
                                    c->h_str_off.begin()) - 1;
          v[2 * q] = int32_t(x);
          v[2 * q + 1] = int32_t(c->h_str_off[x + 1] - c->h_str_off[x]) - 1;
        }
        // one in-neighbour: it fills both entries (the row pass takes the
        // minimum over both)
        if (v[0] < 0) {
          v[0] = v[2];
          v[1] = v[3];
        }
        if (v[2] < 0) {
          v[2] = v[0];
          v[3] = v[1];
        }
        pbnb[st] = make_int4(v[0], v[1], v[2], v[3]);
      }
      // the row pass's chunk table is built from these per tile schedule
      c->h_pbnb = std::move(pbnb);
      c->h_lclass = lclass;
    }
  }
  if (!c->batch_fast) {
    // general batched kernel: first char of every char's string
    std::vector<uint32_t> cfirst(V);
    for (int64_t st = 0; st < n_s; ++st)
      for (uint32_t p = c->h_str_off[st]; p < c->h_str_off[st + 1]; ++p) cfirst[p] = c->h_str_off[st];
    upload(c->d_cfirst, cfirst, s);
  }
  upload(c->d_tile_char, tile_char, s);
  upload(c->d_tile_str, tile_str, s);
  upload(c->d_str_first, c->h_str_off, s);
  upload(c->d_sw_vert, sw_vert, s);
  c->h_sw_vert = sw_vert;
  upload(c->d_lvl_off, c->h_lvl_off, s);
  {
    // the batched level sweep reads up to kSweepPad offsets past its level
    std::vector<uint32_t> pad(up_off);
    pad.resize(n_s + 1 + kSweepPad, up_off[n_s]);
    upload(c->d_up_off, pad, s);
  }
  upload(c->d_up_in, up_in, s);
  {
    std::vector<uint32_t> pad(dn_off);
    pad.resize(n_s + 1 + kSweepPad, dn_off[n_s]);
    upload(c->d_dn_off, pad, s);
  }
  upload(c->d_dn_in, dn_in, s);
  c->ensure_staged_tables(g);
  build_top(c, c->top_plain, c->j_top, up_off, up_in, dn_off, dn_in, pos_of_og);
  {
    // batched sweep: first char of the vertex at each sweep position, and the
    // pull lists with the last char of each source string (padded)
    // longest pull list per level; the sweep loads a vertex's pulls in
    // chunks and may read up to (longest list + chunk) entries past the
    // start of any list of its level, hence the padding
    auto maxp = [&](const std::vector<uint32_t>& off, std::vector<uint32_t>& out) {
      out.assign(size_t(c->J), 0u);
      uint32_t all = 0;
      for (int32_t j = 0; j < c->J; ++j)
        for (uint32_t k = c->h_lvl_off[j]; k < c->h_lvl_off[j + 1]; ++k) out[j] = std::max(out[j], off[k + 1] - off[k]);
      for (uint32_t m : out) all = std::max(all, m);
      return all;
    };
    const uint32_t pad = std::max(maxp(up_off, c->h_up_maxp), maxp(dn_off, c->h_dn_maxp)) + 2 * kSweepChunk;
    // a pull reads the last char of its source string x (char rows) or, in
    // packed rows, plane PL of x (plane F when x has one char); a vertex's
    // own value is its first char (plane F)
    auto pull_src = [&](uint32_t x) -> int32_t {
      if (!c->packed) return int32_t(c->h_str_off[x + 1] - 1);
      return int32_t(3 * x + (c->h_str_off[x + 1] - c->h_str_off[x] > 1 ? 1 : 0));
    };
    auto by_last = [&](const std::vector<int2>& in) {
      std::vector<int2> out(in.size() + pad, make_int2(0, 0));
      for (size_t e = 0; e < in.size(); ++e) out[e] = make_int2(pull_src(uint32_t(in[e].x) & 0x7fffffffu), in[e].y);
      return out;
    };
    const std::vector<int2> bup = by_last(up_in), bdn = by_last(dn_in);
    upload(c->d_bup_in, bup, s);
    upload(c->d_bdn_in, bdn, s);
    std::vector<uint32_t> vfirst(n_s + kSweepPad, 0u);
    for (int64_t k = 0; k < n_s; ++k) vfirst[k] = c->packed ? 3 * sw_vert[k] : c->h_str_off[sw_vert[k]];
    upload(c->d_vfirst, vfirst, s);
    // levels whose lists hold at most R = 2, 4 or 8 pulls of weight < 2^16:
    // one record of 4, 8 or 16 words per vertex, {first, src_0 .. src_R-1,
    // w_0 | w_1 << 16, ...}; a missing pull is the vertex's own first char at
    // weight 0.  Level j's records start at word rbase[j]; zero records pad
    // each level by the widest wave, so a wave reads only in-bounds chars.
    auto records = [&](const std::vector<uint32_t>& off, const std::vector<int2>& pin,
                       const std::vector<uint32_t>& lmax, std::vector<uint8_t>& R, std::vector<uint64_t>& rbase) {
      std::vector<uint32_t> rec;
      R.assign(size_t(c->J), 0);
      rbase.assign(size_t(c->J), 0);
      for (int32_t j = 0; j < c->J; ++j) {
        const int r = lmax[j] <= 2 ? 2 : lmax[j] <= 4 ? 4 : lmax[j] <= 8 ? 8 : 0;
        bool fits = r > 0;
        for (uint32_t e = off[c->h_lvl_off[j]]; fits && e < off[c->h_lvl_off[j + 1]]; ++e)
          fits = pin[e].y >= 0 && pin[e].y < 65536;
        if (!fits) continue;
        const int W = sweep_record_words(r);
        R[j] = uint8_t(r);
        rbase[j] = rec.size();
        const uint32_t nv = c->h_lvl_off[j + 1] - c->h_lvl_off[j];
        rec.resize(rec.size() + size_t(nv + kSweepPad) * W, 0u);
        for (uint32_t k = c->h_lvl_off[j]; k < c->h_lvl_off[j + 1]; ++k) {
          uint32_t* w = rec.data() + rbase[j] + size_t(k - c->h_lvl_off[j]) * W;
          const uint32_t n = off[k + 1] - off[k], f = vfirst[k];
          w[0] = f;
          for (int q = 0; q < r; ++q) {
            const bool has = uint32_t(q) < n;
            w[1 + q] = has ? uint32_t(pin[off[k] + q].x) : f;
            w[1 + r + q / 2] |= (has ? uint32_t(pin[off[k] + q].y) : 0u) << (16 * (q & 1));
          }
        }
      }
      if (rec.empty()) rec.assign(4, 0u);
      return rec;
    };
    upload(c->d_bup_rec, records(up_off, bup, c->h_up_maxp, c->h_up_rec, c->h_up_rbase), s);
    upload(c->d_bdn_rec, records(dn_off, bdn, c->h_dn_maxp, c->h_dn_rec, c->h_dn_rbase), s);
    // The top of the hierarchy in one launch.  Levels L0 .. J - 1 (at most
    // kTopClosureMax vertices) pull only from each other and from lower
    // levels, whose values are final when the up sweep reaches L0: their up
    // then down sweep is a fixed min-plus map from source values to the
    // final first values of the top vertices.  Running the pulls on symbols
    // (one-char top vertices relay what reached them; a multi-char source
    // gives its stale last value, a lower vertex its pull value) gives each
    // top vertex a list of {source, summed weight} terms, and one level
    // launch over those lists replaces the top levels' launches.  Weights
    // compose linearly (each pull adds w * ins), so the result is the
    // level-by-level sweep's.  Terms whose weight * ins reaches 2^30 are
    // dropped: no row value reaches them.  Terms are keyed by the pull
    // lists' string codes (string id, | 2^31 for the stale last value of a
    // multi-char string), which the plain sweep reads from F / Lst; the
    // batched sweep gets the matching row elements.
    // The batched sweep (64 slots a vertex) and the plain sweep (one read)
    // get closures of their own size: a longer closure list costs the
    // batched launch more than the level launches it replaces, the plain
    // path's level launches cost their launch latency.
    c->top_l0 = c->top_l0_plain = c->J;
    auto build_top = [&](int64_t cap, bool plain) {
      const int32_t J = c->J;
      int32_t L0 = J;
      while (L0 > 1 && c->h_lvl_off[J] - c->h_lvl_off[L0 - 1] <= uint32_t(cap)) --L0;
      {
        if (L0 > J - 2) return;
        const uint32_t kT0 = c->h_lvl_off[L0], kT1 = c->h_lvl_off[J];
        using Terms = std::vector<std::pair<uint32_t, int64_t>>;
        std::vector<Terms> terms(kT1 - kT0);
        std::unordered_map<uint32_t, int64_t> acc;  // the terms of the vertex being pulled
        auto add = [&](uint32_t src, int64_t o) {
          const auto it = acc.find(src);
          if (it == acc.end())
            acc.emplace(src, o);
          else
            it->second = std::min(it->second, o);
        };
        int64_t n_terms = 0;
        bool fits = true;
        for (uint32_t k = kT0; k < kT1; ++k) terms[k - kT0].emplace_back(sw_vert[k], 0);
        auto pull = [&](uint32_t k, const std::vector<uint32_t>& off, const std::vector<int2>& pin) {
          if (!fits) return;
          Terms& tk = terms[k - kT0];
          acc.clear();
          for (const auto& t : tk) acc.emplace(t.first, t.second);
          for (uint32_t e = off[k]; e < off[k + 1]; ++e) {
            const uint32_t x = uint32_t(pin[e].x) & 0x7fffffffu;
            const int64_t w = pin[e].y;
            const uint32_t px = pos_of_og[x];
            if (pin[e].x >= 0 && px >= kT0) {
              // the first value of a top vertex relays its terms (every
              // pull of a one-char string; a multi-char string's twin entry
              // with exact closure)
              for (const auto& t : terms[px - kT0]) add(t.first, t.second + w);
            } else {
              add(uint32_t(pin[e].x), w);
            }
          }
          n_terms += int64_t(acc.size()) - int64_t(tk.size());
          tk.assign(acc.begin(), acc.end());
          fits = n_terms <= kTopClosureTermsMax;
        };
        // row element of a term: the first value of a string, or the stale
        // last value of a multi-char one
        auto element = [&](uint32_t code) -> int32_t {
          const uint32_t x = code & 0x7fffffffu;
          if (code >> 31) return pull_src(x);
          return int32_t(c->packed ? 3 * x : c->h_str_off[x]);
        };
        for (int32_t j = L0; j < J; ++j)
          for (uint32_t k = c->h_lvl_off[j]; k < c->h_lvl_off[j + 1]; ++k) pull(k, up_off, up_in);
        for (int32_t j = J - 2; j >= L0; --j)
          for (uint32_t k = c->h_lvl_off[j]; k < c->h_lvl_off[j + 1]; ++k) pull(k, dn_off, dn_in);
        if (fits) {
        std::vector<uint32_t> toff(size_t(n_s) + 1 + kSweepPad, 0u);
        std::vector<int2> tin;
        uint32_t tmax = 0;
        for (uint32_t k = 0; k <= uint32_t(n_s); ++k) {
          toff[k] = uint32_t(tin.size());
          if (k < kT0 || k >= kT1) continue;
          for (const auto& t : terms[k - kT0]) {
            if (t.first == sw_vert[k] && t.second == 0) continue;  // the vertex's own value
            if (int64_t(c->isc) * t.second >= (int64_t(1) << 30)) continue;
            tin.push_back(make_int2(plain ? int32_t(t.first) : element(t.first), int32_t(t.second)));
          }
          tmax = std::max(tmax, uint32_t(tin.size()) - toff[k]);
        }
        for (size_t k = size_t(n_s) + 1; k < toff.size(); ++k) toff[k] = uint32_t(tin.size());
        tin.resize(tin.size() + tmax + 2 * kSweepChunk, make_int2(0, 0));
        upload(plain ? c->d_top_off_plain : c->d_top_off, toff, s);
        upload(plain ? c->d_top_in_plain : c->d_top_in, tin, s);
        if (plain) {
          c->top_l0_plain = L0;
        } else {
          c->top_l0 = L0;
          c->top_maxp = tmax;
        }
        }
      }
    };
    build_top(kTopClosureMax, false);
    build_top(kTopClosurePlainMax, true);
  }
  c->h_up_off = std::move(up_off);
  c->h_dn_off = std::move(dn_off);
  c->d_F.alloc(sizeof(int32_t) * n_s);
  c->d_Lst.alloc(sizeof(int32_t) * n_s);
  c->d_row[0].alloc(sizeof(int32_t) * V);
  c->d_row[1].alloc(sizeof(int32_t) * V);
  hip_check(hipStreamSynchronize(s), "context upload");
}

// Row launches for storage type T (int32_t, or int16_t when narrow).
template <typename T>
static void launch_row_t(chsga_ctx* c, const void* H, void* D, uint8_t rc) {
  const hipStream_t s = c->stream;
  hipLaunchKernelGGL((c->compact ? k_row_tile8<true, T> : k_row_tile8<false, T>), dim3(unsigned(c->n_tiles)),
                     dim3(CHSGA_BLOCK), 0, s, (const T*)H, (T*)D, (T*)c->d_F.p, (T*)c->d_Lst.p,
                     (const uint8_t*)c->d_labels.p, (const uint32_t*)c->d_in_off.p, (const uint32_t*)c->d_in_src.p,
                     (const int2*)c->d_first_nb.p, (const uint32_t*)c->d_first_bits.p,
                     (const uint32_t*)c->d_tile_char.p, (const uint32_t*)c->d_tile_str.p, c->V, rc, c->alphabet,
                     c->msc, c->ssc, c->dsc, c->isc, LocalRow{});
}

template <typename T>
static void launch_sweeps_t(chsga_ctx* c, int mode) {
  const hipStream_t s = c->stream;
  T* F = (T*)c->d_F.p;
  const T* L = (const T*)c->d_Lst.p;
  auto level = [&](uint32_t k0, uint32_t k1, const void* off, const void* in) {
    hipLaunchKernelGGL(k_sweep_level<T>, dim3((k1 - k0 + CHSGA_BLOCK - 1) / CHSGA_BLOCK), dim3(CHSGA_BLOCK), 0, s, F, L,
                       (const uint32_t*)c->d_sw_vert.p, (const uint32_t*)off, (const int2*)in, k0, k1, c->isc);
    c->launches += 1;
  };
  if (mode == 3 && c->top_closure && c->top_l0_plain < c->J) {
    // levels below top_l0_plain one by one, the top levels as one closure launch
    const int32_t jt = c->top_l0_plain;
    for (int32_t j = 1; j < jt; ++j) {
      const uint32_t k0 = c->h_lvl_off[j], k1 = c->h_lvl_off[j + 1];
      if (c->h_up_off[k1] != c->h_up_off[k0]) level(k0, k1, c->d_up_off.p, c->d_up_in.p);
    }
    {
      constexpr int G = 16;  // lanes per top vertex
      const uint32_t k0 = c->h_lvl_off[jt], k1 = c->h_lvl_off[c->J];
      const uint64_t threads = uint64_t(k1 - k0) * G;
      hipLaunchKernelGGL((k_sweep_group<T, G>), dim3(unsigned((threads + CHSGA_BLOCK - 1) / CHSGA_BLOCK)),
                         dim3(CHSGA_BLOCK), 0, s, F, L, (const uint32_t*)c->d_sw_vert.p,
                         (const uint32_t*)c->d_top_off_plain.p, (const int2*)c->d_top_in_plain.p, k0, k1, c->isc);
      c->launches += 1;
    }
    for (int32_t j = jt - 1; j >= 0; --j) {
      const uint32_t k0 = c->h_lvl_off[j], k1 = c->h_lvl_off[j + 1];
      if (c->h_dn_off[k1] != c->h_dn_off[k0]) level(k0, k1, c->d_dn_off.p, c->d_dn_in.p);
    }
    return;
  }
  if (mode & 1)
    for (int32_t j = 1; j < c->j_top; ++j) {
      const uint32_t k0 = c->h_lvl_off[j], k1 = c->h_lvl_off[j + 1];
      if (c->h_up_off[k1] == c->h_up_off[k0]) continue;
      hipLaunchKernelGGL(k_sweep_level<T>, dim3((k1 - k0 + CHSGA_BLOCK - 1) / CHSGA_BLOCK), dim3(CHSGA_BLOCK), 0, s, F,
                         L, (const uint32_t*)c->d_sw_vert.p, (const uint32_t*)c->d_up_off.p,
                         (const int2*)c->d_up_in.p, k0, k1, c->isc);
      c->launches += 1;
    }
  const TopSet& TS = c->top_plain;
  if (TS.NT > 0) {
    hipLaunchKernelGGL(k_sweep_top_lds<T>, dim3(1), dim3(1024), size_t(2) * TS.NT * sizeof(int32_t), s, F, L,
                       (const uint32_t*)c->d_sw_vert.p, TS.K0, TS.NT, (const uint32_t*)c->d_lvl_off.p, TS.j_top, c->J,
                       (const uint32_t*)TS.ext_off.p, (const int2*)TS.ext.p, (const uint32_t*)TS.upi_off.p,
                       (const int2*)TS.upi.p, (const uint32_t*)TS.dni_off.p, (const int2*)TS.dni.p, c->isc,
                       int32_t(mode));
    c->launches += 1;
  }
  if (mode & 2)
    for (int32_t j = std::min(c->j_top, c->J) - 1; j >= 0; --j) {
      const uint32_t k0 = c->h_lvl_off[j], k1 = c->h_lvl_off[j + 1];
      if (c->h_dn_off[k1] == c->h_dn_off[k0]) continue;
      hipLaunchKernelGGL(k_sweep_level<T>, dim3((k1 - k0 + CHSGA_BLOCK - 1) / CHSGA_BLOCK), dim3(CHSGA_BLOCK), 0, s, F,
                         L, (const uint32_t*)c->d_sw_vert.p, (const uint32_t*)c->d_dn_off.p,
                         (const int2*)c->d_dn_in.p, k0, k1, c->isc);
      c->launches += 1;
    }
}

template <typename T>
static void launch_post_t(chsga_ctx* c, void* D) {
  hipLaunchKernelGGL(k_post<T>, dim3(unsigned((c->n_s + CHSGA_BLOCK - 1) / CHSGA_BLOCK)), dim3(CHSGA_BLOCK), 0,
                     c->stream, (T*)D, (const T*)c->d_F.p, (const uint32_t*)c->d_str_first.p, c->n_s, c->isc);
  c->launches += 1;
}

void chsga_ctx::launch_row(const void* H, void* D, uint8_t rc) {
  const hipStream_t s = stream;
  if (profile) hip_check(hipEventRecord(ev_pool[ev_used++], s), "event");
  if (narrow)
    launch_row_t<int16_t>(this, H, D, rc);
  else
    launch_row_t<int32_t>(this, H, D, rc);
  if (profile) hip_check(hipEventRecord(ev_pool[ev_used++], s), "event");
  launches += 1;
  launch_insertion(D, rc);
  if (profile) hip_check(hipEventRecord(ev_pool[ev_used++], s), "event");
}

void chsga_ctx::launch_sweeps(int mode) {
  if (narrow)
    launch_sweeps_t<int16_t>(this, mode);
  else
    launch_sweeps_t<int32_t>(this, mode);
}

void chsga_ctx::launch_insertion(void* D, uint8_t rc) {
  if (cut_on && (narrow ? launch_cut_t<int16_t>(this, D, rc) : launch_cut_t<int32_t>(this, D, rc))) return;
  launch_sweeps(3);
  if (narrow)
    launch_post_t<int16_t>(this, D);
  else
    launch_post_t<int32_t>(this, D);
}

void chsga_ctx::launch_init(void* H, uint8_t c0) {
  const dim3 g(unsigned((V + CHSGA_BLOCK - 1) / CHSGA_BLOCK));
  if (narrow)
    hipLaunchKernelGGL(k_init_row<int16_t>, g, dim3(CHSGA_BLOCK), 0, stream, (int16_t*)H, (const uint8_t*)d_labels.p,
                       V, c0, alphabet, msc, ssc);
  else
    hipLaunchKernelGGL(k_init_row<int32_t>, g, dim3(CHSGA_BLOCK), 0, stream, (int32_t*)H, (const uint8_t*)d_labels.p,
                       V, c0, alphabet, msc, ssc);
  launches += 1;
}

void chsga_ctx::launch_argmin(const void* D, unsigned long long* key) {
  if (narrow)
    hipLaunchKernelGGL(k_argmin<int16_t>, dim3(512), dim3(CHSGA_BLOCK), 0, stream, (const int16_t*)D, V, key);
  else
    hipLaunchKernelGGL(k_argmin<int32_t>, dim3(512), dim3(CHSGA_BLOCK), 0, stream, (const int32_t*)D, V, key);
}

// Largest value any row of a read of length len can store: row 0 is at most
// max(match, subst) and every later row at most the row above plus del.
bool chsga_ctx::fits_narrow(int64_t max_len) const {
  if (!allow_narrow) return false;
  const int64_t bound = int64_t(std::max(msc, ssc)) + (max_len - 1) * int64_t(dsc);
  return bound <= 32767;
}

// Forward pass of one read into the two-row ping-pong buffer; returns the
// buffer index holding the last row.  ckpt (optional) receives row 0 and
// every row i with i % k == 0.
int chsga_ctx::forward(const char* read, int64_t len, int64_t k, int32_t** ckpt_rows) {
  if (ckpt_rows && narrow) throw std::logic_error("checkpointed forward pass runs with int32 rows");
  void* R[2] = {d_row[0].p, d_row[1].p};
  int cur = 0;
  launch_init(R[cur], uint8_t(read[0]));
  if (ckpt_rows)
    hip_check(hipMemcpyAsync(ckpt_rows[0], R[cur], sizeof(int32_t) * V, hipMemcpyDefault, stream), "ckpt");
  for (int64_t i = 1; i < len; ++i) {
    if (profile && ev_used + 3 > ev_pool.size()) flush_profile();
    launch_row(R[cur], R[cur ^ 1], uint8_t(read[i]));
    cur ^= 1;
    if (ckpt_rows && i % k == 0)
      hip_check(hipMemcpyAsync(ckpt_rows[i / k], R[cur], sizeof(int32_t) * V, hipMemcpyDefault, stream),
                "ckpt");  // device or pinned host row
  }
  rows += len;
  return cur;
}

void chsga_ctx::flush_profile() {
  if (!profile || ev_used == 0) return;
  for (size_t i = 0; i + 2 < ev_used; i += 3) {
    // each triple is on one stream; concurrent batch groups use several
    hip_check(hipEventSynchronize(ev_pool[i + 2]), "event sync");
    float a = 0, b = 0;
    hip_check(hipEventElapsedTime(&a, ev_pool[i], ev_pool[i + 1]), "elapsed");
    hip_check(hipEventElapsedTime(&b, ev_pool[i + 1], ev_pool[i + 2]), "elapsed");
    prof_row_ms += a;
    prof_ins_ms += b;
    prof_rows += 1;
  }
  ev_used = 0;
}

chsga_ctx::~chsga_ctx() {
  for (hipEvent_t e : ev_pool) (void)hipEventDestroy(e);
  if (stream) (void)hipStreamDestroy(stream);
}

#define ENGINE_GUARD(...)                  \
  try {                                     \
    __VA_ARGS__;                                   \
    return CHSGA_OK;                        \
  } catch (const std::invalid_argument& e) { \
    set_error(e.what());                    \
    return CHSGA_EINVAL;                    \
  } catch (const HipError& e) {             \
    set_error(e.what());                    \
    return CHSGA_EDEVICE;                   \
  } catch (const std::bad_alloc&) {         \
    set_error("host allocation failed");    \
    return CHSGA_ENOMEM;                    \
  } catch (const std::exception& e) {       \
    set_error(e.what());                    \
    return CHSGA_EINTERNAL;                 \
  }

extern "C" {

const char* chsga_last_error(void) { return g_err.c_str(); }

int chsga_device_count(int* count) {
  if (!count) {
    set_error("null argument");
    return CHSGA_EINVAL;
  }
  ENGINE_GUARD({
    int n = 0;
    hip_check(hipGetDeviceCount(&n), "hipGetDeviceCount");
    *count = n;
  })
}

int chsga_mem_info(chsga_ctx* ctx, int64_t* free_bytes, int64_t* total_bytes) {
  if (!ctx || !free_bytes || !total_bytes) {
    set_error("null argument");
    return CHSGA_EINVAL;
  }
  ENGINE_GUARD({
    hip_check(hipSetDevice(ctx->device), "hipSetDevice");
    size_t fr = 0, to = 0;
    hip_check(hipMemGetInfo(&fr, &to), "hipMemGetInfo");
    *free_bytes = int64_t(fr);
    *total_bytes = int64_t(to);
  })
}

int chsga_ctx_create(const chsga_graph_desc* g, int device, chsga_ctx** out) {
  return chsga_ctx_create_ex(g, device, 0u, out);
}

int chsga_ctx_create_ex(const chsga_graph_desc* g, int device, uint32_t flags, chsga_ctx** out) {
  if (!out) {
    set_error("null output pointer");
    return CHSGA_EINVAL;
  }
  *out = nullptr;
  chsga_ctx* c = nullptr;
  ENGINE_GUARD({
    if (flags & ~uint32_t(CHSGA_CREATE_CHAR_ROWS | CHSGA_CREATE_EXACT_CLOSURE | CHSGA_CREATE_CUT_CLOSURE))
      throw std::invalid_argument("unknown context flags");
    c = new chsga_ctx();
    static std::atomic<uint64_t> next_serial{1};
    c->serial = next_serial++;
    try {
      c->device = device;
      c->allow_packed = (flags & CHSGA_CREATE_CHAR_ROWS) == 0;
      c->cut = (flags & CHSGA_CREATE_CUT_CLOSURE) != 0;
      c->exact = c->cut || (flags & CHSGA_CREATE_EXACT_CLOSURE) != 0;
      build_context(c, g);
      if (c->cut) {
        upload_cut_tables(c, c->stream);  // build_context leaves the context's device current
        c->cut_on = true;
      }
    } catch (...) {
      delete c;
      throw;
    }
    *out = c;
  })
}

void chsga_ctx_destroy(chsga_ctx* ctx) {
  if (!ctx) return;
  if (ctx->worker.joinable()) ctx->worker.join();
  (void)hipSetDevice(ctx->device);
  delete ctx;
}

int chsga_ctx_set_option(chsga_ctx* ctx, int option, int64_t value) {
  if (!ctx) {
    set_error("null context");
    return CHSGA_EINVAL;
  }
  ENGINE_GUARD({
    switch (option) {
      case CHSGA_OPT_PROFILE:
        ctx->profile = value != 0;
        if (ctx->profile && ctx->ev_pool.empty()) {
          ctx->ev_pool.resize(3 * 512);
          for (auto& e : ctx->ev_pool) hip_check(hipEventCreate(&e), "hipEventCreate");
        }
        break;
      case CHSGA_OPT_USE_GRAPHS:
        ctx->use_graphs = value != 0;
        break;
      case CHSGA_OPT_NARROW:
        ctx->allow_narrow = value != 0;
        break;
      case CHSGA_OPT_STAGED_WALK:
        if (value != 0 && ctx->exact)
          throw std::invalid_argument("the staged walk follows the stale-last sweep; exact-closure rows need none");
        ctx->force_staged = value != 0;
        break;
      case CHSGA_OPT_BATCHED_BACKWARD:
        ctx->batched_backward = value != 0;
        break;
      case CHSGA_OPT_CKPT_HBM_ROWS:
        ctx->ckpt_hbm_rows = value < 0 ? -1 : value;
        break;
      case CHSGA_OPT_SEG_ROWS:
        ctx->seg_rows_cap = value < 0 ? -1 : value;
        break;
      case CHSGA_OPT_TOP_CLOSURE:
        ctx->top_closure = value != 0;
        break;
      case CHSGA_OPT_RING_ROWS:
        ctx->ring_rows_cap = value < 0 ? -1 : value;
        break;
      case CHSGA_OPT_FORWARD_TAIL:
        if (value < 0) throw std::invalid_argument("CHSGA_OPT_FORWARD_TAIL must be >= 0");
        ctx->forward_tail = value;
        break;
      case CHSGA_OPT_BACKWARD_TAIL:
        if (value < 0) throw std::invalid_argument("CHSGA_OPT_BACKWARD_TAIL must be >= 0");
        ctx->backward_tail = value;
        break;
      case CHSGA_OPT_BACKWARD_BAND:
        if (value < 0) throw std::invalid_argument("CHSGA_OPT_BACKWARD_BAND must be >= 0");
        ctx->backward_band = value;
        break;
      case CHSGA_OPT_CUT_CLOSURE:
        if (value != 0 && !ctx->cut)
          throw std::invalid_argument("no cut tables: create the context with CHSGA_CREATE_CUT_CLOSURE");
        ctx->cut_on = value != 0;
        break;
      case CHSGA_OPT_PACKED_ROWS:
        if ((value != 0) != ctx->packed)
          throw std::invalid_argument(ctx->packed ? "the batched rows are packed"
                                                  : "the batched rows are not packed (chsga_ctx_create_ex decides)");
        break;
      default:
        throw std::invalid_argument("unknown option");
    }
  })
}

int chsga_score_reads(chsga_ctx* ctx, const char* const* reads, const int64_t* lens, int64_t n, int32_t* scores_out,
                      double* total_sec, double* sub_sec, double* ins_sec, chsga_stats* stats) {
  if (!ctx || (n > 0 && (!reads || !lens || !scores_out))) {
    set_error("null argument");
    return CHSGA_EINVAL;
  }
  ENGINE_GUARD({
    hip_check(hipSetDevice(ctx->device), "hipSetDevice");
    for (int64_t r = 0; r < n; ++r)
      if (lens[r] <= 0 || !reads[r]) throw std::invalid_argument("empty read");
    const int64_t t0 = now_ns();
    ctx->launches = 0;
    ctx->rows = 0;
    ctx->prof_row_ms = ctx->prof_ins_ms = 0;
    ctx->prof_rows = 0;
    ctx->ev_used = 0;
    DevMem keys;
    keys.alloc(sizeof(unsigned long long) * std::max<int64_t>(1, n));
    unsigned long long* dk = (unsigned long long*)keys.p;
    hipLaunchKernelGGL(k_fill_u64, dim3(unsigned((n + 255) / 256 + 1)), dim3(256), 0, ctx->stream, dk, n, ~0ull);
    int64_t cells = 0;
    for (int64_t r = 0; r < n; ++r) {
      const int64_t r0 = now_ns();
      const double pr0 = ctx->prof_row_ms, pi0 = ctx->prof_ins_ms;
      ctx->narrow = ctx->fits_narrow(lens[r]);
      const int cur = ctx->forward(reads[r], lens[r], 0, nullptr);
      ctx->launch_argmin(ctx->d_row[cur].p, dk + r);
      ctx->narrow = false;
      cells += lens[r] * ctx->V;
      if (total_sec || sub_sec || ins_sec || ctx->profile) {
        ctx->flush_profile();
        hip_check(hipStreamSynchronize(ctx->stream), "sync");
        const double sec = double(now_ns() - r0) * 1e-9;
        if (total_sec) total_sec[r] = sec;
        if (sub_sec) sub_sec[r] = ctx->profile ? (ctx->prof_row_ms - pr0) * 1e-3 : 0.0;
        if (ins_sec) ins_sec[r] = ctx->profile ? (ctx->prof_ins_ms - pi0) * 1e-3 : 0.0;
      }
    }
    std::vector<unsigned long long> hk(n);
    if (n)
      hip_check(hipMemcpyAsync(hk.data(), dk, sizeof(unsigned long long) * n, hipMemcpyDeviceToHost, ctx->stream),
                "copy scores");
    hip_check(hipStreamSynchronize(ctx->stream), "sync");
    hip_check(hipGetLastError(), "kernel launch");
    for (int64_t r = 0; r < n; ++r) scores_out[r] = int32_t(hk[r] >> 32);
    if (stats) {
      std::memset(stats, 0, sizeof(*stats));
      stats->wall_ms = double(now_ns() - t0) * 1e-6;
      stats->row_kernel_ms = ctx->prof_row_ms;
      stats->row_kernel_launches = ctx->prof_rows;
      stats->sweep_ms = ctx->prof_ins_ms;
      stats->cells = cells;
      stats->kernel_launches = ctx->launches;
      stats->rows = ctx->rows;
    }
  })
}

int chsga_score_reads_noalloc(chsga_ctx* ctx, const char* const* reads, const int64_t* lens, int64_t n,
                              int32_t* scores_out) {
  if (!ctx || (n > 0 && (!reads || !lens || !scores_out)) || n < 0) {
    set_error("null argument");
    return CHSGA_EINVAL;
  }
  if (n > CHSGA_NOALLOC_MAX_READS) {
    set_error("chsga_score_reads_noalloc: more than CHSGA_NOALLOC_MAX_READS (65536) reads in one call");
    return CHSGA_EINVAL;
  }
  ENGINE_GUARD({
    hip_check(hipSetDevice(ctx->device), "hipSetDevice");
    if (!ctx->d_scratch) throw std::logic_error("no-allocation scratch not set up for this context");
    for (int64_t r = 0; r < n; ++r)
      if (lens[r] <= 0 || !reads[r]) throw std::invalid_argument("empty read");
    if (n == 0) return CHSGA_OK;
    // the device's scratch array: argmin keys, then the int32 scores
    unsigned long long* keys = static_cast<unsigned long long*>(ctx->d_scratch);
    int32_t* sc = reinterpret_cast<int32_t*>(keys + CHSGA_NOALLOC_MAX_READS);
    const bool prof = ctx->profile;
    ctx->profile = false;
    hipLaunchKernelGGL(k_fill_u64, dim3(unsigned((n + 255) / 256)), dim3(256), 0, ctx->stream, keys, n, ~0ull);
    for (int64_t r = 0; r < n; ++r) {
      ctx->narrow = ctx->fits_narrow(lens[r]);
      const int cur = ctx->forward(reads[r], lens[r], 0, nullptr);
      ctx->launch_argmin(ctx->d_row[cur].p, keys + r);
      ctx->narrow = false;
    }
    ctx->profile = prof;
    hipLaunchKernelGGL(k_key_scores, dim3(unsigned((n + 255) / 256)), dim3(256), 0, ctx->stream,
                       (const unsigned long long*)keys, sc, n);
    hip_check(hipMemcpyAsync(scores_out, sc, sizeof(int32_t) * size_t(n), hipMemcpyDeviceToHost, ctx->stream),
              "copy scores");
    hip_check(hipStreamSynchronize(ctx->stream), "sync");
    hip_check(hipGetLastError(), "kernel launch");
  })
}

int chsga_last_row(chsga_ctx* ctx, const char* read, int64_t len, int32_t* row_out) {
  if (!ctx || !read || len <= 0 || !row_out) {
    set_error("bad argument");
    return CHSGA_EINVAL;
  }
  ENGINE_GUARD({
    hip_check(hipSetDevice(ctx->device), "hipSetDevice");
    const bool prof = ctx->profile;
    ctx->profile = false;
    ctx->narrow = false;
    const int cur = ctx->forward(read, len, 0, nullptr);
    ctx->profile = prof;
    hip_check(hipMemcpyAsync(row_out, ctx->d_row[cur].p, sizeof(int32_t) * ctx->V, hipMemcpyDeviceToHost,
                             ctx->stream),
              "copy row");
    hip_check(hipStreamSynchronize(ctx->stream), "sync");
  })
}

}  // extern "C"

#include "batch.hip.inc"
#include "traceback.hip.inc"
#include "mlp.hip.inc"
#include "pasgal.hip.inc"
