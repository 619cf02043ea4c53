// Host-side formats, generator and C ABI (chsga/host.h).
#include <zlib.h>

#include <algorithm>
#include <array>
#include <functional>
#include <unordered_map>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <memory>
#include <sstream>

#include "host_internal.hpp"

namespace chsga {

static thread_local std::string g_err;
void set_error(const std::string& msg) { g_err = msg; }

double Rng::normal() {
  // Box-Muller
  double u1 = uniform();
  while (u1 <= 0.0) u1 = uniform();
  const double u2 = uniform();
  return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
}

int64_t Rng::geometric(double p) {
  if (p >= 1.0) return 0;
  if (p <= 0.0) return INT64_MAX / 4;
  double u = uniform();
  while (u <= 0.0) u = uniform();
  return int64_t(std::floor(std::log(u) / std::log1p(-p)));
}

void build_in_csr(int64_t n, const std::vector<uint64_t>& off, const std::vector<uint32_t>& dst,
                  std::vector<uint64_t>& in_off, std::vector<uint32_t>& in_src) {
  in_off.assign(n + 1, 0);
  for (int64_t u = 0; u < n; ++u)
    for (uint64_t e = off[u]; e < off[u + 1]; ++e) ++in_off[dst[e] + 1];
  for (int64_t v = 0; v < n; ++v) in_off[v + 1] += in_off[v];
  in_src.resize(in_off[n]);
  std::vector<uint64_t> cur(in_off.begin(), in_off.end() - 1);
  for (int64_t u = 0; u < n; ++u)
    for (uint64_t e = off[u]; e < off[u + 1]; ++e) in_src[cur[dst[e]]++] = uint32_t(u);
}

void expand_char_graph(chsga_dataset& ds) {
  const int64_t n = ds.n_s;
  ds.V = int64_t(ds.str_off[n]);
  ds.char_out_off.assign(ds.V + 1, 0);
  ds.char_out_dst.clear();
  ds.char_out_dst.reserve(ds.V + ds.str_out_dst.size());
  int64_t p = 0;
  for (int64_t s = 0; s < n; ++s) {
    const uint64_t a = ds.str_off[s], b = ds.str_off[s + 1];
    for (uint64_t c = a; c < b; ++c) {
      if (c + 1 < b) {
        ds.char_out_dst.push_back(uint32_t(c + 1));
      } else {
        for (uint64_t e = ds.str_out_off[s]; e < ds.str_out_off[s + 1]; ++e)
          ds.char_out_dst.push_back(uint32_t(ds.str_off[ds.str_out_dst[e]]));
      }
      ds.char_out_off[++p] = ds.char_out_dst.size();
    }
  }
  build_in_csr(ds.V, ds.char_out_off, ds.char_out_dst, ds.in_off, ds.in_src);
}

void validate_dataset(const chsga_dataset& ds) {
  if (int64_t(ds.str_off.size()) != ds.n_s + 1) throw std::runtime_error("label offsets: expected n_s+1 values");
  if (ds.str_off[0] != 0) throw std::runtime_error("label offsets must start at 0");
  for (int64_t s = 0; s < ds.n_s; ++s)
    if (ds.str_off[s + 1] <= ds.str_off[s]) throw std::runtime_error("label offsets: empty or decreasing string vertex");
  if (ds.str_off[ds.n_s] != uint64_t(ds.labels.size()))
    throw std::runtime_error("labels length " + std::to_string(ds.labels.size()) + " != last label offset " +
                             std::to_string(ds.str_off[ds.n_s]));
  if (ds.V != int64_t(ds.labels.size())) throw std::runtime_error("char graph vertex count != labels length");
  if (ds.V >= (int64_t(1) << 31)) throw std::runtime_error("more than 2^31 char vertices");
  for (uint32_t v : ds.in_src)
    if (int64_t(v) >= ds.V) throw std::runtime_error("char edge endpoint out of range");
  for (uint32_t v : ds.str_out_dst)
    if (int64_t(v) >= ds.n_s) throw std::runtime_error("string edge endpoint out of range");
  const ContractionHierarchy& ch = ds.ch;
  if (ch.n != ds.n_s) throw std::runtime_error("CH vertex count != string vertex count");
  std::vector<uint8_t> seen(ds.n_s, 0);
  for (uint32_t r : ch.rank) {
    if (int64_t(r) >= ds.n_s || seen[r]) throw std::runtime_error("CH rank is not a permutation");
    seen[r] = 1;
  }
  if (ch.layer_off.empty() || ch.layer_off.back() != uint64_t(ds.n_s)) throw std::runtime_error("CH layer offsets");
  if (ch.cc_off.empty() || ch.cc_off.back() != ch.layer_off.size() - 1) throw std::runtime_error("CH cc offsets");
  // Every CH edge must go from a lower layer to a strictly higher layer of the
  // same component: this is what makes the level-synchronous sweep exact.
  std::vector<uint32_t> layer_of(ds.n_s);
  for (size_t l = 0; l + 1 < ch.layer_off.size(); ++l)
    for (uint64_t c = ch.layer_off[l]; c < ch.layer_off[l + 1]; ++c) layer_of[c] = uint32_t(l);
  for (int64_t c = 0; c < ds.n_s; ++c) {
    for (uint64_t e = ch.up_off[c]; e < ch.up_off[c + 1]; ++e)
      if (int64_t(ch.up_v[e]) >= ds.n_s || layer_of[ch.up_v[e]] <= layer_of[c])
        throw std::runtime_error("CH up-edge does not go to a higher layer");
    for (uint64_t e = ch.dn_off[c]; e < ch.dn_off[c + 1]; ++e)
      if (int64_t(ch.dn_v[e]) >= ds.n_s || layer_of[ch.dn_v[e]] <= layer_of[c])
        throw std::runtime_error("CH down-edge does not come from a higher layer");
  }
}

// ---------------------------------------------------------------------------
// generator
// ---------------------------------------------------------------------------
static const char kBases[4] = {'A', 'C', 'G', 'T'};

static void generate(const chsga_graph_params& p, chsga_dataset& ds) {
  Rng rng(p.seed);
  const int64_t N = p.backbone_bp;
  if (N < 2) throw std::runtime_error("backbone too short");
  std::string bb(N, 'A');
  for (int64_t i = 0; i < N; ++i) bb[i] = kBases[rng.next() & 3];
  const double total = p.snp_rate + p.indel_rate + p.sv_rate;
  std::vector<std::string> strs;
  std::vector<std::pair<uint32_t, uint32_t>> edges;
  std::vector<GenBlock> blocks;
  std::string seg;
  int64_t pos = 0;
  std::vector<uint32_t> prev_tails;  // strings whose last char leads into the next segment
  bool prev_skip_from_seg = false;
  uint32_t prev_seg = UINT32_MAX;
  auto emit_segment = [&](std::string&& s) {
    const uint32_t id = uint32_t(strs.size());
    strs.push_back(std::move(s));
    blocks.push_back({id, 0, false});
    for (uint32_t t : prev_tails) edges.push_back({t, id});
    if (prev_skip_from_seg) edges.push_back({prev_seg, id});
    prev_tails.clear();
    prev_skip_from_seg = false;
    prev_seg = id;
  };
  while (true) {
    const int64_t gap = 1 + (total > 0 ? rng.geometric(total) : N);
    const double r = rng.uniform() * (total > 0 ? total : 1.0);
    int kind = 0;  // 0 snp, 1 indel, 2 sv
    if (r >= p.snp_rate) kind = (r < p.snp_rate + p.indel_rate) ? 1 : 2;
    int64_t elen = 0;
    bool del = false;
    if (kind == 0) {
      elen = 1;
    } else {
      const int32_t lo = kind == 1 ? p.indel_min : p.sv_min, hi = kind == 1 ? p.indel_max : p.sv_max;
      elen = rng.range(std::max(1, lo), std::max(std::max(1, lo), hi));
      del = (rng.next() >> 63) != 0;
    }
    const int64_t consumed = (kind == 0 || del) ? elen : 0;
    if (total <= 0 || pos + gap + consumed >= N) {
      seg.append(bb, pos, N - pos);
      emit_segment(std::move(seg));
      break;
    }
    seg.append(bb, pos, gap);
    pos += gap;
    emit_segment(std::move(seg));
    seg.clear();
    const uint32_t first_alt = uint32_t(strs.size());
    if (kind == 0) {
      const char ref = bb[pos];
      char alt = ref;
      while (alt == ref) alt = kBases[rng.next() & 3];
      strs.push_back(std::string(1, ref));
      strs.push_back(std::string(1, alt));
      pos += 1;
      blocks.push_back({first_alt, 2, false});
    } else {
      std::string a;
      if (del) {
        a = bb.substr(pos, elen);
        pos += elen;
      } else {
        a.resize(elen);
        for (int64_t i = 0; i < elen; ++i) a[i] = kBases[rng.next() & 3];
      }
      strs.push_back(std::move(a));
      blocks.push_back({first_alt, 1, true});
    }
    const GenBlock& bk = blocks.back();
    for (uint32_t j = 0; j < bk.n_alts; ++j) {
      edges.push_back({prev_seg, bk.first_str + j});
      prev_tails.push_back(bk.first_str + j);
    }
    prev_skip_from_seg = bk.has_skip;
  }
  ds.n_s = int64_t(strs.size());
  ds.str_off.assign(ds.n_s + 1, 0);
  ds.labels.clear();
  for (int64_t s = 0; s < ds.n_s; ++s) {
    ds.labels += strs[s];
    ds.str_off[s + 1] = ds.labels.size();
  }
  std::sort(edges.begin(), edges.end());
  edges.erase(std::unique(edges.begin(), edges.end()), edges.end());
  ds.str_out_off.assign(ds.n_s + 1, 0);
  ds.str_out_dst.resize(edges.size());
  for (const auto& e : edges) ++ds.str_out_off[e.first + 1];
  for (int64_t s = 0; s < ds.n_s; ++s) ds.str_out_off[s + 1] += ds.str_out_off[s];
  for (size_t i = 0; i < edges.size(); ++i) ds.str_out_dst[i] = edges[i].second;
  ds.blocks = std::move(blocks);
  expand_char_graph(ds);
}

// Read sampling: a window of a random source-to-sink path with i.i.d. errors.
static void sample_read(const chsga_dataset& ds, const std::vector<uint64_t>& seg_prefix,
                        const std::vector<uint32_t>& seg_blocks, int64_t L, double sub, double ins, double del,
                        double nr, Rng& rng, std::string& out) {
  out.clear();
  if (L <= 0) return;
  const uint64_t total = seg_prefix.back();
  for (int attempt = 0; attempt < 64; ++attempt) {
    out.clear();
    // start: a uniform backbone-segment position, leaving room for the read
    const uint64_t limit = total > uint64_t(2 * L + 64) ? total - uint64_t(2 * L + 64) : 1;
    const uint64_t x = rng.below(attempt < 32 ? limit : total);
    size_t sb = std::upper_bound(seg_prefix.begin(), seg_prefix.end(), x) - seg_prefix.begin() - 1;
    size_t bi = seg_blocks[sb];
    uint64_t off = x - seg_prefix[sb];
    while (bi < ds.blocks.size() && int64_t(out.size()) < L) {
      const GenBlock& bk = ds.blocks[bi];
      uint32_t sid;
      if (bk.n_alts == 0) {
        sid = bk.first_str;
      } else {
        const uint32_t choices = bk.n_alts + (bk.has_skip ? 1 : 0);
        const uint32_t c = uint32_t(rng.below(choices));
        if (c >= bk.n_alts) {
          ++bi;
          off = 0;
          continue;
        }
        sid = bk.first_str + c;
      }
      const uint64_t a = ds.str_off[sid], b = ds.str_off[sid + 1];
      for (uint64_t q = a + off; q < b && int64_t(out.size()) < L; ++q) {
        const char g = ds.labels[q];
        const double u = rng.uniform();
        if (u < sub) {
          char c = g;
          while (c == g) c = kBases[rng.next() & 3];
          out.push_back(c);
        } else if (u < sub + ins) {
          out.push_back(kBases[rng.next() & 3]);
          if (int64_t(out.size()) < L) out.push_back(g);
        } else if (u < sub + ins + del) {
          // graph char skipped by the read
        } else {
          out.push_back(g);
        }
      }
      off = 0;
      ++bi;
    }
    if (int64_t(out.size()) >= L) break;
  }
  out.resize(std::min<size_t>(out.size(), size_t(L)));
  while (int64_t(out.size()) < L) out.push_back(kBases[rng.next() & 3]);
  if (nr > 0)
    for (char& c : out)
      if (rng.uniform() < nr) c = 'N';
}

// ---------------------------------------------------------------------------
// file formats
// ---------------------------------------------------------------------------
static std::string slurp(const char* path) {
  std::FILE* f = std::fopen(path, "rb");
  if (!f) throw std::runtime_error(std::string("cannot open ") + path);
  std::string s;
  char buf[1 << 16];
  size_t k;
  while ((k = std::fread(buf, 1, sizeof(buf), f)) > 0) s.append(buf, k);
  std::fclose(f);
  return s;
}

struct Tok {
  const std::string& s;
  size_t i = 0;
  explicit Tok(const std::string& str) : s(str) {}
  bool next(std::string& t) {
    while (i < s.size() && std::isspace(static_cast<unsigned char>(s[i]))) ++i;
    if (i >= s.size()) return false;
    const size_t a = i;
    while (i < s.size() && !std::isspace(static_cast<unsigned char>(s[i]))) ++i;
    t.assign(s, a, i - a);
    return true;
  }
  bool next_u64(uint64_t& v) {
    while (i < s.size() && std::isspace(static_cast<unsigned char>(s[i]))) ++i;
    if (i >= s.size() || !std::isdigit(static_cast<unsigned char>(s[i]))) return false;
    uint64_t x = 0;
    while (i < s.size() && std::isdigit(static_cast<unsigned char>(s[i]))) x = x * 10 + uint64_t(s[i++] - '0');
    if (i < s.size() && s[i] == '.') {  // weights such as "1.0"
      ++i;
      while (i < s.size() && std::isdigit(static_cast<unsigned char>(s[i]))) ++i;
    }
    v = x;
    return true;
  }
};

// AdjacencyGraph text (src/csr_to_adj.cpp:46-114): header, n, m, the first n
// offsets (the (n+1)-th is implied by m), m targets, then optionally m
// weights.  "WeightedAdjacencyGraph" is accepted as well.
static void read_adj(const char* path, int64_t& n, std::vector<uint64_t>& off, std::vector<uint32_t>& dst) {
  const std::string s = slurp(path);
  Tok t(s);
  std::string hdr;
  if (!t.next(hdr) || (hdr != "AdjacencyGraph" && hdr != "WeightedAdjacencyGraph"))
    throw std::runtime_error(std::string(path) + ": expected AdjacencyGraph header");
  uint64_t nn, m;
  if (!t.next_u64(nn) || !t.next_u64(m)) throw std::runtime_error(std::string(path) + ": bad n/m");
  n = int64_t(nn);
  off.assign(nn + 1, 0);
  for (uint64_t i = 0; i < nn; ++i)
    if (!t.next_u64(off[i])) throw std::runtime_error(std::string(path) + ": truncated offsets");
  off[nn] = m;
  dst.resize(m);
  for (uint64_t i = 0; i < m; ++i) {
    uint64_t v;
    if (!t.next_u64(v)) throw std::runtime_error(std::string(path) + ": truncated targets");
    if (v >= nn) throw std::runtime_error(std::string(path) + ": target out of range");
    dst[i] = uint32_t(v);
  }
  for (uint64_t i = 0; i < m; ++i) {
    uint64_t w;
    if (!t.next_u64(w)) break;  // weights are optional
    if (w != 1) throw std::runtime_error(std::string(path) + ": only unit edge weights are supported");
  }
  for (uint64_t i = 0; i < nn; ++i)
    if (off[i] > off[i + 1]) throw std::runtime_error(std::string(path) + ": offsets not non-decreasing");
}

static void write_adj(const char* path, int64_t n, const std::vector<uint64_t>& off, const std::vector<uint32_t>& dst) {
  std::FILE* f = std::fopen(path, "wb");
  if (!f) throw std::runtime_error(std::string("cannot write ") + path);
  std::fprintf(f, "AdjacencyGraph\n%lld\n%llu\n", (long long)n, (unsigned long long)dst.size());
  for (int64_t i = 0; i < n; ++i) std::fprintf(f, "%llu\n", (unsigned long long)off[i]);
  for (uint32_t v : dst) std::fprintf(f, "%u\n", v);
  for (size_t i = 0; i < dst.size(); ++i) std::fputs("1.0\n", f);
  std::fclose(f);
}

// Contraction-hierarchy file, this build's own versioned binary layout.  The
// fields are those of the reference's PchGraph that ConSGA reads (rank,
// level, layerOffset, ccOffset, offset/E, in_offset/in_E), plus the middle
// vertex of every shortcut.
static const char kChMagic[8] = {'C', 'H', 'S', 'G', 'A', 'C', 'H', '1'};

template <class T>
static void wvec(std::FILE* f, const std::vector<T>& v) {
  if (!v.empty() && std::fwrite(v.data(), sizeof(T), v.size(), f) != v.size())
    throw std::runtime_error("CH write failed");
}
template <class T>
static void rvec(std::FILE* f, std::vector<T>& v, uint64_t n) {
  v.resize(n);
  if (n && std::fread(v.data(), sizeof(T), n, f) != n) throw std::runtime_error("CH file truncated");
}

static void write_ch(const char* path, const ContractionHierarchy& ch) {
  std::FILE* f = std::fopen(path, "wb");
  if (!f) throw std::runtime_error(std::string("cannot write ") + path);
  std::fwrite(kChMagic, 1, 8, f);
  const uint64_t hdr[6] = {uint64_t(ch.n), uint64_t(ch.up_v.size()), uint64_t(ch.dn_v.size()),
                           uint64_t(ch.layer_off.size() - 1), uint64_t(ch.cc_off.size() - 1),
                           uint64_t(ch.has_mid ? 1 : 0)};
  std::fwrite(hdr, sizeof(uint64_t), 6, f);
  wvec(f, ch.rank);
  wvec(f, ch.level);
  wvec(f, ch.layer_off);
  wvec(f, ch.cc_off);
  wvec(f, ch.up_off);
  wvec(f, ch.up_v);
  wvec(f, ch.up_w);
  wvec(f, ch.dn_off);
  wvec(f, ch.dn_v);
  wvec(f, ch.dn_w);
  if (ch.has_mid) {
    wvec(f, ch.up_mid);
    wvec(f, ch.dn_mid);
  }
  std::fclose(f);
}

static void read_ch(const char* path, ContractionHierarchy& ch) {
  std::FILE* f = std::fopen(path, "rb");
  if (!f) throw std::runtime_error(std::string("cannot open ") + path);
  std::unique_ptr<std::FILE, int (*)(std::FILE*)> guard(f, std::fclose);
  char mg[8];
  if (std::fread(mg, 1, 8, f) != 8 || std::memcmp(mg, kChMagic, 8) != 0)
    throw std::runtime_error(std::string(path) + ": not a CHSGACH1 contraction-hierarchy file");
  uint64_t hdr[6];
  if (std::fread(hdr, sizeof(uint64_t), 6, f) != 6) throw std::runtime_error("CH header truncated");
  ch = ContractionHierarchy();
  ch.n = int64_t(hdr[0]);
  rvec(f, ch.rank, hdr[0]);
  rvec(f, ch.level, hdr[0]);
  rvec(f, ch.layer_off, hdr[3] + 1);
  rvec(f, ch.cc_off, hdr[4] + 1);
  rvec(f, ch.up_off, hdr[0] + 1);
  rvec(f, ch.up_v, hdr[1]);
  rvec(f, ch.up_w, hdr[1]);
  rvec(f, ch.dn_off, hdr[0] + 1);
  rvec(f, ch.dn_v, hdr[2]);
  rvec(f, ch.dn_w, hdr[2]);
  ch.has_mid = hdr[5] != 0;
  if (ch.has_mid) {
    rvec(f, ch.up_mid, hdr[1]);
    rvec(f, ch.dn_mid, hdr[2]);
  }
  if (ch.up_off.back() != hdr[1] || ch.dn_off.back() != hdr[2]) throw std::runtime_error("CH edge offsets");
}

// Reads: FASTA / FASTQ (gzip or plain) or one sequence per line; upper-cased
// as src/load_reads.hpp:20-24 does.
static void load_reads(const char* path, int64_t cap, chsga_readset& rs) {
  gzFile g = gzopen(path, "rb");
  if (!g) throw std::runtime_error(std::string("Cannot open reads file: ") + path);
  std::string s;
  char buf[1 << 16];
  int k;
  while ((k = gzread(g, buf, sizeof(buf))) > 0) s.append(buf, size_t(k));
  gzclose(g);
  rs.data.clear();
  rs.off.assign(1, 0);
  size_t i = 0;
  auto line = [&](std::string& out) -> bool {
    if (i >= s.size()) return false;
    size_t j = s.find('\n', i);
    if (j == std::string::npos) j = s.size();
    out.assign(s, i, j - i);
    if (!out.empty() && out.back() == '\r') out.pop_back();
    i = j + 1;
    return true;
  };
  auto push = [&](const std::string& seq) {
    for (char c : seq) {
      if (std::isspace(static_cast<unsigned char>(c))) continue;
      rs.data.push_back(char(std::toupper(static_cast<unsigned char>(c))));
    }
    rs.off.push_back(int64_t(rs.data.size()));
  };
  std::string ln, seq;
  size_t first = s.find_first_not_of(" \t\r\n");
  const char lead = first == std::string::npos ? 0 : s[first];
  i = first == std::string::npos ? s.size() : first;
  if (lead == '>') {
    bool have = false;
    while (line(ln)) {
      if (!ln.empty() && ln[0] == '>') {
        if (have) {
          push(seq);
          if (cap > 0 && int64_t(rs.off.size()) - 1 >= cap) break;
        }
        seq.clear();
        have = true;
      } else {
        seq += ln;
      }
    }
    if (have && !(cap > 0 && int64_t(rs.off.size()) - 1 >= cap)) push(seq);
  } else if (lead == '@') {
    while (line(ln)) {
      if (ln.empty()) continue;
      if (ln[0] != '@') throw std::runtime_error(std::string(path) + ": malformed FASTQ record");
      seq.clear();
      while (line(ln) && (ln.empty() || ln[0] != '+')) seq += ln;
      size_t ql = 0;
      while (ql < seq.size() && line(ln)) ql += ln.size();
      push(seq);
      if (cap > 0 && int64_t(rs.off.size()) - 1 >= cap) break;
    }
  } else {
    while (line(ln)) {
      if (ln.empty()) continue;
      push(ln);
      if (cap > 0 && int64_t(rs.off.size()) - 1 >= cap) break;
    }
  }
  if (rs.off.size() == 1) throw std::runtime_error(std::string("No reads loaded from: ") + path);
}

}  // namespace chsga

using namespace chsga;

#define CHSGA_GUARD(...)                        \
  try {                                          \
    __VA_ARGS__;                                        \
    return CHSGA_OK;                             \
  } catch (const std::bad_alloc&) {              \
    set_error("host allocation failed");         \
    return CHSGA_ENOMEM;                         \
  } catch (const std::exception& e) {            \
    set_error(e.what());                         \
    return CHSGA_EIO;                            \
  }

extern "C" {

const char* chsga_host_last_error(void) { return g_err.c_str(); }

int chsga_host_generate(const chsga_graph_params* p, chsga_dataset** out) {
  if (!p || !out) {
    set_error("null argument");
    return CHSGA_EINVAL;
  }
  CHSGA_GUARD({
    auto ds = std::make_unique<chsga_dataset>();
    generate(*p, *ds);
    build_ch(*ds, p->seed);
    validate_dataset(*ds);
    *out = ds.release();
  })
}

int chsga_host_build_ch(chsga_dataset* ds, uint64_t seed) {
  if (!ds) {
    set_error("null dataset");
    return CHSGA_EINVAL;
  }
  CHSGA_GUARD({
    build_ch(*ds, seed);
    validate_dataset(*ds);
  })
}

int chsga_host_load(const char* char_adj, const char* str_adj, const char* ch_path, const char* labels,
                    const char* offsets, chsga_dataset** out) {
  if (!char_adj || !str_adj || !labels || !offsets || !out) {
    set_error("missing input path");
    return CHSGA_EINVAL;
  }
  CHSGA_GUARD({
    auto ds = std::make_unique<chsga_dataset>();
    {
      // labels: every byte except '\n' (src/ConSGA.cpp:166-169)
      const std::string s = slurp(labels);
      ds->labels.reserve(s.size());
      for (char c : s)
        if (c != '\n') ds->labels.push_back(c);
    }
    {
      // label offsets: whitespace-separated integers (src/ConSGA.cpp:182-184)
      const std::string s = slurp(offsets);
      Tok t(s);
      uint64_t v;
      while (t.next_u64(v)) ds->str_off.push_back(v);
    }
    int64_t nc = 0;
    read_adj(char_adj, nc, ds->char_out_off, ds->char_out_dst);
    ds->V = nc;
    build_in_csr(ds->V, ds->char_out_off, ds->char_out_dst, ds->in_off, ds->in_src);
    read_adj(str_adj, ds->n_s, ds->str_out_off, ds->str_out_dst);
    if (ch_path)
      read_ch(ch_path, ds->ch);
    else
      build_ch(*ds, 1);
    validate_dataset(*ds);
    *out = ds.release();
  })
}

int chsga_host_save(const chsga_dataset* ds, const char* prefix) {
  if (!ds || !prefix) {
    set_error("null argument");
    return CHSGA_EINVAL;
  }
  CHSGA_GUARD({
    const std::string p(prefix);
    write_adj((p + ".char.adj").c_str(), ds->V, ds->char_out_off, ds->char_out_dst);
    write_adj((p + ".str.adj").c_str(), ds->n_s, ds->str_out_off, ds->str_out_dst);
    write_ch((p + ".ch").c_str(), ds->ch);
    {
      std::ofstream f(p + ".labels", std::ios::binary);
      f << ds->labels;
      if (!f) throw std::runtime_error("cannot write labels");
    }
    {
      std::FILE* f = std::fopen((p + ".label_offsets").c_str(), "wb");
      if (!f) throw std::runtime_error("cannot write label offsets");
      for (int64_t s = 0; s < ds->n_s; ++s) std::fprintf(f, "%llu ", (unsigned long long)ds->str_off[s]);
      std::fprintf(f, "%llu\n", (unsigned long long)ds->str_off[ds->n_s]);
      std::fclose(f);
    }
  })
}

// ---------------------------------------------------------------------------
// converters: GFA and the string-labelled CSR text of handlegraph_to_csr
// ---------------------------------------------------------------------------

// Complement of one sequence byte as a handle graph reports a reverse
// strand: IUPAC codes complemented, case kept, other bytes unchanged.
static char complement(char c) {
  switch (c) {
    case 'A': return 'T';
    case 'C': return 'G';
    case 'G': return 'C';
    case 'T': return 'A';
    case 'U': return 'A';
    case 'R': return 'Y';
    case 'Y': return 'R';
    case 'K': return 'M';
    case 'M': return 'K';
    case 'B': return 'V';
    case 'V': return 'B';
    case 'D': return 'H';
    case 'H': return 'D';
    case 'a': return 't';
    case 'c': return 'g';
    case 'g': return 'c';
    case 't': return 'a';
    case 'u': return 'a';
    case 'r': return 'y';
    case 'y': return 'r';
    case 'k': return 'm';
    case 'm': return 'k';
    case 'b': return 'v';
    case 'v': return 'b';
    case 'd': return 'h';
    case 'h': return 'd';
    default: return c;  // N, S, W and others are their own complement
  }
}

// Dataset from a string graph (sequences, out-lists): labels and offsets,
// the char expansion, the CH.
static void dataset_from_strings(const std::vector<std::string>& seq, std::vector<std::vector<uint32_t>>& out,
                                 chsga_dataset& ds) {
  const int64_t n = int64_t(seq.size());
  if (n == 0) throw std::runtime_error("empty graph");
  ds.n_s = n;
  ds.str_off.assign(size_t(n) + 1, 0);
  ds.labels.clear();
  for (int64_t s = 0; s < n; ++s) {
    if (seq[s].empty()) throw std::runtime_error("string vertex " + std::to_string(s) + " has an empty sequence");
    ds.labels += seq[s];
    ds.str_off[s + 1] = ds.labels.size();
  }
  ds.str_out_off.assign(size_t(n) + 1, 0);
  ds.str_out_dst.clear();
  for (int64_t s = 0; s < n; ++s) {
    std::sort(out[s].begin(), out[s].end());
    out[s].erase(std::unique(out[s].begin(), out[s].end()), out[s].end());
    for (uint32_t t : out[s]) {
      if (int64_t(t) >= n) throw std::runtime_error("edge to vertex " + std::to_string(t) + " out of range");
      ds.str_out_dst.push_back(t);
    }
    ds.str_out_off[s + 1] = ds.str_out_dst.size();
  }
  expand_char_graph(ds);
  build_ch(ds, 1);
  validate_dataset(ds);
}

// GFA 1 (S and L lines, overlaps * or 0M), listed as handlegraph_to_csr
// lists a handle graph (src/handlegraph_to_csr.cpp:39-110): segments ordered
// by id (the numeric name, or file order when a name is not a number),
// segment k's forward strand as string vertex 2k and its reverse complement
// as 2k + 1; a link a -> b also gives flip(b) -> flip(a).
static void load_gfa(const char* path, chsga_dataset& ds) {
  const std::string text = slurp(path);
  std::vector<std::string> name, seq;
  std::vector<std::array<std::string, 4>> links;
  std::istringstream in(text);
  std::string line;
  int64_t ln = 0;
  while (std::getline(in, line)) {
    ++ln;
    if (!line.empty() && line.back() == '\r') line.pop_back();
    if (line.empty() || (line[0] != 'S' && line[0] != 'L')) continue;
    std::vector<std::string> f;
    for (size_t a = 0;;) {
      const size_t b = line.find('\t', a);
      f.push_back(line.substr(a, b == std::string::npos ? std::string::npos : b - a));
      if (b == std::string::npos) break;
      a = b + 1;
    }
    auto bad = [&](const std::string& why) {
      throw std::runtime_error(std::string(path) + ":" + std::to_string(ln) + ": " + why);
    };
    if (f[0] == "S") {
      if (f.size() < 3) bad("S line needs a name and a sequence");
      if (f[2] == "*" || f[2].empty()) bad("segment " + f[1] + " has no sequence");
      name.push_back(f[1]);
      seq.push_back(f[2]);
    } else if (f[0] == "L") {
      if (f.size() < 5) bad("L line needs from, orientation, to, orientation");
      if ((f[2] != "+" && f[2] != "-") || (f[4] != "+" && f[4] != "-")) bad("orientation must be + or -");
      if (f.size() > 5 && f[5] != "*" && f[5] != "0M") bad("only blunt links (overlap 0M or *) are supported");
      links.push_back({f[1], f[2], f[3], f[4]});
    }
  }
  const size_t k = name.size();
  std::vector<size_t> order(k);
  for (size_t i = 0; i < k; ++i) order[i] = i;
  bool numeric = true;
  for (const std::string& s : name)
    numeric = numeric && !s.empty() && s.size() < 19 && std::all_of(s.begin(), s.end(), [](char c) {
      return c >= '0' && c <= '9';
    });
  if (numeric)
    std::stable_sort(order.begin(), order.end(),
                     [&](size_t a, size_t b) { return std::stoull(name[a]) < std::stoull(name[b]); });
  std::unordered_map<std::string, uint32_t> idx;  // name -> position in id order
  for (size_t i = 0; i < k; ++i)
    if (!idx.emplace(name[order[i]], uint32_t(i)).second) throw std::runtime_error("duplicate segment " + name[order[i]]);
  std::vector<std::string> sv(2 * k);
  for (size_t i = 0; i < k; ++i) {
    const std::string& s = seq[order[i]];
    sv[2 * i] = s;
    sv[2 * i + 1].assign(s.rbegin(), s.rend());
    for (char& c : sv[2 * i + 1]) c = complement(c);
  }
  std::vector<std::vector<uint32_t>> out(2 * k);
  for (const auto& l : links) {
    const auto a = idx.find(l[0]), b = idx.find(l[2]);
    if (a == idx.end() || b == idx.end()) throw std::runtime_error("link names an unknown segment");
    const uint32_t h1 = 2 * a->second + (l[1] == "-" ? 1u : 0u);
    const uint32_t h2 = 2 * b->second + (l[3] == "-" ? 1u : 0u);
    out[h1].push_back(h2);
    out[h2 ^ 1u].push_back(h1 ^ 1u);
  }
  dataset_from_strings(sv, out, ds);
}

// The string-labelled CSR text handlegraph_to_csr writes (write_csr,
// src/handlegraph_to_csr.cpp:112-150): n, m, n labels, n + 1 offsets, m
// destinations, whitespace separated.
static void load_string_csr(const char* path, chsga_dataset& ds) {
  const std::string text = slurp(path);
  Tok t(text);
  uint64_t n = 0, m = 0;
  if (!t.next_u64(n) || !t.next_u64(m)) throw std::runtime_error(std::string(path) + ": expected n and m");
  std::vector<std::string> seq(n);
  for (uint64_t i = 0; i < n; ++i)
    if (!t.next(seq[i])) throw std::runtime_error(std::string(path) + ": truncated labels");
  std::vector<uint64_t> off(n + 1);
  for (uint64_t i = 0; i <= n; ++i)
    if (!t.next_u64(off[i])) throw std::runtime_error(std::string(path) + ": truncated offsets");
  if (off[0] != 0 || off[n] != m) throw std::runtime_error(std::string(path) + ": offsets must run from 0 to m");
  std::vector<std::vector<uint32_t>> out(n);
  for (uint64_t i = 0; i < n; ++i) {
    if (off[i + 1] < off[i]) throw std::runtime_error(std::string(path) + ": decreasing offsets");
    for (uint64_t e = off[i]; e < off[i + 1]; ++e) {
      uint64_t d;
      if (!t.next_u64(d)) throw std::runtime_error(std::string(path) + ": truncated adjacency");
      out[i].push_back(uint32_t(d));
    }
  }
  dataset_from_strings(seq, out, ds);
}

// A CSR text in that format: n, m, the labels, n + 1 offsets, m destinations.
static void write_csr_text(const std::string& path, int64_t n, const std::function<void(std::FILE*, int64_t)>& label,
                           const std::vector<uint64_t>& off, const std::vector<uint32_t>& dst) {
  std::FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) throw std::runtime_error("cannot write " + path);
  std::fprintf(f, "%lld\n%llu\n", (long long)n, (unsigned long long)dst.size());
  for (int64_t v = 0; v < n; ++v) {
    label(f, v);
    std::fputc(' ', f);
  }
  std::fputc('\n', f);
  for (int64_t v = 0; v <= n; ++v) std::fprintf(f, "%llu ", (unsigned long long)off[v]);
  std::fputc('\n', f);
  for (uint32_t d : dst) std::fprintf(f, "%u ", d);
  std::fputc('\n', f);
  if (std::fclose(f) != 0) throw std::runtime_error("cannot write " + path);
}

int chsga_host_load_gfa(const char* path, chsga_dataset** out) {
  if (!path || !out) {
    set_error("null argument");
    return CHSGA_EINVAL;
  }
  CHSGA_GUARD({
    auto ds = std::make_unique<chsga_dataset>();
    load_gfa(path, *ds);
    *out = ds.release();
  })
}

int chsga_host_load_string_csr(const char* path, chsga_dataset** out) {
  if (!path || !out) {
    set_error("null argument");
    return CHSGA_EINVAL;
  }
  CHSGA_GUARD({
    auto ds = std::make_unique<chsga_dataset>();
    load_string_csr(path, *ds);
    *out = ds.release();
  })
}

// Reversal for PaSGAL's phase 1-R (src/PaSGALoverrides/align.hpp:146-245
// walks chars V-1..0 over out-edges): string s becomes n_s - 1 - s with its
// label reversed, edge u -> v becomes v' -> u', so char j becomes V - 1 - j;
// then the char expansion and a CH of the reversed string graph.
int chsga_host_reverse(const chsga_dataset* ds, uint64_t seed, chsga_dataset** out) {
  if (!ds || !out) {
    set_error("null argument");
    return CHSGA_EINVAL;
  }
  CHSGA_GUARD({
    const int64_t n = ds->n_s;
    auto r = std::make_unique<chsga_dataset>();
    r->n_s = n;
    r->str_off.assign(n + 1, 0);
    r->labels.reserve(ds->labels.size());
    for (int64_t t = 0; t < n; ++t) {
      const int64_t s = n - 1 - t;
      const uint64_t a = ds->str_off[s], b = ds->str_off[s + 1];
      for (uint64_t c = b; c > a; --c) r->labels.push_back(ds->labels[c - 1]);
      r->str_off[t + 1] = r->labels.size();
    }
    std::vector<std::vector<uint32_t>> adj(n);
    for (int64_t u = 0; u < n; ++u)
      for (uint64_t e = ds->str_out_off[u]; e < ds->str_out_off[u + 1]; ++e)
        adj[n - 1 - ds->str_out_dst[e]].push_back(uint32_t(n - 1 - u));
    r->str_out_off.assign(n + 1, 0);
    for (int64_t t = 0; t < n; ++t) {
      std::sort(adj[t].begin(), adj[t].end());
      r->str_out_dst.insert(r->str_out_dst.end(), adj[t].begin(), adj[t].end());
      r->str_out_off[t + 1] = r->str_out_dst.size();
    }
    expand_char_graph(*r);
    if (r->in_src.size() != ds->in_src.size())
      throw std::runtime_error("the char graph is not the expansion of the string graph: cannot reverse it");
    build_ch(*r, seed);
    validate_dataset(*r);
    *out = r.release();
  })
}

int chsga_host_write_csr(const chsga_dataset* ds, const char* string_csr, const char* char_csr) {
  if (!ds || !string_csr || !char_csr) {
    set_error("null argument");
    return CHSGA_EINVAL;
  }
  CHSGA_GUARD({
    write_csr_text(
        string_csr, ds->n_s,
        [&](std::FILE* f, int64_t s) {
          std::fwrite(ds->labels.data() + ds->str_off[s], 1, size_t(ds->str_off[s + 1] - ds->str_off[s]), f);
        },
        ds->str_out_off, ds->str_out_dst);
    write_csr_text(
        char_csr, ds->V, [&](std::FILE* f, int64_t p) { std::fputc(ds->labels[size_t(p)], f); }, ds->char_out_off,
        ds->char_out_dst);
  })
}

int chsga_host_save_ch(const chsga_dataset* ds, const char* path) {
  if (!ds || !path) {
    set_error("null argument");
    return CHSGA_EINVAL;
  }
  CHSGA_GUARD(write_ch(path, ds->ch))
}

int chsga_host_build_ch_file(const char* str_adj, const char* label_offsets, const char* out_ch, uint64_t seed) {
  if (!str_adj || !out_ch) {
    set_error("missing input path");
    return CHSGA_EINVAL;
  }
  CHSGA_GUARD({
    chsga_dataset ds;
    read_adj(str_adj, ds.n_s, ds.str_out_off, ds.str_out_dst);
    if (label_offsets) {
      const std::string s = slurp(label_offsets);
      Tok t(s);
      uint64_t v;
      while (t.next_u64(v)) ds.str_off.push_back(v);
      if (int64_t(ds.str_off.size()) != ds.n_s + 1)
        throw std::runtime_error("label offsets: expected n+1 values for the string graph");
    } else {
      ds.str_off.resize(ds.n_s + 1);
      for (int64_t i = 0; i <= ds.n_s; ++i) ds.str_off[i] = uint64_t(i);
    }
    build_ch(ds, seed);
    write_ch(out_ch, ds.ch);
  })
}

int chsga_host_describe(const chsga_dataset* ds, int32_t alphabet, chsga_graph_desc* g) {
  if (!ds || !g) {
    set_error("null argument");
    return CHSGA_EINVAL;
  }
  std::memset(g, 0, sizeof(*g));
  g->V = ds->V;
  g->labels = ds->labels.data();
  g->in_off = ds->in_off.data();
  g->in_src = ds->in_src.data();
  g->E = int64_t(ds->in_src.size());
  g->n_s = ds->n_s;
  g->str_off = ds->str_off.data();
  g->str_out_off = ds->str_out_off.data();
  g->str_out_dst = ds->str_out_dst.data();
  g->m_str = int64_t(ds->str_out_dst.size());
  const ContractionHierarchy& ch = ds->ch;
  g->num_layers = int64_t(ch.layer_off.size()) - 1;
  g->layer_off = ch.layer_off.data();
  g->num_cc = int64_t(ch.cc_off.size()) - 1;
  g->cc_off = ch.cc_off.data();
  g->rank = ch.rank.data();
  g->up_off = ch.up_off.data();
  g->up_v = ch.up_v.data();
  g->up_w = ch.up_w.data();
  g->m_up = int64_t(ch.up_v.size());
  g->dn_off = ch.dn_off.data();
  g->dn_v = ch.dn_v.data();
  g->dn_w = ch.dn_w.data();
  g->m_dn = int64_t(ch.dn_v.size());
  g->up_mid = ch.has_mid ? ch.up_mid.data() : nullptr;
  g->dn_mid = ch.has_mid ? ch.dn_mid.data() : nullptr;
  g->alphabet = alphabet;
  g->match_score = 0;
  g->subst_score = 1;
  g->del_score = 1;
  g->ins_score = 1;
  return CHSGA_OK;
}

int chsga_host_info(const chsga_dataset* ds, chsga_dataset_info* out) {
  if (!ds || !out) {
    set_error("null argument");
    return CHSGA_EINVAL;
  }
  out->V = ds->V;
  out->E = int64_t(ds->in_src.size());
  out->n_s = ds->n_s;
  out->m_str = int64_t(ds->str_out_dst.size());
  out->m_up = int64_t(ds->ch.up_v.size());
  out->m_dn = int64_t(ds->ch.dn_v.size());
  out->num_layers = int64_t(ds->ch.layer_off.size()) - 1;
  out->num_cc = int64_t(ds->ch.cc_off.size()) - 1;
  out->max_level = 0;
  for (uint32_t l : ds->ch.level) out->max_level = std::max<int64_t>(out->max_level, l);
  out->max_str_len = 0;
  for (int64_t s = 0; s < ds->n_s; ++s)
    out->max_str_len = std::max<int64_t>(out->max_str_len, int64_t(ds->str_off[s + 1] - ds->str_off[s]));
  return CHSGA_OK;
}

void chsga_host_free(chsga_dataset* ds) { delete ds; }

int chsga_host_gen_reads(const chsga_dataset* ds, const chsga_read_params* p, chsga_readset** out) {
  if (!ds || !p || !out) {
    set_error("null argument");
    return CHSGA_EINVAL;
  }
  if (ds->blocks.empty()) {
    set_error("read generation needs a generated dataset (block structure)");
    return CHSGA_EINVAL;
  }
  CHSGA_GUARD({
    std::vector<uint64_t> seg_prefix{0};
    std::vector<uint32_t> seg_blocks;
    for (size_t b = 0; b < ds->blocks.size(); ++b)
      if (ds->blocks[b].n_alts == 0) {
        const uint32_t s = ds->blocks[b].first_str;
        seg_prefix.push_back(seg_prefix.back() + (ds->str_off[s + 1] - ds->str_off[s]));
        seg_blocks.push_back(uint32_t(b));
      }
    auto rs = std::make_unique<chsga_readset>();
    std::string r;
    for (int64_t i = 0; i < p->n_reads; ++i) {
      Rng rng(p->seed * 0x9E3779B97F4A7C15ull + uint64_t(i) * 0xD1B54A32D192ED03ull + 17);
      int64_t L;
      double es = 1.0;
      if (p->len_mode == 0) {
        L = int64_t(p->len_a);
      } else if (p->len_mode == 1) {
        L = int64_t(std::llround(std::exp(std::log(p->len_a) + p->len_b * rng.normal())));
        L = std::min(std::max(L, p->len_min), p->len_max);
      } else {
        if (rng.uniform() < p->mix_frac) {
          L = rng.range(p->len_min, p->len_max);
          es = p->err_scale_a;
        } else {
          L = rng.range(p->len2_min, p->len2_max);
          es = p->err_scale_b;
        }
      }
      sample_read(*ds, seg_prefix, seg_blocks, L, p->sub_rate * es, p->ins_rate * es, p->del_rate * es, p->n_rate,
                  rng, r);
      rs->data += r;
      rs->off.push_back(int64_t(rs->data.size()));
    }
    *out = rs.release();
  })
}

int chsga_host_load_reads(const char* path, int64_t cap, chsga_readset** out) {
  if (!path || !out) {
    set_error("null argument");
    return CHSGA_EINVAL;
  }
  CHSGA_GUARD({
    auto rs = std::make_unique<chsga_readset>();
    load_reads(path, cap, *rs);
    *out = rs.release();
  })
}

int chsga_host_save_reads(const chsga_readset* rs, const char* path) {
  if (!rs || !path) {
    set_error("null argument");
    return CHSGA_EINVAL;
  }
  CHSGA_GUARD({
    std::FILE* f = std::fopen(path, "wb");
    if (!f) throw std::runtime_error(std::string("cannot write ") + path);
    for (size_t i = 0; i + 1 < rs->off.size(); ++i) {
      const int64_t a = rs->off[i], b = rs->off[i + 1];
      std::fprintf(f, "@read_%zu\n", i);
      std::fwrite(rs->data.data() + a, 1, size_t(b - a), f);
      std::fputs("\n+\n", f);
      for (int64_t q = a; q < b; ++q) std::fputc('I', f);
      std::fputc('\n', f);
    }
    std::fclose(f);
  })
}

// The verification of src/ConSGA_traceback.cpp:184-225 (path rescored with
// the edit costs, last consuming row = |r| - 1), plus the structure the
// rescoring assumes: the path opens with the row-0 MATCHSUB that
// align_read_with_traceback appends (src/traceback.hpp:1020-1027), and every
// step moves along a char edge (MATCHSUB, INSERT) or stays on its char
// (DELETE) with the row advancing as the op says.
int chsga_host_verify_path(const chsga_graph_desc* g, const char* read, int64_t len, const chsga_align_op* ops,
                           int64_t n_ops, int32_t score, int64_t* bad_index) {
  if (!g || !read || len <= 0 || !g->labels || !g->in_off || !g->in_src) {
    set_error("null argument");
    return CHSGA_EINVAL;
  }
  int64_t bad = 0;
  int code = 0;
  auto edge = [&](uint32_t u, uint32_t v) {
    const uint32_t* a = g->in_src + g->in_off[v];
    const uint32_t* b = g->in_src + g->in_off[v + 1];
    return std::find(a, b, u) != b;
  };
  auto sig = [&](char a, char b) {
    const bool m = g->alphabet == CHSGA_ALPHABET_MATCH_N ? (a == 'N' || b == 'N' || a == b) : a == b;
    return m ? g->match_score : g->subst_score;
  };
  int64_t cost = 0, consumed = -1;
  if (n_ops <= 0 || !ops) {
    code = 1;
  } else if (ops[0].row != 0 || ops[0].op != CHSGA_OP_MATCHSUB) {
    code = 2;
  } else {
    for (int64_t j = 0; j < n_ops && !code; ++j) {
      const chsga_align_op& o = ops[j];
      bad = j;
      if (int64_t(o.vertex) >= g->V || o.row < 0 || o.row >= len || o.op > CHSGA_OP_INSERT) {
        code = 4;
        break;
      }
      if (j > 0) {
        const chsga_align_op& p = ops[j - 1];
        const bool step = o.op == CHSGA_OP_DELETE     ? o.row == p.row + 1 && o.vertex == p.vertex
                          : o.op == CHSGA_OP_MATCHSUB ? o.row == p.row + 1 && edge(p.vertex, o.vertex)
                                                      : o.row == p.row && edge(p.vertex, o.vertex);
        if (!step) {
          code = 3;
          break;
        }
      }
      if (o.op == CHSGA_OP_INSERT) {
        cost += g->ins_score;
      } else {
        cost += o.op == CHSGA_OP_DELETE ? g->del_score : sig(g->labels[o.vertex], read[o.row]);
        consumed = o.row;
      }
    }
    if (!code && consumed != len - 1) code = 5;
    if (!code && cost != score) code = 6;
    if (code >= 5) bad = n_ops - 1;
  }
  if (bad_index) *bad_index = code ? bad : -1;
  return code;
}

int64_t chsga_reads_count(const chsga_readset* rs) { return rs ? int64_t(rs->off.size()) - 1 : 0; }
const char* chsga_reads_data(const chsga_readset* rs) { return rs ? rs->data.data() : nullptr; }
const int64_t* chsga_reads_offsets(const chsga_readset* rs) { return rs ? rs->off.data() : nullptr; }
void chsga_reads_free(chsga_readset* rs) { delete rs; }
const char* chsga_ds_labels(const chsga_dataset* ds) { return ds ? ds->labels.data() : nullptr; }
const uint64_t* chsga_ds_str_off(const chsga_dataset* ds) { return ds ? ds->str_off.data() : nullptr; }

}  // extern "C"
