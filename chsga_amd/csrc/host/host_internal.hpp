// Host-side data model shared by the format readers, the generator, the CH
// builder and the drop-in drivers.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "chsga/host.h"

namespace chsga {

// Deterministic 64-bit generator (xoshiro256**, seeded through splitmix64),
// so that every seeded graph and read set is the same on every machine.
struct Rng {
  uint64_t s[4];
  explicit Rng(uint64_t seed) {
    uint64_t z = seed;
    for (int i = 0; i < 4; ++i) {
      z += 0x9E3779B97F4A7C15ull;
      uint64_t x = z;
      x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
      x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
      s[i] = x ^ (x >> 31);
    }
  }
  static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t next() {
    const uint64_t r = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return r;
  }
  double uniform() { return double(next() >> 11) * (1.0 / 9007199254740992.0); }
  uint64_t below(uint64_t n) { return n ? next() % n : 0; }
  int64_t range(int64_t lo, int64_t hi) { return lo + int64_t(below(uint64_t(hi - lo + 1))); }
  double normal();
  int64_t geometric(double p);  // number of failures before the first success
};

// One CH edge: other endpoint, weight, contracted middle vertex (-1: original).
struct ChEdge {
  uint32_t v;
  int32_t w;
  int32_t mid;
};

// A contraction hierarchy in CH-id space, laid out like the reference's
// PchGraph fields used by ConSGA (src/align.hpp:294-343, ConSGA.cpp:203-206).
struct ContractionHierarchy {
  int64_t n = 0;
  std::vector<uint32_t> rank;   // og string id -> CH id
  std::vector<uint32_t> level;  // by CH id
  std::vector<uint64_t> layer_off;
  std::vector<uint64_t> cc_off;
  std::vector<uint64_t> up_off;
  std::vector<uint32_t> up_v;
  std::vector<int32_t> up_w, up_mid;
  std::vector<uint64_t> dn_off;
  std::vector<uint32_t> dn_v;
  std::vector<int32_t> dn_w, dn_mid;
  bool has_mid = false;
};

// Block structure of a generated chain-of-bubbles graph (used to sample reads
// along random paths).  Blocks alternate segment, bubble, segment, ...
struct GenBlock {
  uint32_t first_str;  // string id of the segment, or of the first alternative
  uint32_t n_alts;     // 0 for a segment
  bool has_skip;
};

}  // namespace chsga

struct chsga_dataset {
  // string graph
  int64_t n_s = 0;
  std::vector<uint64_t> str_off;  // label offsets, n_s + 1
  std::vector<uint64_t> str_out_off;
  std::vector<uint32_t> str_out_dst;
  std::string labels;  // V chars
  // char graph
  int64_t V = 0;
  std::vector<uint64_t> char_out_off;
  std::vector<uint32_t> char_out_dst;
  std::vector<uint64_t> in_off;
  std::vector<uint32_t> in_src;
  chsga::ContractionHierarchy ch;
  std::vector<chsga::GenBlock> blocks;  // generated datasets only
};

struct chsga_readset {
  std::string data;
  std::vector<int64_t> off{0};
};

namespace chsga {

void set_error(const std::string& msg);
// char graph from the string graph, after src/handlegraph_to_csr.cpp:165-247
void expand_char_graph(chsga_dataset& ds);
// in-CSR with in-neighbours in ascending order, after src/align.hpp:220-242
void build_in_csr(int64_t n, const std::vector<uint64_t>& off, const std::vector<uint32_t>& dst,
                  std::vector<uint64_t>& in_off, std::vector<uint32_t>& in_src);
void build_ch(chsga_dataset& ds, uint64_t seed);
void validate_dataset(const chsga_dataset& ds);

}  // namespace chsga
