// Shared code of the drop-in drivers: input loading through the host C ABI,
// the engine context, and the reference's report helpers.
#pragma once

#include <getopt.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <iomanip>
#include <iostream>
#include <string>
#include <unordered_map>
#include <vector>

#include "chsga/abi.h"
#include "chsga/host.h"

#ifndef CHSGA_ALPHABET
#define CHSGA_ALPHABET 5
#endif

namespace drv {

inline double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

[[noreturn]] inline void die(const std::string& what, const char* detail) {
  std::cerr << what << (detail ? detail : "") << "\n";
  std::exit(EXIT_FAILURE);
}

struct Reads {
  chsga_readset* rs = nullptr;
  std::vector<const char*> ptr;
  std::vector<int64_t> len;
  ~Reads() { chsga_reads_free(rs); }
  size_t size() const { return len.size(); }
  std::string str(size_t i) const { return std::string(ptr[i], size_t(len[i])); }
};

// load_reads (src/load_reads.hpp:7-40); cap <= 0 means no cap (the reference
// stops after 5 reads, src/load_reads.hpp:5; pass -n 5 to mimic that)
inline void load_reads(const char* path, int64_t cap, Reads& r) {
  if (chsga_host_load_reads(path, cap, &r.rs) != CHSGA_OK) {
    std::cerr << "terminate called after throwing an instance of 'std::runtime_error'\n  what():  "
              << chsga_host_last_error() << "\n";
    std::exit(EXIT_FAILURE);
  }
  const int64_t n = chsga_reads_count(r.rs);
  const char* d = chsga_reads_data(r.rs);
  const int64_t* off = chsga_reads_offsets(r.rs);
  for (int64_t i = 0; i < n; ++i) {
    r.ptr.push_back(d + off[i]);
    r.len.push_back(off[i + 1] - off[i]);
  }
}

// t-based CI of src/ParSGAoverrides/csr_align.hpp:52-99 (mean and sample sd
// over reads 1..N-1, two-sided 95% t critical value, df capped at 20)
inline double t_critical_95(int df) {
  static const double t95[] = {0,     12.706, 4.303, 3.182, 2.776, 2.571, 2.447, 2.365, 2.306, 2.262, 2.228,
                               2.201, 2.179,  2.160, 2.145, 2.131, 2.120, 2.110, 2.101, 2.093, 2.086};
  if (df <= 0) return 0.0;
  if (df >= 20) return t95[20];
  return t95[df];
}

inline void print_mean_with_moe_t(const std::string& label, const std::vector<double>& v) {
  if (v.size() <= 1) {
    std::cout << label << ": insufficient data (need >1 after skipping first)\n";
    return;
  }
  const size_t n = v.size() - 1;
  double m = 0;
  for (size_t i = 1; i < v.size(); ++i) m += v[i];
  m /= double(n);
  double s = 0;
  if (v.size() > 2) {
    double acc = 0;
    for (size_t i = 1; i < v.size(); ++i) acc += (v[i] - m) * (v[i] - m);
    s = std::sqrt(acc / double(n - 1));
  }
  const int df = int(n) - 1;
  const double moe = t_critical_95(df) * (n > 0 ? s / std::sqrt(double(n)) : 0.0);
  std::cout << std::fixed << std::setprecision(4);
  std::cout << label << ": " << m << " \xC2\xB1 " << moe << " sec (n=" << n << ", df=" << df << ")\n";
}

// printVector (src/ParSGAoverrides/csr_align.hpp:12-22)
inline void print_vector(const std::string& name, const std::vector<int32_t>& v) {
  std::cout << name << ": ";
  for (size_t i = 0; i < v.size(); ++i) {
    std::cout << static_cast<int64_t>(v[i]);
    if (i + 1 < v.size()) std::cout << " ";
  }
  std::cout << '\n';
}

struct Inputs {
  chsga_dataset* ds = nullptr;
  chsga_ctx* ctx = nullptr;
  chsga_graph_desc desc{};
  ~Inputs() {
    chsga_ctx_destroy(ctx);
    chsga_host_free(ds);
  }
};

inline void load_inputs(const char* I, const char* i, const char* c, const char* l, const char* o, Inputs& in,
                        double* ch_time) {
  const double t0 = now_s();
  if (chsga_host_load(I, i, c, l, o, &in.ds) != CHSGA_OK) die("Error: ", chsga_host_last_error());
  if (ch_time) *ch_time = now_s() - t0;
  chsga_host_describe(in.ds, CHSGA_ALPHABET, &in.desc);
}

// closure 'E': ParSGA's semantics (CHSGA_CREATE_EXACT_CLOSURE) instead of
// ConSGA's; 'X': the same values, the plain path closing rows on match-cut
// components (CHSGA_CREATE_CUT_CLOSURE); 0: ConSGA's
inline void open_engine(Inputs& in, int device, char closure = 0) {
  const uint32_t flags = closure == 'X' ? CHSGA_CREATE_CUT_CLOSURE : closure == 'E' ? CHSGA_CREATE_EXACT_CLOSURE : 0u;
  if (chsga_ctx_create_ex(&in.desc, device, flags, &in.ctx) != CHSGA_OK) die("Error: ", chsga_last_error());
}

}  // namespace drv
