// csr_to_adj: a CSR text (n, m, n labels, n + 1 offsets, m destinations) as
// the ConSGA drivers' inputs, with the command line and outputs of
// src/csr_to_adj.cpp: the adjacency graph (header, n, m, n offsets, m
// destinations, m unit weights), the labels concatenated, and the label
// offsets (the running sums, space separated).
//   csr_to_adj IN.csr OUT.adj OUT.labels OUT.label_offsets
#include <cstdint>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

int main(int argc, char** argv) {
  if (argc < 5) {
    std::cerr << "Usage: csr2adj IN.csr OUT.adj OUT.labels OUT.label_offsets\n";
    return 2;
  }
  std::ifstream in(argv[1]);
  if (!in) {
    std::cerr << "cannot open " << argv[1] << "\n";
    return 3;
  }
  uint64_t n = 0, m = 0;
  if (!(in >> n >> m)) {
    std::cerr << argv[1] << ": expected n and m\n";
    return 4;
  }
  std::vector<std::string> lab(n);
  for (auto& l : lab)
    if (!(in >> l)) {
      std::cerr << argv[1] << ": truncated labels\n";
      return 5;
    }
  std::vector<uint64_t> off(n + 1);
  for (auto& o : off)
    if (!(in >> o)) {
      std::cerr << argv[1] << ": truncated offsets\n";
      return 6;
    }
  for (uint64_t i = 0; i < n; ++i)
    if (off[i + 1] < off[i]) {
      std::cerr << argv[1] << ": decreasing offsets at " << i + 1 << "\n";
      return 6;
    }
  if (off[n] != m) std::cerr << "Warning: last offset " << off[n] << " != m " << m << "\n";
  std::vector<uint64_t> dst(m);
  for (auto& d : dst)
    if (!(in >> d)) {
      std::cerr << argv[1] << ": truncated adjacency\n";
      return 7;
    }
  std::ofstream adj(argv[2]), labels(argv[3]), offsets(argv[4]);
  if (!adj || !labels || !offsets) {
    std::cerr << "cannot open an output file\n";
    return 3;
  }
  adj << "AdjacencyGraph\n" << n << '\n' << m << '\n';
  for (uint64_t i = 0; i < n; ++i) adj << off[i] << '\n';
  for (uint64_t d : dst) adj << d << '\n';
  for (uint64_t i = 0; i < m; ++i) adj << "1.0\n";
  uint64_t cur = 0;
  for (const auto& l : lab) {
    labels << l;
    offsets << cur << ' ';
    cur += l.size();
  }
  offsets << cur << '\n';
  return (adj && labels && offsets) ? 0 : 3;
}
