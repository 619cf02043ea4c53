// build_pch: contraction hierarchy of a string graph, written in this build's
// CH file layout (the -c input of the ConSGA drivers).  Flags follow
// src/CHoverrides/build_pch.cpp:31-93: -i input string graph, -o output,
// -v label offsets (vertex-weighted shortcuts w1 + w2 + len(mid) - 1).
// -p/-s/-t/-q/-d/-k are accepted and ignored (witness-search and verification
// knobs of the reference's builder); -b (degree-bounded contraction) is
// rejected because it leaves levels that break the level-synchronous sweep.
// Extra: -S <seed> for the tie-break priorities.
#include <getopt.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>

#include "chsga/host.h"

int main(int argc, char** argv) {
  const char *in = nullptr, *out = nullptr, *offs = nullptr;
  unsigned long long seed = 1;
  int c;
  while ((c = getopt(argc, argv, "i:o:p:s:t:q:b:dk:v:S:")) != -1) {
    switch (c) {
      case 'i': in = optarg; break;
      case 'o': out = optarg; break;
      case 'v': offs = optarg; break;
      case 'S': seed = std::strtoull(optarg, nullptr, 10); break;
      case 'p': case 's': case 't': case 'q': case 'd': case 'k': break;
      case 'b':
        std::fprintf(stderr, "Error: degree-bounded contraction (-b) is not supported\n");
        return EXIT_FAILURE;
      default:
        std::fprintf(stderr, "Error: Unknown option %c\n", optopt);
        return EXIT_FAILURE;
    }
  }
  if (!in || !out) {
    std::fprintf(stderr, "Usage: %s -i input_graph -o output_ch_graph [-v label_offsets]\n", argv[0]);
    return EXIT_FAILURE;
  }
  const auto t0 = std::chrono::high_resolution_clock::now();
  if (chsga_host_build_ch_file(in, offs, out, seed) != CHSGA_OK) {
    std::fprintf(stderr, "Error: %s\n", chsga_host_last_error());
    return EXIT_FAILURE;
  }
  const auto t1 = std::chrono::high_resolution_clock::now();
  std::printf("[ build time ] : %f sec\n", std::chrono::duration<double>(t1 - t0).count());
  return 0;
}
