// gfa_to_csr: the string-labelled and char-labelled CSR texts of a GFA 1
// graph, the two outputs of src/handlegraph_to_csr.cpp (which reads a libbdsg
// PackedGraph; GFA takes its place here).  Every segment appears in both
// orientations, as the reference lists a handle graph's handles.
//   gfa_to_csr <graph.gfa> <string_csr_out> <char_csr_out>
#include <cstdio>
#include <cstdlib>

#include "chsga/host.h"

int main(int argc, char** argv) {
  if (argc != 4) {
    std::fprintf(stderr, "Usage: %s <graph.gfa> <string_csr_out> <char_csr_out>\n", argv[0]);
    return 1;
  }
  chsga_dataset* ds = nullptr;
  if (chsga_host_load_gfa(argv[1], &ds) != CHSGA_OK) {
    std::fprintf(stderr, "Error: %s\n", chsga_host_last_error());
    return 1;
  }
  const int rc = chsga_host_write_csr(ds, argv[2], argv[3]);
  if (rc != CHSGA_OK) std::fprintf(stderr, "Error: %s\n", chsga_host_last_error());
  chsga_host_free(ds);
  return rc == CHSGA_OK ? 0 : 1;
}
