// PaSGAL_match: the reference's baseline local aligner (Makefile:197-213 builds
// PaSGAL with src/PaSGALoverrides/align.hpp), aligning reads to the ConSGA
// inputs on the gfx950 engine (chsga_pasgal_align) for comparison with the
// ConSGA drivers on the same graph.
//
// PaSGAL's own graph readers (.vg, .txt) live in its submodule, which the
// reference leaves empty; this driver reads the ConSGA inputs instead:
//   -I char_adj -i string_adj [-c contracted] -l labels -o offsets
// and PaSGAL's query/output pair:
//   -q reads (FASTA/FASTQ[.gz])   -O output file
// Extra flags:
//   -s          phase 1 only (PaSGAL built with PASGAL_SCORE_ONLY, align.hpp:571-575)
//   -S m,x,i,d  match, mismatch, ins, del (PaSGAL's defaults 1,1,1,1)
//   -n <cap>    stop after <cap> reads;  -g <dev> GPU index;  -t <threads> accepted, unused
// Output (printResultsToFile, align.hpp:730-750), one line per read: name, length,
// qryRowStart, qryRowEnd, strand, the string vertex of refColumnStart and of
// refColumnEnd, score, CIGAR.  Reads are named read_<index> (the loader keeps
// no names).  In phase-1-only mode the start fields are -1 and the CIGAR is empty.
#include <algorithm>
#include <cstring>
#include <fstream>

#include "common.hpp"

int main(int argc, char** argv) {
  char *orig_char = nullptr, *orig_str = nullptr, *contr = nullptr, *labels_f = nullptr, *offsets_f = nullptr,
       *reads_f = nullptr, *out_f = nullptr;
  int64_t cap = 0;
  int device = 0;
  chsga_pasgal_params P{1, 1, 1, 1, 0};
  int opt;
  while ((opt = getopt(argc, argv, "I:i:c:l:o:q:O:sS:n:g:t:")) != -1) {
    switch (opt) {
      case 'I': orig_char = optarg; break;
      case 'i': orig_str = optarg; break;
      case 'c': contr = optarg; break;
      case 'l': labels_f = optarg; break;
      case 'o': offsets_f = optarg; break;
      case 'q': reads_f = optarg; break;
      case 'O': out_f = optarg; break;
      case 's': P.score_only = 1; break;
      case 'S':
        if (std::sscanf(optarg, "%d,%d,%d,%d", &P.match, &P.mismatch, &P.ins, &P.del) != 4)
          drv::die("Error: -S takes match,mismatch,ins,del", nullptr);
        break;
      case 'n': cap = std::atoll(optarg); break;
      case 'g': device = std::atoi(optarg); break;
      case 't': break;
      default:
        std::cerr << "Usage: " << argv[0]
                  << " -I orig_char -i orig_string [-c contr_string] -l labels_file -o offsets_file"
                     " -q reads_file -O output_file [-s] [-S m,x,i,d]\n";
        return EXIT_FAILURE;
    }
  }
  if (!orig_char || !orig_str || !labels_f || !offsets_f || !reads_f || !out_f) {
    std::cerr << "Error: missing required arguments\n";
    return EXIT_FAILURE;
  }
  drv::Reads reads;
  drv::load_reads(reads_f, cap, reads);
  std::cout << "INFO, psgl::alignToDAG, total count of reads = " << reads.size() << std::endl;
  drv::Inputs in;
  drv::load_inputs(orig_char, orig_str, contr, labels_f, offsets_f, in, nullptr);
  drv::open_engine(in, device, 'E');
  chsga_dataset* rds = nullptr;
  chsga_ctx* rctx = nullptr;
  chsga_graph_desc rdesc{};
  if (!P.score_only) {
    if (chsga_host_reverse(in.ds, 1, &rds) != CHSGA_OK) drv::die("Error: ", chsga_host_last_error());
    chsga_host_describe(rds, CHSGA_ALPHABET, &rdesc);
    if (chsga_ctx_create_ex(&rdesc, device, CHSGA_CREATE_EXACT_CLOSURE, &rctx) != CHSGA_OK)
      drv::die("Error: ", chsga_last_error());
  }
  const size_t n = reads.size();
  std::vector<chsga_pasgal_result> res(n);
  chsga_stats st{};
  const double t0 = drv::now_s();
  if (chsga_pasgal_align(in.ctx, rctx, &P, reads.ptr.data(), reads.len.data(), int64_t(n), res.data(), &st) !=
      CHSGA_OK)
    drv::die("Error: ", chsga_last_error());
  const double sec = drv::now_s() - t0;
  std::cout << "TIMER, psgl::alignToDAG, GPU time of phases 1" << (P.score_only ? "" : ", 1-R and 2")
            << " (s) = " << sec << ", GCUPS = " << (sec > 0 ? double(st.cells) / sec * 1e-9 : 0.0) << std::endl;
  // originalVertexId: the string vertex of a char
  const uint64_t* so = in.desc.str_off;
  auto og = [&](int64_t p) -> int64_t {
    if (p < 0) return -1;
    return int64_t(std::upper_bound(so, so + in.desc.n_s + 1, uint64_t(p)) - so) - 1;
  };
  std::ofstream os(out_f);
  if (!os) drv::die("Error: cannot write ", out_f);
  int64_t flagged = 0;
  for (size_t q = 0; q < n; ++q) {
    const chsga_pasgal_result& r = res[q];
    os << "read_" << q << '\t' << reads.len[q] << '\t' << r.qry_start << '\t' << r.qry_end << '\t' << r.strand
       << '\t' << og(r.ref_start) << '\t' << og(r.ref_end) << '\t' << r.score << '\t' << (r.cigar ? r.cigar : "")
       << '\n';
    flagged += r.flags != 0;
  }
  if (flagged) std::cout << "INFO, psgl::alignToDAG, reads with result flags set = " << flagged << std::endl;
  chsga_pasgal_free(res.data(), int64_t(n));
  chsga_ctx_destroy(rctx);
  chsga_host_free(rds);
  return 0;
}
