// ConSGA_<alphabet>_dag: drop-in for the reference's score-only driver
// (src/ConSGA.cpp:117-233): same flags -I -i -c -r -l -o and the same report,
// with the alignment run by the gfx950 engine through include/chsga/abi.h.
//
// Extra flags (no collision with the reference's getopt string):
//   -n <cap>    stop after <cap> reads (0 = all; the reference stops at 5)
//   -B          batched path (slot-interleaved rows, device work queue)
//   -g <dev>    GPU index
//   -E          exact insertion closure (ParSGA's scores) instead of ConSGA's
//   -X          the same scores, rows closed on match-cut components (ParSGA's
//               cut-graph insertion) on the plain path
#include "common.hpp"

int main(int argc, char** argv) {
  char *orig_char = nullptr, *orig_str = nullptr, *contr = nullptr, *reads_f = nullptr, *labels_f = nullptr,
       *offsets_f = nullptr;
  int64_t cap = 0;
  bool batched = false;
  char closure = 0;
  int device = 0;
  int opt;
  while ((opt = getopt(argc, argv, "I:C:i:c:r:l:o:a:s:N:M:y:n:BEXg:")) != -1) {
    switch (opt) {
      case 'I': orig_char = optarg; break;
      case 'i': orig_str = optarg; break;
      case 'c': contr = optarg; break;
      case 'r': reads_f = optarg; break;
      case 'l': labels_f = optarg; break;
      case 'o': offsets_f = optarg; break;
      case 'n': cap = std::atoll(optarg); break;
      case 'B': batched = true; break;
      case 'E':
      case 'X': closure = char(opt); break;
      case 'g': device = std::atoi(optarg); break;
      default:
        std::cerr << "Usage: " << argv[0]
                  << " -I orig_char -i orig_string -c contr_string"
                     " -r reads_file -l labels_file -o offsets_file"
                     " \n";
        return EXIT_FAILURE;
    }
  }
  if (!orig_char || !orig_str || !contr || !reads_f || !labels_f || !offsets_f) {
    std::cerr << "Error: missing required arguments\n";
    return EXIT_FAILURE;
  }
  drv::Reads reads;
  drv::load_reads(reads_f, cap, reads);
  const double input_start = drv::now_s();
  drv::Inputs in;
  double ch_time = 0;
  drv::load_inputs(orig_char, orig_str, contr, labels_f, offsets_f, in, &ch_time);
  drv::open_engine(in, device, closure);
  std::cout << "input ch time: " << ch_time << '\n';
  std::cout << "input time: " << drv::now_s() - input_start << '\n' << std::flush;

  const size_t n = reads.size();
  std::vector<int32_t> scores(n);
  std::vector<double> total(n), sub(n), ins(n);
  const auto t0 = std::chrono::high_resolution_clock::now();
  int rc;
  if (batched) {
    rc = chsga_score_batch(in.ctx, reads.ptr.data(), reads.len.data(), int64_t(n), nullptr, scores.data(), nullptr);
  } else {
    chsga_ctx_set_option(in.ctx, CHSGA_OPT_PROFILE, 1);  // per-phase times, as the reference reports
    rc = chsga_score_reads(in.ctx, reads.ptr.data(), reads.len.data(), int64_t(n), scores.data(), total.data(),
                           sub.data(), ins.data(), nullptr);
  }
  if (rc != CHSGA_OK) drv::die("Error: ", chsga_last_error());
  const auto t1 = std::chrono::high_resolution_clock::now();
  if (!batched) {
    std::cout << "Average times (skipping first read), 95% CI:\n";
    drv::print_mean_with_moe_t(" Total", total);
    drv::print_mean_with_moe_t(" Substitutions", sub);
    drv::print_mean_with_moe_t(" Insertions", ins);
  }
  const double ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
  std::cout << "[total over " << n << " reads] : " << ms << " ms\n";
  drv::print_vector("score", scores);
  return 0;
}
