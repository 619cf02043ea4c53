// ConSGA_traceback_<alphabet>_dag: drop-in for src/ConSGA_traceback.cpp:53-226.
// Same flags (-b, -k, -d plus the inputs), same verification report and the
// same -b benchmark block.  Checkpoint rows live in HBM, so -d names no files
// here; it is accepted and echoed for compatibility.
//
// Extra flags: -n <cap> (read cap, 0 = all), -g <dev> (GPU index), -B (batched
// forward pass keeping every k-th row of every read, then per-read segment
// recompute and walk: chsga_traceback_batch), -E (exact insertion closure:
// ParSGA's scores and paths instead of ConSGA's), -X (the same, rows closed
// on match-cut components on the plain path: ParSGA's cut-graph insertion).
#include "common.hpp"

// print_mean_with_moe of src/ConSGA_traceback.cpp:26-51 (z = 1.96)
static void print_mean_with_moe_z(const char* label, const std::vector<double>& times) {
  if (times.size() <= 1) {
    std::cout << " " << label << ":        " << (times.empty() ? 0.0 : times[0]) << " sec\n";
    return;
  }
  std::vector<double> trimmed(times.begin() + 1, times.end());
  double mean = 0.0;
  for (double t : trimmed) mean += t;
  mean /= static_cast<double>(trimmed.size());
  double var = 0.0;
  for (double t : trimmed) var += (t - mean) * (t - mean);
  var /= static_cast<double>(trimmed.size() - 1);
  const double moe = 1.96 * std::sqrt(var / static_cast<double>(trimmed.size()));
  std::cout << " " << label << ":        " << std::fixed << std::setprecision(4) << mean << " \xC2\xB1 " << moe
            << " sec (n=" << trimmed.size() << ", df=" << int(trimmed.size()) - 1 << ")\n";
}

int main(int argc, char** argv) {
  char *orig_char = nullptr, *orig_str = nullptr, *contr = nullptr, *reads_f = nullptr, *labels_f = nullptr,
       *offsets_f = nullptr;
  int64_t k = 500, cap = 0;
  std::string dir = "/tmp/consga_traceback_ckpts";
  bool benchmark = false, batched = false;
  char closure = 0;
  int device = 0;
  int opt;
  while ((opt = getopt(argc, argv, "bBEXI:i:c:r:l:o:k:d:n:g:")) != -1) {
    switch (opt) {
      case 'b': benchmark = true; break;
      case 'B': batched = true; break;
      case 'E':
      case 'X': closure = char(opt); break;
      case 'I': orig_char = optarg; break;
      case 'i': orig_str = optarg; break;
      case 'c': contr = optarg; break;
      case 'r': reads_f = optarg; break;
      case 'l': labels_f = optarg; break;
      case 'o': offsets_f = optarg; break;
      case 'k': k = std::atoll(optarg); break;
      case 'd': dir = optarg; break;
      case 'n': cap = std::atoll(optarg); break;
      case 'g': device = std::atoi(optarg); break;
      default:
        std::cerr << "Usage: " << argv[0]
                  << " [-b] -I orig_char -i orig_string -c contracted"
                     " -r reads -l labels -o offsets"
                     " [-k checkpoint_every_k_rows] [-d ckpt_dir]\n";
        return EXIT_FAILURE;
    }
  }
  if (!orig_char || !orig_str || !contr || !reads_f || !labels_f || !offsets_f) {
    std::cerr << "Error: missing required arguments\n";
    return EXIT_FAILURE;
  }
  if (k <= 0) {
    std::cerr << "Error: -k must be a positive number of rows\n";
    return EXIT_FAILURE;
  }
  drv::Reads reads;
  drv::load_reads(reads_f, cap, reads);
  drv::Inputs in;
  drv::load_inputs(orig_char, orig_str, contr, labels_f, offsets_f, in, nullptr);
  drv::open_engine(in, device, closure);
  const int match_score = in.desc.match_score, subst_score = in.desc.subst_score, del_score = in.desc.del_score,
            ins_score = in.desc.ins_score;
  const size_t n = reads.size();

  if (benchmark && batched) {
    std::vector<chsga_path*> paths(n, nullptr);
    const auto t0 = std::chrono::high_resolution_clock::now();
    if (chsga_traceback_batch(in.ctx, reads.ptr.data(), reads.len.data(), int64_t(n), k, nullptr, 0, paths.data(),
                              nullptr) != CHSGA_OK)
      drv::die("Error: ", chsga_last_error());
    const auto t1 = std::chrono::high_resolution_clock::now();
    for (chsga_path* p : paths) chsga_path_free(p);
    std::cout << "Average times (skipping first read), 95% CI:\n";
    std::cout << " Total:        NA sec (batched)\n";
    std::cout << " Substitutions: NA sec\n";
    std::cout << " Insertions:    NA sec\n";
    std::cout << "[total over " << n << " reads] : " << std::chrono::duration<double, std::milli>(t1 - t0).count()
              << " ms\n";
    return 0;
  }
  if (benchmark) {
    std::vector<double> per_read;
    const auto t0 = std::chrono::high_resolution_clock::now();
    for (size_t ri = 0; ri < n; ++ri) {
      const auto r0 = std::chrono::high_resolution_clock::now();
      chsga_path* p = nullptr;
      if (chsga_traceback_read(in.ctx, reads.ptr[ri], reads.len[ri], k, &p, nullptr) != CHSGA_OK)
        drv::die("Error: ", chsga_last_error());
      chsga_path_free(p);
      const auto r1 = std::chrono::high_resolution_clock::now();
      per_read.push_back(std::chrono::duration<double>(r1 - r0).count());
    }
    const auto t1 = std::chrono::high_resolution_clock::now();
    std::cout << "Average times (skipping first read), 95% CI:\n";
    print_mean_with_moe_z(" Total", per_read);
    std::cout << " Substitutions: NA sec\n";
    std::cout << " Insertions:    NA sec\n";
    std::cout << "[total over " << n << " reads] : " << std::chrono::duration<double, std::milli>(t1 - t0).count()
              << " ms\n";
    return 0;
  }

  std::cout << "=== Forward-pass scores (align_combo) ===\n";
  std::vector<int32_t> scores(n);
  std::vector<double> total(n), sub(n), ins(n);
  chsga_ctx_set_option(in.ctx, CHSGA_OPT_PROFILE, 1);
  if (chsga_score_reads(in.ctx, reads.ptr.data(), reads.len.data(), int64_t(n), scores.data(), total.data(),
                        sub.data(), ins.data(), nullptr) != CHSGA_OK)
    drv::die("Error: ", chsga_last_error());
  chsga_ctx_set_option(in.ctx, CHSGA_OPT_PROFILE, 0);
  std::cout << "Average times (skipping first read), 95% CI:\n";
  drv::print_mean_with_moe_t(" Total", total);
  drv::print_mean_with_moe_t(" Substitutions", sub);
  drv::print_mean_with_moe_t(" Insertions", ins);
  for (size_t i = 0; i < n; ++i) std::cout << "read[" << i << "] score=" << scores[i] << "\n";

  std::cout << "\n=== Traceback results (align_combo_with_traceback) ===\n";
  std::cout << "  (checkpoint every k=" << k << " rows, dir=" << dir << ")\n";
  std::vector<chsga_path*> paths(n, nullptr);
  const auto t0 = std::chrono::high_resolution_clock::now();
  if (batched) {
    if (chsga_traceback_batch(in.ctx, reads.ptr.data(), reads.len.data(), int64_t(n), k, nullptr, 0, paths.data(),
                              nullptr) != CHSGA_OK)
      drv::die("Error: ", chsga_last_error());
  } else {
    for (size_t ri = 0; ri < n; ++ri)
      if (chsga_traceback_read(in.ctx, reads.ptr[ri], reads.len[ri], k, &paths[ri], nullptr) != CHSGA_OK)
        drv::die("Error: ", chsga_last_error());
  }
  const auto t1 = std::chrono::high_resolution_clock::now();
  const double ms = std::chrono::duration<double, std::milli>(t1 - t0).count();

  bool all_ok = true;
  const char* labels = chsga_ds_labels(in.ds);
  for (size_t ri = 0; ri < n; ++ri) {
    const chsga_path* p = paths[ri];
    const std::string read = reads.str(ri);
    // ops_to_cigar (src/traceback.hpp:1081-1110) and the rescoring of
    // src/ConSGA_traceback.cpp:190-206
    std::string cigar;
    char prev = 0;
    int run = 0;
    int path_score = 0, read_pos = -1;
    for (int64_t j = 0; j < p->n_ops; ++j) {
      const chsga_align_op& o = p->ops[j];
      const char c = o.op == CHSGA_OP_DELETE ? 'D' : (o.op == CHSGA_OP_MATCHSUB ? 'M' : 'I');
      if (c == prev) {
        ++run;
      } else {
        if (prev) cigar += std::to_string(run) + prev;
        prev = c;
        run = 1;
      }
      switch (o.op) {
        case CHSGA_OP_DELETE:
          path_score += del_score;
          read_pos = int(o.row);
          break;
        case CHSGA_OP_MATCHSUB: {
          const char a = labels[o.vertex], b = read[size_t(o.row)];
          const bool m = (CHSGA_ALPHABET == 5) ? (a == b) : (a == 'N' || b == 'N' || a == b);
          path_score += m ? match_score : subst_score;
          read_pos = int(o.row);
          break;
        }
        case CHSGA_OP_INSERT:
          path_score += ins_score;
          break;
      }
    }
    if (prev) cigar += std::to_string(run) + prev;
    const bool score_ok = path_score == scores[ri];
    const bool len_ok = read_pos == int(read.size()) - 1;
    std::cout << "read[" << ri << "] = \"" << read << "\"\n"
              << "  CIGAR      : " << cigar << "\n"
              << "  path_score : " << path_score << "  ref_score  : " << scores[ri]
              << "  score_ok   : " << (score_ok ? "YES" : "FAIL") << "\n"
              << "  last_row   : " << read_pos << "  expected   : " << (read.size() - 1)
              << "  row_ok     : " << (len_ok ? "YES" : "FAIL") << "\n";
    if (!score_ok || !len_ok) all_ok = false;
    chsga_path_free(paths[ri]);
  }
  std::cout << "\n[traceback time] " << ms << " ms\n";
  std::cout << "\n[RESULT] " << (all_ok ? "ALL CHECKS PASSED" : "SOME CHECKS FAILED") << "\n";
  return all_ok ? 0 : 1;
}
