// gsparc/cmdline_opts.hpp -- the reference's command line
// (include/gsparc/cmdline_opts.hpp, src/gsparc/cmdline_opts.cpp:9-171):
// the same flags, getters and stderr echo lines (figure scripts parse them).
//
//   -X/--X path  -Y/--Y path  -c/--contract-mode n  -x m...  -y m...
//   -a/--add-contract-mode n  -z m...  -m/--multi-sptc 0|1  -n iters
//   -g gpus  -d -1|0|1  -q (quantum directory input)  -b  -h
#ifndef GSPARC_CMDLINE_OPTS_HPP_
#define GSPARC_CMDLINE_OPTS_HPP_
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace gsparc {

class CommandLineOptions {
 public:
  // Parses argv; returns false (after a usage line on stderr) on a malformed
  // command line.
  bool Parse(int argc, char** argv) {
    for (int i = 1; i < argc; ++i) {
      const std::string a = argv[i];
      auto next = [&]() -> const char* { return i + 1 < argc ? argv[++i] : nullptr; };
      auto list = [&](std::vector<int>& dst, int n, const char* label) {
        std::fprintf(stderr, "# contract modes of %s: %d, contract modes: ", label, n);
        dst.clear();
        for (int k = 0; k < n; ++k) {
          const char* v = next();
          if (!v) return false;
          dst.push_back(std::atoi(v));
          std::fprintf(stderr, "%d ", dst.back());
        }
        std::fprintf(stderr, "\n");
        return true;
      };
      const char* v = nullptr;
      if (a == "-X" || a == "--X") {
        if (!(v = next())) return usage();
        x_ = v;
        std::fprintf(stderr, "tensor X: %s\n", v);
      } else if (a == "-Y" || a == "--Y") {
        if (!(v = next())) return usage();
        y_ = v;
        std::fprintf(stderr, "tensor Y: %s\n", v);
      } else if (a == "-c" || a == "--contract-mode") {
        if (!(v = next())) return usage();
        nc_ = std::atoi(v);
      } else if (a == "-a" || a == "--add-contract-mode") {
        if (!(v = next())) return usage();
        nac_ = std::atoi(v);
      } else if (a == "-x" || a == "--x") {
        if (!list(cx_, nc_, "X")) return usage();
      } else if (a == "-y" || a == "--y") {
        if (!list(cy_, nc_, "Y")) return usage();
      } else if (a == "-z" || a == "--z") {
        if (!list(cz_, nac_, "Z (second SpTC)")) return usage();
      } else if (a == "-m" || a == "--multi-sptc") {
        if (!(v = next())) return usage();
        multi_ = std::atoi(v);
        std::fprintf(stderr, "multi_sptc: %d\n", multi_);
      } else if (a == "-n" || a == "--n") {
        if (!(v = next())) return usage();
        iters_ = std::atoi(v);
      } else if (a == "-g" || a == "--g") {
        if (!(v = next())) return usage();
        gpus_ = std::atoi(v);
        std::fprintf(stderr, "num_gpus: %d\n", gpus_);
      } else if (a == "-d" || a == "--d") {
        if (!(v = next())) return usage();
        dense_ = std::atoi(v);
        std::fprintf(stderr, "dense_accumulator: %d\n", dense_);
      } else if (a == "-q" || a == "--q") {
        quantum_ = 1;
      } else if (a == "-b" || a == "--b") {
        // accepted and ignored, as in the reference
      } else {
        return usage();
      }
    }
    if (x_.empty() || y_.empty() || nc_ < 0 || (int)cx_.size() != nc_ ||
        (int)cy_.size() != nc_ || iters_ < 1 || (multi_ && (int)cz_.size() != nac_))
      return usage();
    return true;
  }

  const char* get_first_input_path() const { return x_.c_str(); }
  const char* get_second_input_path() const { return y_.c_str(); }
  const int* get_first_input_cmodes() const { return cx_.data(); }
  const int* get_second_input_cmodes() const { return cy_.data(); }
  const int* get_third_cmodes() const { return cz_.data(); }
  int get_num_cmodes() const { return nc_; }
  int get_gpu_count() const { return gpus_; }
  int get_quantum() const { return quantum_; }
  int get_num_iter() const { return iters_; }
  int get_add_num_cmodes() const { return nac_; }
  int is_multi_sptc() const { return multi_; }
  int is_dense_accumulator() const { return dense_; }

 private:
  static bool usage() {
    std::fprintf(stderr,
                 "usage: gsparc -X x.tns -Y y.tns -c <nc> -x <modes...> -y <modes...> "
                 "[-m 1 -a <n> -z <modes...>] [-q] [-n iters] [-d -1|0|1] [-g 1]\n");
    return false;
  }
  std::string x_, y_;
  int nc_ = -1, nac_ = 0, gpus_ = 1, quantum_ = 0, iters_ = 1, multi_ = 0, dense_ = 0;
  std::vector<int> cx_, cy_, cz_;
};

}  // namespace gsparc
#endif
