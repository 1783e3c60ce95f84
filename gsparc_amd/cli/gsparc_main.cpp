// gsparc -- command-line driver with the reference's flags and stderr lines
// (cmdline_opts.cpp:9-171, main.cu:20-432), running the SpTC on the MI355X
// engine through include/gsparc/sptc.hpp.
//
//   gsparc -X x.tns -Y y.tns -c <nc> -x <modes...> -y <modes...> [-n iters]
//          [-d dense_accumulator] [-g 1]
//
// .tns input: one nonzero per line, 1-based indices then the value; dims are
// the observed maximum index + 1 (io_manager.tpp:74-259).  The quantum input
// format (-q) and the chained second contraction (-m/-a/-z) are not part of
// this engine and are rejected.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "gsparc/sptc.hpp"

namespace {

bool read_tns(const char* path, gsparc::SparseTensor<uint64_t, float>* t) {
  FILE* f = std::fopen(path, "rb");
  if (!f) return false;
  std::fseek(f, 0, SEEK_END);
  const long len = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  std::string buf((size_t)(len > 0 ? len : 0), '\0');
  const size_t got = std::fread(&buf[0], 1, buf.size(), f);
  std::fclose(f);
  buf.resize(got);
  // order = fields on the first data line minus the value
  size_t p = 0;
  while (p < buf.size() && (buf[p] == '#' || buf[p] == '\n')) {
    while (p < buf.size() && buf[p] != '\n') ++p;
    ++p;
  }
  int fields = 0;
  for (size_t q = p; q < buf.size() && buf[q] != '\n';) {
    while (q < buf.size() && (buf[q] == ' ' || buf[q] == '\t' || buf[q] == '\r')) ++q;
    if (q >= buf.size() || buf[q] == '\n') break;
    ++fields;
    while (q < buf.size() && !(buf[q] == ' ' || buf[q] == '\t' || buf[q] == '\n' || buf[q] == '\r')) ++q;
  }
  if (fields < 2 || fields - 1 > SPTC_MAX_ORDER) return false;
  t->order = (unsigned short)(fields - 1);
  t->indices.assign(t->order, {});
  t->dims.assign(t->order, 0);
  const char* s = buf.c_str();
  const char* end = s + buf.size();
  const char* c = s + p;
  while (c < end) {
    while (c < end && (*c == ' ' || *c == '\t' || *c == '\r' || *c == '\n')) ++c;
    if (c >= end) break;
    if (*c == '#') {
      while (c < end && *c != '\n') ++c;
      continue;
    }
    for (int m = 0; m < t->order; ++m) {
      char* e;
      const unsigned long long v = std::strtoull(c, &e, 10);
      if (e == c || v == 0) return false;  // 1-based indices
      t->indices[m].push_back(v - 1);
      if (v > t->dims[m]) t->dims[m] = v;
      c = e;
    }
    char* e;
    const double v = std::strtod(c, &e);
    if (e == c) return false;
    t->values.push_back((float)v);
    c = e;
  }
  t->nnz = t->values.size();
  return true;
}

int usage() {
  std::fprintf(stderr,
               "usage: gsparc -X x.tns -Y y.tns -c <nc> -x <modes...> -y <modes...> "
               "[-n iters] [-d -1|0|1] [-g 1]\n");
  return 1;
}

}  // namespace

int main(int argc, char** argv) {
  std::string px, py;
  int nc = -1, iters = 1, dense = 0, gpus = 1;
  std::vector<int> cx, cy;
  for (int i = 1; i < argc; ++i) {
    const std::string a = argv[i];
    auto next = [&]() -> const char* { return i + 1 < argc ? argv[++i] : nullptr; };
    if (a == "-X" || a == "--X") {
      const char* v = next();
      if (!v) return usage();
      px = v;
      std::fprintf(stderr, "tensor X: %s\n", v);
    } else if (a == "-Y" || a == "--Y") {
      const char* v = next();
      if (!v) return usage();
      py = v;
      std::fprintf(stderr, "tensor Y: %s\n", v);
    } else if (a == "-c" || a == "--contract-mode") {
      const char* v = next();
      if (!v) return usage();
      nc = std::atoi(v);
    } else if (a == "-x" || a == "-y" || a == "--x" || a == "--y") {
      if (nc < 0) return usage();
      std::vector<int>& dst = (a.back() == 'x') ? cx : cy;
      std::fprintf(stderr, "# contract modes of %c: %d, contract modes: ",
                   a.back() == 'x' ? 'X' : 'Y', nc);
      for (int k = 0; k < nc; ++k) {
        const char* v = next();
        if (!v) return usage();
        dst.push_back(std::atoi(v));
        std::fprintf(stderr, "%d ", dst.back());
      }
      std::fprintf(stderr, "\n");
    } else if (a == "-n" || a == "--n") {
      const char* v = next();
      if (!v) return usage();
      iters = std::atoi(v);
    } else if (a == "-g" || a == "--g") {
      const char* v = next();
      if (!v) return usage();
      gpus = std::atoi(v);
      std::fprintf(stderr, "num_gpus: %d\n", gpus);
    } else if (a == "-d" || a == "--d") {
      const char* v = next();
      if (!v) return usage();
      dense = std::atoi(v);
      std::fprintf(stderr, "dense_accumulator: %d\n", dense);
    } else if (a == "-q" || a == "-m" || a == "-a" || a == "-z" || a == "-b") {
      std::fprintf(stderr, "gsparc: %s (quantum input / chained contraction) is not supported\n",
                   a.c_str());
      return 1;
    } else if (a == "-h" || a == "--help") {
      return usage();
    } else {
      std::fprintf(stderr, "gsparc: unknown argument %s\n", a.c_str());
      return usage();
    }
  }
  if (px.empty() || py.empty() || nc < 0 || (int)cx.size() != nc || (int)cy.size() != nc ||
      iters < 1)
    return usage();
  if (gpus != 1) {
    std::fprintf(stderr, "gsparc: this engine runs on one GPU (-g 1)\n");
    return 1;
  }
  gsparc::SparseTensor<uint64_t, float> X, Y;
  if (!read_tns(px.c_str(), &X) || !read_tns(py.c_str(), &Y)) {
    std::fprintf(stderr, "gsparc: cannot read the input tensors\n");
    return 1;
  }
  gsparc::Timers sum;
  uint64_t result_nnz = 0;
  try {
    for (int it = 0; it < iters; ++it) {
      // main.cu:293-376's sequence
      gsparc::TensorManager tm(gpus);
      tm.FindPartitionNum(X.nnz, Y.nnz, 0, 0);
      gsparc::SLITOM SX(1), SY(2), SZ;
      tm.ConvertTensor(&X, &SX, nc, cx.data());
      tm.ConvertTensor(&Y, &SY, nc, cy.data());
      tm.SortSlitomXY(&SX, &SY);
      tm.Partition(&SX);
      tm.Partition(&SY);
      gsparc::SpTC(&SX, &SY, &SZ, dense);
      result_nnz += SZ.nnz;
      const gsparc::Timers& t = SZ.timers;
      sum.slitom += t.slitom;
      sum.partition += t.partition;
      sum.idxmatch += t.idxmatch;
      sum.contraction += t.contraction;
      sum.dynamic_partition += t.dynamic_partition;
      sum.total += t.total;
    }
  } catch (const gsparc::Error& e) {
    std::fprintf(stderr, "gsparc: %s (status %d)\n", e.what(), (int)e.status);
    return 2;
  }
  const double n = iters;
  std::fprintf(stderr, "--------Break Down--------\n");
  std::fprintf(stderr, "SLITOM time: %f\n", sum.slitom / n);
  std::fprintf(stderr, "Partition time: %f\n", sum.partition / n);
  std::fprintf(stderr, "Idxmatch time: %f\n", sum.idxmatch / n);
  std::fprintf(stderr, "Contraction time: %f\n", sum.contraction / n);
  std::fprintf(stderr, "Dynamic Partition time: %f\n", sum.dynamic_partition / n);
  std::fprintf(stderr, "--------Elapsed Time--------\n");
  std::fprintf(stderr, "Result NNZ: %llu\n", (unsigned long long)(result_nnz / iters));
  std::fprintf(stderr, "SLITOM time: %f\n", sum.slitom / n);
  std::fprintf(stderr, "Partition time: %f\n", sum.partition / n);
  std::fprintf(stderr, "SpTC(IdxMatch + Contraction + Data Copy) time: %f\n",
               (sum.idxmatch + sum.contraction + sum.dynamic_partition) / n);
  std::fprintf(stderr, "Total time: %f\n", sum.total / n);
  return 0;
}
