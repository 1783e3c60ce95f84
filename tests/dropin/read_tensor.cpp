// Reads a tensor with gsparc::IOManager (a .tns file, or with -q a quantum
// directory) and prints what the tests compare: order, nnz, dims, and per
// mode the sum of the 0-based indices, then the FP64 sum of the values.
#include <cstdio>
#include <cstring>
#include <memory>

#include "gsparc/io_manager.hpp"

int main(int argc, char** argv) {
  if (argc < 2) return 2;
  const bool quantum = argc > 2 && std::strcmp(argv[2], "-q") == 0;
  using tensor_t = gsparc::SparseTensor<uint64_t, float>;
  gsparc::IOManager<tensor_t> io;
  std::unique_ptr<tensor_t> shell(new tensor_t(1));
  tensor_t* t = shell.get();
  try {
    io.ParseFromFile(argv[1], &t, 1, quantum);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "%s\n", e.what());
    return 1;
  }
  std::unique_ptr<tensor_t> own(t);
  std::printf("order %d nnz %llu\n", t->order, (unsigned long long)t->nnz);
  for (int m = 0; m < t->order; ++m) {
    unsigned long long s = 0;
    for (uint64_t i = 0; i < t->nnz; ++i) s += t->indices[m][i];
    std::printf("mode %d dim %llu sum %llu\n", m, (unsigned long long)t->dims[m], s);
  }
  double v = 0;
  for (uint64_t i = 0; i < t->nnz; ++i) v += t->values[i];
  std::printf("values %.17g\n", v);
  return 0;
}
