// gsparc -- command-line driver with the reference's flags and stderr lines
// (cmdline_opts.cpp:9-171, main.cu:20-432), running main.cu's sequence on
// the MI355X engine through the reference-shaped headers of
// include/gsparc/: IOManager (.tns or quantum directory), TensorManager
// (FindPartitionNum, ConvertTensor, SortSlitomXY, Partition,
// convert_and_sort_64), then gsparc::SpTC, or for -m 1 the chained
// multi_SpTC pair, iterated -n times.
//
//   gsparc -X x.tns -Y y.tns -c <nc> -x <modes...> -y <modes...>
//          [-m 1 -a <nc2> -z <modes...>] [-q] [-n iters] [-d -1|0|1] [-g 1]
//
// The chained form contracts the first output's trailing nc2 modes (its Y
// free modes) with modes -z of the second input (main.cu:323-371).
#include <cstdio>
#include <memory>
#include <vector>

#include "gsparc/sptc.hpp"

int main(int argc, char** argv) {
  using namespace gsparc;
  CommandLineOptions options;
  if (!options.Parse(argc, argv)) return 1;
  if (options.get_gpu_count() != 1) {
    std::fprintf(stderr, "gsparc: this engine runs on one GPU (-g 1)\n");
    return 1;
  }
  using index_t = uint64_t;
  using value_t = float;
  using tensor_t = SparseTensor<index_t, value_t>;
  using lindex_t = uint64_t;
  using slitom_t = SLITOM<uint64_t, uint8_t, lindex_t, value_t>;
  using tensor_manager_t = TensorManager<tensor_t, slitom_t>;

  IOManager<tensor_t> io;
  // ParseFromFile replaces the tensor it is given (io_manager.tpp:23-34)
  auto parse = [&](const char* path, int number) {
    std::unique_ptr<tensor_t> shell(new tensor_t(number));
    tensor_t* t = shell.get();
    io.ParseFromFile(path, &t, 1, options.get_quantum() != 0);
    return std::unique_ptr<tensor_t>(t);
  };
  int rc = 0;
  try {
    std::unique_ptr<tensor_t> X = parse(options.get_first_input_path(), 1);
    std::unique_ptr<tensor_t> Y = parse(options.get_second_input_path(), 2);
    std::unique_ptr<tensor_t> W;
    if (options.is_multi_sptc()) W = parse(options.get_second_input_path(), 2);
    tensor_t* tx = X.get();
    tensor_t* ty = Y.get();
    tensor_t* tz = W.get();

    std::unique_ptr<CudaMemoryPool> pool_owner(new CudaMemoryPool(0));
    CudaMemoryPool* pool = pool_owner.get();
    CudaMemoryPool* pools[1] = {pool};
    Timer slitom_timer, partition_timer;
    Timer copy, idxmatch, dprtn, contraction;
    Timer *copy_t[1] = {&copy}, *idx_t[1] = {&idxmatch}, *dprtn_t[1] = {&dprtn},
          *contr_t[1] = {&contraction};
    double reorder = 0.0;
    uint64_t result_nnz = 0;
    const int iters = options.get_num_iter();
    for (int it = 0; it < iters; ++it) {
      pool->reset();
      tensor_manager_t tm(options.get_gpu_count());
      const bool sort_gpu = tm.FindPartitionNum(tx->nnz, ty->nnz, pool->getPoolSize(), tx->nbits);
      slitom_t SX(1, sort_gpu), SY(2, sort_gpu), SZ;
      tm.ConvertTensor(tx, &SX, options.get_num_cmodes(), options.get_first_input_cmodes(),
                       &slitom_timer);
      tm.ConvertTensor(ty, &SY, options.get_num_cmodes(), options.get_second_input_cmodes(),
                       &slitom_timer);
      tm.SortSlitomXY(&SX, &SY, pools, &slitom_timer);
      partition_timer.start();
      tm.Partition(&SX);
      tm.Partition(&SY);
      partition_timer.stop();
      if (SY.nprtn > 1) {
        partition_timer.start();
        tm.convert_and_sort_64(&SY, pools);
        partition_timer.stop();
      }
      if (options.is_multi_sptc()) {
        multi_SpTC<tensor_t, slitom_t>(&SX, &SY, &SZ, pools, 1, copy_t, idx_t, dprtn_t, contr_t);
        const double first = slitom_timer.getTotalTime() + partition_timer.getTotalTime() +
                             copy.getTotalTime() + idxmatch.getTotalTime() +
                             contraction.getTotalTime() + dprtn.getTotalTime();
        std::fprintf(stderr, "First SpTC time: %f\n", first);
        // the intermediate stays on the device; its keys need no reordering
        // (main.cu:352-358 rewrites them into the same order)
        slitom_t Z;
        tm.ConvertTensor(tz, &Z, options.get_add_num_cmodes(), options.get_third_cmodes(),
                         &slitom_timer);
        tm.SortSlitomXY(&SZ, &Z, pools, &slitom_timer);
        multi_SpTC<tensor_t, slitom_t>(&SZ, &SY, &Z, pools, 1, copy_t, idx_t, dprtn_t, contr_t);
        result_nnz += Z.nnz;
      } else {
        SpTC<tensor_t, slitom_t>(&SX, &SY, &SZ, pools, 1, copy_t, idx_t, dprtn_t, contr_t,
                                 options.is_dense_accumulator());
        result_nnz += SZ.nnz;
      }
    }
    const double n = iters;
    const double slitom_s = slitom_timer.getTotalTime(), part_s = partition_timer.getTotalTime();
    const double copy_s = copy.getTotalTime(), idx_s = idxmatch.getTotalTime(),
                 con_s = contraction.getTotalTime(), dp_s = dprtn.getTotalTime();
    const double total = slitom_s + part_s + copy_s + idx_s + con_s + dp_s;
    std::fprintf(stderr, "--------Break Down--------\n");
    std::fprintf(stderr, "SLITOM time: %f\n", slitom_s / n);
    std::fprintf(stderr, "Partition time: %f\n", part_s / n);
    std::fprintf(stderr, "Idxmatch time: %f\n", idx_s / n);
    std::fprintf(stderr, "Contraction time: %f\n", con_s / n);
    std::fprintf(stderr, "Dynamic Partition time: %f\n", dp_s / n);
    if (options.is_multi_sptc()) {
      std::fprintf(stderr, "Bit reordering time: %f\n", reorder);
      std::fprintf(stderr, "Bit reordering ratio: %f%%\n", total > 0 ? reorder / (total / n) * 100 : 0.0);
    }
    std::fprintf(stderr, "--------Elapsed Time--------\n");
    std::fprintf(stderr, "Result NNZ: %llu\n", (unsigned long long)(result_nnz / iters));
    std::fprintf(stderr, "SLITOM time: %f\n", slitom_s / n);
    std::fprintf(stderr, "Partition time: %f\n", part_s / n);
    std::fprintf(stderr, "SpTC(IdxMatch + Contraction + Data Copy) time: %f\n",
                 (copy_s + idx_s + con_s + dp_s) / n);
    std::fprintf(stderr, "Total time: %f\n", total / n);
  } catch (const Error& e) {
    std::fprintf(stderr, "gsparc: %s (status %d)\n", e.what(), (int)e.status);
    rc = 2;
  } catch (const std::exception& e) {
    std::fprintf(stderr, "gsparc: %s\n", e.what());
    rc = 1;
  }
  return rc;
}
