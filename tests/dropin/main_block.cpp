// Harness that compiles the reference's own SpTC iteration body -- the
// block of main.cu that runs FindPartitionNum, ConvertTensor(_extra),
// SortSlitomXY, Partition(_128), convert_and_sort_64/128 and SpTC /
// multi_SpTC / extra::SpTC (main.cu:95-381) -- unchanged, against the
// headers of include/gsparc/ and libsptc_hip.so.  The block is read from the
// reference checkout at build time (MAIN_BLOCK names the extracted file);
// nothing of it is kept in this repository.  Around it, this file sets up
// what main() holds before the loop (options, readers, pools, timers) and
// prints the reference's "Result NNZ" line after it.
#include <cstdio>
#include <cstdlib>
#include <numeric>
#include <vector>

#include "gsparc/sptc.hpp"

#ifndef MAIN_BLOCK
#error "define MAIN_BLOCK as the path of the extracted main.cu block"
#endif

int main(int argc, char** argv) {
  using namespace gsparc;
  CommandLineOptions* options = new CommandLineOptions;
  if (!options->Parse(argc, argv)) return 1;
  using index_t = uint64_t;
  using value_t = float;
  using tensor_t = SparseTensor<index_t, value_t>;
  using io_manager_t = IOManager<tensor_t>;
  using cuda_memory_t = CudaMemoryPool;
  try {
    io_manager_t* io_manager = new io_manager_t;
    tensor_t* tensor_x = new tensor_t(1);
    tensor_t* tensor_y = new tensor_t(2);
    tensor_t* tensor_z = new tensor_t(2);
    double reorder_bit_multi_sptc = 0.0;
    (void)reorder_bit_multi_sptc;
    io_manager->ParseFromFile(options->get_first_input_path(), &tensor_x, 1,
                              options->get_quantum());
    io_manager->ParseFromFile(options->get_second_input_path(), &tensor_y, 1,
                              options->get_quantum());
    if (options->is_multi_sptc())
      io_manager->ParseFromFile(options->get_second_input_path(), &tensor_z, 1,
                                options->get_quantum());
    tensor_x->ComputeNBits();
    tensor_y->ComputeNBits();
    cuda_memory_t** cuda_memory = new cuda_memory_t*[options->get_gpu_count()];
    for (int i = 0; i < options->get_gpu_count(); i++) cuda_memory[i] = new cuda_memory_t(i);
    const int num_iter = options->get_num_iter();
    uint64_t result_nnz = 0;
    Timer* slitom_timer = new Timer;
    Timer* partition_timer = new Timer;
    Timer** idxmatch_timers = new Timer*[options->get_gpu_count()];
    Timer** dynamicprtn_timers = new Timer*[options->get_gpu_count()];
    Timer** contraction_timers = new Timer*[options->get_gpu_count()];
    Timer** copy_timers = new Timer*[options->get_gpu_count()];
    for (int i = 0; i < options->get_gpu_count(); i++) {
      idxmatch_timers[i] = new Timer;
      dynamicprtn_timers[i] = new Timer;
      contraction_timers[i] = new Timer;
      copy_timers[i] = new Timer;
    }
    for (int iter = 0; iter < num_iter; iter++) {
      for (int i = 0; i < options->get_gpu_count(); i++) cuda_memory[i]->reset();
#include MAIN_BLOCK
    }
    std::fprintf(stderr, "Result NNZ: %llu\n", (unsigned long long)(result_nnz / num_iter));
    std::fprintf(stderr, "SLITOM time: %f\n", slitom_timer->getTotalTime() / num_iter);
    std::fprintf(stderr, "Partition time: %f\n", partition_timer->getTotalTime() / num_iter);
  } catch (const Error& e) {
    std::fprintf(stderr, "main_block: %s (status %d)\n", e.what(), (int)e.status);
    return 2;
  }
  return 0;
}
