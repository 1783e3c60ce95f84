// gsparc/sparse_tensor.hpp -- the reference's COO tensor
// (include/gsparc/sparse_tensor.hpp:12-44, sparse_tensor.tpp): order, dims,
// nnz, per-mode 0-based index arrays and values as raw arrays, so code that
// fills or reads them compiles unchanged.  Arrays are malloc'd and owned:
// the destructor frees dims, values and every indices[m].
#ifndef GSPARC_SPARSE_TENSOR_HPP_
#define GSPARC_SPARSE_TENSOR_HPP_
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <new>

namespace gsparc {

template <typename T>
T* allocate(size_t n) {
  T* p = static_cast<T*>(std::malloc(sizeof(T) * (n ? n : 1)));
  if (!p) throw std::bad_alloc();
  return p;
}

template <typename T>
void deallocate(T* p) {
  std::free(p);
}

// bits(d) = 64 - clz(d - 1) (sparse_tensor.tpp:39-47); a dim of 1 needs none.
inline int mode_bits(uint64_t d) { return d <= 1 ? 0 : 64 - __builtin_clzll(d - 1); }

template <typename IndexType, typename ValueType>
class SparseTensor {
 public:
  using index_t = IndexType;
  using value_t = ValueType;

  explicit SparseTensor(int input_number) : input_number(input_number) {}
  // Allocates dims (copied), indices[order][nnz] and values[nnz].
  SparseTensor(unsigned short order, index_t* dims, uint64_t nnz, int input_number)
      : order(order), nnz(nnz), input_number(input_number) {
    this->dims = allocate<index_t>(order);
    std::memcpy(this->dims, dims, sizeof(index_t) * order);
    indices = allocate<index_t*>(order);
    for (int m = 0; m < order; ++m) indices[m] = allocate<index_t>(nnz);
    values = allocate<value_t>(nnz);
    ComputeNBits();
  }
  SparseTensor(const SparseTensor&) = delete;
  SparseTensor& operator=(const SparseTensor&) = delete;
  ~SparseTensor() {
    if (indices)
      for (int m = 0; m < order; ++m) deallocate(indices[m]);
    deallocate(indices);
    deallocate(values);
    deallocate(dims);
  }
  void ComputeNBits() {
    nbits = 0;
    for (int m = 0; m < order; ++m) nbits += mode_bits((uint64_t)dims[m]);
  }

  unsigned short order = 0;
  index_t* dims = nullptr;
  uint64_t nnz = 0;
  int nbits = 0;
  value_t* values = nullptr;
  index_t** indices = nullptr;
  int input_number = 0;
};

}  // namespace gsparc
#endif
