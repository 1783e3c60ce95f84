// gsparc/io_manager.hpp -- the reference's tensor readers
// (include/gsparc/io_manager.hpp, io_manager.tpp:23-484), multithreaded:
//
//  * .tns text: one nonzero per line, `i_1 ... i_N value`, indices 1-based
//    (`based` is subtracted, io_manager.tpp:210); the order is the number of
//    fields on the first data line minus one; dims = observed maximum index
//    + 1 (:213,250-256); duplicates are kept.  The buffer is cut into one
//    range per thread at line boundaries, lines are counted per range, and
//    each range is parsed into its own slice of the arrays (:83-233).  Blank
//    lines and lines starting with '#' are skipped (the reference counts
//    every newline as a nonzero).
//  * quantum directory (`-q`): <dir>/indices.txt holds one line per mode
//    with the nnz 1-based indices of that mode, <dir>/values.txt the nnz
//    values (:310-484); modes are parsed in parallel, dims[m] = the largest
//    index of line m.  Values are read as double and stored as float (the
//    reference scans them with %le into float storage, :480).
//
// Malformed input (a missing value, a non-numeric or zero 1-based index, a
// mode line of the wrong length) throws std::runtime_error.
#ifndef GSPARC_IO_MANAGER_HPP_
#define GSPARC_IO_MANAGER_HPP_
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "sparse_tensor.hpp"

namespace gsparc {

namespace io_detail {

inline std::string read_file(const char* path) {
  FILE* f = std::fopen(path, "rb");
  if (!f) throw std::runtime_error(std::string("[ERROR] Cannot open file \"") + path + "\"");
  std::fseek(f, 0, SEEK_END);
  const long len = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  std::string buf(len > 0 ? (size_t)len : 0, '\0');
  const size_t got = buf.empty() ? 0 : std::fread(&buf[0], 1, buf.size(), f);
  std::fclose(f);
  buf.resize(got);
  return buf;
}

inline int threads_for(size_t bytes) {
  const unsigned h = std::thread::hardware_concurrency();
  const size_t by_size = std::max<size_t>(1, bytes >> 20);  // at least 1 MiB each
  return (int)std::max<size_t>(1, std::min<size_t>(h ? h : 1, std::min<size_t>(by_size, 64)));
}

inline bool is_space(char c) { return c == ' ' || c == '\t' || c == '\r'; }

// A data line: not blank and not a comment.
inline bool data_line(const char* b, const char* e) {
  while (b < e && is_space(*b)) ++b;
  return b < e && *b != '#' && *b != '\n';
}

template <typename F>
void run_threads(int n, F&& f) {
  std::vector<std::thread> ts;
  for (int t = 1; t < n; ++t) ts.emplace_back([&f, t] { f(t); });
  f(0);
  for (auto& th : ts) th.join();
}

}  // namespace io_detail

template <typename TensorType>
class IOManager {
  using sptensor_t = TensorType;
  using index_t = typename sptensor_t::index_t;
  using value_t = typename sptensor_t::value_t;

 public:
  IOManager() = default;
  // Reads `file_path` into a new tensor stored in *tensor (keeping the
  // input number of the tensor passed in; that object is left to the
  // caller, as in the reference).  based: 1 for the 1-based .tns files.
  bool ParseFromFile(const char* file_path, sptensor_t** tensor, int based, bool is_quantum) {
    sptensor_t* t = new sptensor_t((*tensor)->input_number);
    try {
      if (is_quantum) read_quantum(file_path, t);
      else read_tns(file_path, t, based);
    } catch (...) {
      delete t;
      throw;
    }
    *tensor = t;
    return true;
  }

 private:
  static void read_tns(const char* path, sptensor_t* t, int based) {
    using namespace io_detail;
    const std::string buf = read_file(path);
    const char* s = buf.data();
    const size_t L = buf.size();
    // order: fields of the first data line minus the value
    size_t p = 0;
    while (p < L) {
      size_t q = p;
      while (q < L && s[q] != '\n') ++q;
      if (data_line(s + p, s + q)) break;
      p = q + 1;
    }
    int fields = 0;
    for (size_t q = p; q < L && s[q] != '\n';) {
      while (q < L && is_space(s[q])) ++q;
      if (q >= L || s[q] == '\n') break;
      ++fields;
      while (q < L && !is_space(s[q]) && s[q] != '\n') ++q;
    }
    if (fields < 2) throw std::runtime_error(std::string("no nonzeros in ") + path);
    const int order = fields - 1;
    // ranges cut at line starts
    const int T = threads_for(L);
    std::vector<size_t> cut(T + 1, L);
    cut[0] = 0;
    for (int k = 1; k < T; ++k) {
      size_t c = std::max(cut[k - 1], L * k / T);
      while (c < L && c > 0 && s[c - 1] != '\n') ++c;
      cut[k] = c;
    }
    std::vector<uint64_t> lines(T, 0);
    run_threads(T, [&](int k) {
      uint64_t n = 0;
      for (size_t b = cut[k]; b < cut[k + 1];) {
        size_t e = b;
        while (e < cut[k + 1] && s[e] != '\n') ++e;
        if (data_line(s + b, s + e)) ++n;
        b = e + 1;
      }
      lines[k] = n;
    });
    std::vector<uint64_t> first(T + 1, 0);
    for (int k = 0; k < T; ++k) first[k + 1] = first[k] + lines[k];
    const uint64_t nnz = first[T];
    t->order = (unsigned short)order;
    t->nnz = nnz;
    t->dims = allocate<index_t>(order);
    t->indices = allocate<index_t*>(order);
    for (int m = 0; m < order; ++m) t->indices[m] = allocate<index_t>(nnz);
    t->values = allocate<value_t>(nnz);
    std::vector<std::vector<uint64_t>> dmax(T, std::vector<uint64_t>(order, 0));
    std::vector<std::string> err(T);
    run_threads(T, [&](int k) {
      uint64_t i = first[k];
      for (size_t b = cut[k]; b < cut[k + 1] && err[k].empty();) {
        size_t e = b;
        while (e < cut[k + 1] && s[e] != '\n') ++e;
        if (data_line(s + b, s + e)) {
          const char* c = s + b;
          for (int m = 0; m < order; ++m) {
            char* end;
            const unsigned long long v = std::strtoull(c, &end, 10);
            if (end == c || end > s + e || v < (unsigned long long)based) {
              err[k] = "bad index on line of nonzero " + std::to_string(i);
              break;
            }
            const uint64_t idx = v - based;
            t->indices[m][i] = (index_t)idx;
            dmax[k][m] = std::max<uint64_t>(dmax[k][m], idx + 1);
            c = end;
          }
          if (!err[k].empty()) break;
          char* end;
          const double v = std::strtod(c, &end);
          if (end == c || end > s + e) {
            err[k] = "missing value on line of nonzero " + std::to_string(i);
            break;
          }
          t->values[i] = (value_t)v;
          ++i;
        }
        b = e + 1;
      }
    });
    for (int k = 0; k < T; ++k)
      if (!err[k].empty()) throw std::runtime_error(std::string(path) + ": " + err[k]);
    for (int m = 0; m < order; ++m) {
      uint64_t d = 0;
      for (int k = 0; k < T; ++k) d = std::max(d, dmax[k][m]);
      t->dims[m] = (index_t)d;
    }
    t->ComputeNBits();
  }

  static void read_quantum(const char* dir, sptensor_t* t) {
    using namespace io_detail;
    const std::string ib = read_file((std::string(dir) + "/indices.txt").c_str());
    const std::string vb = read_file((std::string(dir) + "/values.txt").c_str());
    // one mode per non-empty line
    std::vector<std::pair<size_t, size_t>> modes;
    for (size_t b = 0; b < ib.size();) {
      size_t e = b;
      while (e < ib.size() && ib[e] != '\n') ++e;
      if (data_line(ib.data() + b, ib.data() + e)) modes.emplace_back(b, e);
      b = e + 1;
    }
    const int order = (int)modes.size();
    if (order < 1) throw std::runtime_error(std::string(dir) + ": no modes in indices.txt");
    // nnz from the first mode line
    uint64_t nnz = 0;
    {
      const char* c = ib.data() + modes[0].first;
      const char* e = ib.data() + modes[0].second;
      while (c < e) {
        while (c < e && is_space(*c)) ++c;
        if (c >= e) break;
        ++nnz;
        while (c < e && !is_space(*c)) ++c;
      }
    }
    t->order = (unsigned short)order;
    t->nnz = nnz;
    t->dims = allocate<index_t>(order);
    t->indices = allocate<index_t*>(order);
    for (int m = 0; m < order; ++m) t->indices[m] = allocate<index_t>(nnz);
    t->values = allocate<value_t>(nnz);
    std::vector<std::string> err(order + 1);
    const int T = std::max(1, std::min(order + 1, threads_for(ib.size() + vb.size()) + 1));
    run_threads(T, [&](int k) {
      for (int m = k; m <= order; m += T) {
        if (m == order) {  // values.txt
          const char* c = vb.data();
          const char* e = vb.data() + vb.size();
          for (uint64_t i = 0; i < nnz; ++i) {
            char* end;
            const double v = std::strtod(c, &end);
            if (end == c) {
              err[m] = "values.txt holds fewer values than indices";
              break;
            }
            t->values[i] = (value_t)v;
            c = end;
          }
          (void)e;
          continue;
        }
        const char* c = ib.data() + modes[m].first;
        const char* e = ib.data() + modes[m].second;
        uint64_t dmax = 0, i = 0;
        while (true) {
          while (c < e && is_space(*c)) ++c;
          if (c >= e) break;
          char* end;
          const double v = std::strtod(c, &end);
          if (end == c || v < 1.0 || i >= nnz) {
            err[m] = "bad index or length on mode line " + std::to_string(m);
            break;
          }
          const uint64_t idx = (uint64_t)v;
          t->indices[m][i++] = (index_t)(idx - 1);
          dmax = std::max(dmax, idx);
          c = end;
        }
        if (err[m].empty() && i != nnz) err[m] = "mode line " + std::to_string(m) + " is short";
        t->dims[m] = (index_t)dmax;
      }
    });
    for (auto& e : err)
      if (!e.empty()) throw std::runtime_error(std::string(dir) + ": " + e);
    t->ComputeNBits();
  }
};

}  // namespace gsparc
#endif
