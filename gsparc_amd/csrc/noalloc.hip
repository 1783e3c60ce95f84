// Index match that allocates nothing: sptc_index_match_noalloc.
//
// The same computation as sptc_index_match (the reference's IndexMatch,
// indexMatch.cuh:231-287) for callers that cannot allocate device memory
// (code under HIP graph capture, tight serving loops, harnesses that forbid
// hipMalloc).  Its working memory -- the Y contract table and the scan's tile
// sums -- lives in `scratch`, a fixed device array that each GPU holds for
// the life of the process.  Calls on one GPU share that GPU's array, so
// calls that overlap on one GPU (two streams, two threads) race; every call
// rewrites what it reads (the direct table is cleared by a fill kernel).
#include <array>
#include <cstring>
#include <mutex>
#include <string>

#include "kernels.hpp"
#include "scan.hpp"

namespace sptc {

sptc_status api_fail(sptc_status s, const std::string& m);  // api.cpp

namespace {

// Table region: a direct table (two 8-byte entries per contract key, up to
// 2^21 keys) or the distinct-key table (two 8-byte entries per Y nonzero,
// fewer than 2^21 of them).  Tile-sum region: scan_tmp_count(n) + 1 16-byte
// entries (n below 65535 * 2048).
constexpr uint64_t kTableBytes = 32ull << 20;
constexpr uint64_t kTmpBytes = 1ull << 20;
constexpr uint64_t kScratchBytes = kTableBytes + kTmpBytes;
constexpr int kNoAllocDirectBits = 21;
constexpr int kMaxDevices = 64;

__device__ __align__(256) uint8_t scratch[kScratchBytes];

std::mutex g_mu;
std::array<uint8_t*, kMaxDevices> g_ptr{};

// The device address of `scratch` on the current device (its host-side name
// is not one), looked up once per device.
uint8_t* scratch_ptr(std::string* err) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
    *err = "no current device";
    return nullptr;
  }
  std::lock_guard<std::mutex> g(g_mu);
  if (!g_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      *err = std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e);
      return nullptr;
    }
    g_ptr[dev] = static_cast<uint8_t*>(p);
  }
  return g_ptr[dev];
}

}  // namespace

}  // namespace sptc

using namespace sptc;

extern "C" sptc_status sptc_index_match_noalloc(const uint64_t* xkeys, uint64_t nx,
                                                const uint64_t* ykeys, uint64_t ny,
                                                int32_t ncbits, int32_t y_nfbits,
                                                uint64_t* mpos, uint64_t* mcnt,
                                                uint64_t* prefix, uint64_t* ir_nnz,
                                                void* stream) {
  if (ncbits < 0 || ncbits > 64 || y_nfbits < 0 || y_nfbits > 64 || !prefix ||
      (nx && (!xkeys || !mpos || !mcnt)) || (ny && !ykeys))
    return api_fail(SPTC_ERR_BAD_ARGS, "sptc_index_match_noalloc: bad arguments");
  const bool direct = ncbits <= kNoAllocDirectBits;
  const uint64_t table_bytes = direct ? 16ull << ncbits : 16 * (ny + 1);
  const uint64_t tmp_entries = scan_tmp_count(nx > ny ? nx : ny) + 1;
  if (table_bytes > kTableBytes || tmp_entries * sizeof(U2) > kTmpBytes)
    return api_fail(SPTC_ERR_BAD_ARGS,
                    "sptc_index_match_noalloc: sizes above its fixed workspace (contract "
                    "field of at most 21 bits or fewer than 2^21 Y nonzeros, and at most "
                    "65535 * 2048 nonzeros per operand)");
  std::string err;
  uint8_t* base = scratch_ptr(&err);
  if (!base) return api_fail(SPTC_ERR_HIP, err);
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  try {
    uint64_t* t0 = reinterpret_cast<uint64_t*>(base);
    U2* tmp = reinterpret_cast<U2*>(base + kTableBytes);
    YTable tab;
    std::memset(&tab, 0, sizeof(tab));
    if (direct) {
      const uint64_t K = 1ull << ncbits;
      fill_u64(t0, 0, 2 * K, s);
      launch_ytable_direct(ykeys, ny, y_nfbits, K, t0, t0 + K, s);
      tab.direct = 1;
      tab.lo = t0;
      tab.hi = t0 + K;
    } else {
      run_yunique(ykeys, ny, y_nfbits, t0, t0 + ny + 1, tmp, s);
      U2 t;
      SPTC_HIP(hipMemcpyAsync(&t, tmp + scan_tmp_count(ny), sizeof(U2), hipMemcpyDeviceToHost,
                              s));
      SPTC_HIP(hipStreamSynchronize(s));
      tab.direct = 0;
      tab.ukey = t0;
      tab.ustart = t0 + ny + 1;
      tab.U = t.a;
    }
    MatchOut mo;
    std::memset(&mo, 0, sizeof(mo));
    mo.mpos = mpos;
    mo.mcnt = mcnt;
    mo.prefix = prefix;
    run_match(xkeys, nullptr, nx, ncbits, tab, mo, 1, tmp, s);
    U2 t;
    SPTC_HIP(hipMemcpyAsync(&t, tmp + scan_tmp_count(nx), sizeof(U2), hipMemcpyDeviceToHost, s));
    if (nx == 0) fill_u64(prefix, 0, 1, s);
    SPTC_HIP(hipStreamSynchronize(s));
    if (ir_nnz) *ir_nnz = t.a;
  } catch (const HipError& he) {
    return api_fail(SPTC_ERR_HIP, std::string(he.what) + " failed: " + hipGetErrorString(he.err));
  }
  return SPTC_OK;
}
