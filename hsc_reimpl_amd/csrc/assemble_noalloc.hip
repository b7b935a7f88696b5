// The assembly before the HSC path without allocation: hsc_assemble_bcoo's
// kernels on a workspace that is a file-scope device array of this file
// instead of library-allocated memory, for callers that cannot allocate
// (code under stream capture, serving loops that forbid hipMalloc).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <mutex>

#include "../../include/hsc.h"
#include "common.hpp"
#include "kernels.hpp"

// The workspace is this file-scope device
// array, one copy per GPU, held until the process ends.  Host code never
// passes the array's name to a kernel or copy: the name is a host-side stand-in,
// not a device address.  Each device's address is looked up once, with that
// device current, by hipGetSymbolAddress, and that pointer is used instead.
constexpr size_t kAsmScratchBytes = 64ull << 20;      // 64 MiB per GPU
constexpr long long kAsmNoAllocMaxCount = 1LL << 20;  // triplets per call
constexpr int kAsmMaxDevices = 64;

static __device__ __align__(256) char scratch[kAsmScratchBytes];

static std::mutex g_scratch_mu;
static char* g_scratch_ptr[kAsmMaxDevices] = {};

extern "C" int32_t hsc_assemble_bcoo_noalloc(const int32_t* rows, const int32_t* cols, const double* values,
                                             int64_t count, int32_t fold_lower, int32_t* out_rows, int32_t* out_cols,
                                             double* out_values, int64_t* out_count, hsc_stream stream) {
  if (!out_count || count < 0 || (count > 0 && (!rows || !cols || !values || !out_rows || !out_cols || !out_values)))
    return HSC_ERR_INVALID_ARGUMENT;
  *out_count = 0;
  if (count > kAsmNoAllocMaxCount) {
    std::fprintf(stderr, "hsc_assemble_bcoo_noalloc: %lld triplets exceed the limit of %lld\n", (long long)count,
                 kAsmNoAllocMaxCount);
    return HSC_ERR_CAPACITY;
  }
  if (count == 0) return HSC_OK;
  hipStream_t st = (hipStream_t)stream;
  int dev = 0;
  hipError_t e = hipStreamGetDevice(st, &dev);
  if (e != hipSuccess || dev < 0 || dev >= kAsmMaxDevices) {
    std::fprintf(stderr, "hsc_assemble_bcoo_noalloc: no usable device for the stream\n");
    return HSC_ERR_HIP;
  }
  hsc::AsmLayout L;
  e = hsc::asm_layout((int)count, st, &L);
  if (e == hipSuccess && L.total > kAsmScratchBytes) {
    std::fprintf(stderr, "hsc_assemble_bcoo_noalloc: workspace of %zu bytes exceeds the %zu-byte scratch\n", L.total,
                 kAsmScratchBytes);
    return HSC_ERR_CAPACITY;
  }
  // one lookup per device, with that device current; the calls of a process
  // are serialised (they share each device's array)
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (e == hipSuccess && !g_scratch_ptr[dev]) {
    int prev = -1;
    e = hipGetDevice(&prev);
    if (e == hipSuccess && prev != dev) e = hipSetDevice(dev);
    void* ptr = nullptr;
    if (e == hipSuccess) e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e == hipSuccess) g_scratch_ptr[dev] = static_cast<char*>(ptr);
    if (prev >= 0 && prev != dev) (void)hipSetDevice(prev);
  }
  long long n = 0;
  if (e == hipSuccess)
    e = hsc::assemble_in(rows, cols, values, count, fold_lower ? 1 : 0, out_rows, out_cols, out_values, &n,
                         g_scratch_ptr[dev], L, st);
  if (e != hipSuccess) {
    std::fprintf(stderr, "hsc_assemble_bcoo_noalloc: %s\n", hipGetErrorString(e));
    return HSC_ERR_HIP;
  }
  *out_count = n;
  return HSC_OK;
}
