// qr_compact! without a workspace argument, for callers that cannot allocate
// device memory (code under graph capture, tight serving loops, harnesses
// that forbid hipMalloc): the QR driver's workspace is a file-scope device
// array of kScratchBytes, one per GPU, held for the life of the process.
//
// Host code never passes the array's name to a kernel (in host code it is a
// stand-in, not a device address): its address on each device comes from
// hipGetSymbolAddress, taken once per device with that device current and
// kept in g_scratch_ptr.  Every call on a GPU shares that GPU's array and
// finds it as the previous call left it; the QR driver clears what must start
// at zero (status words, flags) with kernels at the start of each call, as it
// does for a caller's workspace.  Calls that overlap on one GPU (two streams,
// two threads) race on the array: the caller serialises them.
#include <mutex>

#include "mak_common.hpp"
#include "mak_kernels.hpp"
#include "mak_mi355x_internal.hpp"

namespace mak {

namespace {

constexpr size_t kScratchBytes = (size_t)64 << 20;  // 64 MiB per GPU
constexpr int kMaxDevices = 64;

__device__ __align__(256) unsigned char scratch[kScratchBytes];

std::once_flag g_scratch_once[kMaxDevices];
void* g_scratch_ptr[kMaxDevices];

// The array's device address on the stream's device (nullptr on failure).
void* scratch_on(hipStream_t s) {
  int dev = -1;
  if (s != nullptr) {
    if (hipStreamGetDevice(s, &dev) != hipSuccess) return nullptr;
  } else if (hipGetDevice(&dev) != hipSuccess) {
    return nullptr;
  }
  if (dev < 0 || dev >= kMaxDevices) return nullptr;
  std::call_once(g_scratch_once[dev], [dev] {
    int cur = -1;
    if (hipGetDevice(&cur) != hipSuccess) return;
    if (cur != dev && hipSetDevice(dev) != hipSuccess) return;
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) == hipSuccess) g_scratch_ptr[dev] = p;
    if (cur != dev) (void)hipSetDevice(cur);
  });
  return g_scratch_ptr[dev];
}

}  // namespace

}  // namespace mak

extern "C" {

int mak_qr_compact_noalloc_d(void* s, int m, int n, double* A, int lda, double* Q, int ldq, double* R, int ldr,
                             int positive) {
  if (m < 0) return -2;
  if (n < 0) return -3;
  if (mak::qr_workspace_bytes<double>(m, n, 0) > mak::kScratchBytes) return 7;  // beyond the static scratch
  void* ws = mak::scratch_on((hipStream_t)s);
  if (ws == nullptr) return 1;
  return mak::qr_driver<double>((hipStream_t)s, 0, m, n, A, lda, Q, ldq, R, ldr, positive, ws, mak::kScratchBytes,
                                nullptr);
}

}  // extern "C"
