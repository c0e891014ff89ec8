// Hermitian eigendecomposition on the device, for the polar-based SVD and the
// Sylvester solves of the polar derivatives: rocSOLVER's syevd / heevd
// (divide and conquer), strided-batched.  rocSOLVER and rocBLAS are opened at
// the first call (dlopen) and only their entry points are kept, so the QR /
// polar path of this library carries no link-time dependency on them; the
// function table is written once under std::call_once and only read after.
#include <dlfcn.h>

#include <mutex>

#include <rocsolver/rocsolver.h>

#include "mak_common.hpp"
#include "mak_kernels.hpp"
#include "mak_mi355x_internal.hpp"

namespace mak {

namespace {

struct RocsolverApi {
  decltype(&rocblas_create_handle) create = nullptr;
  decltype(&rocblas_destroy_handle) destroy = nullptr;
  decltype(&rocblas_set_stream) set_stream = nullptr;
  decltype(&rocsolver_dsyevd_strided_batched) dsyevd = nullptr;
  decltype(&rocsolver_zheevd_strided_batched) zheevd = nullptr;
  bool ok = false;
};

const RocsolverApi& rocsolver_api() {
  static RocsolverApi api;
  static std::once_flag once;
  std::call_once(once, [] {
    void* blas = dlopen("librocblas.so.5", RTLD_NOW | RTLD_GLOBAL);
    if (!blas) blas = dlopen("librocblas.so", RTLD_NOW | RTLD_GLOBAL);
    void* solver = dlopen("librocsolver.so.0", RTLD_NOW | RTLD_GLOBAL);
    if (!solver) solver = dlopen("librocsolver.so", RTLD_NOW | RTLD_GLOBAL);
    if (!blas || !solver) return;
    api.create = reinterpret_cast<decltype(api.create)>(dlsym(blas, "rocblas_create_handle"));
    api.destroy = reinterpret_cast<decltype(api.destroy)>(dlsym(blas, "rocblas_destroy_handle"));
    api.set_stream = reinterpret_cast<decltype(api.set_stream)>(dlsym(blas, "rocblas_set_stream"));
    api.dsyevd = reinterpret_cast<decltype(api.dsyevd)>(dlsym(solver, "rocsolver_dsyevd_strided_batched"));
    api.zheevd = reinterpret_cast<decltype(api.zheevd)>(dlsym(solver, "rocsolver_zheevd_strided_batched"));
    api.ok = api.create && api.destroy && api.set_stream && api.dsyevd && api.zheevd;
  });
  return api;
}

}  // namespace

bool eigh_available() { return rocsolver_api().ok; }

size_t eigh_workspace_bytes(int n, int batch) {
  // E (off-diagonal, n per matrix) and the info words
  return ((size_t)n * batch * sizeof(double) + 255) / 256 * 256 + (size_t)batch * sizeof(int) + 256;
}

// A (batch matrices, n x n, ld lda, stride strideA): on entry Hermitian (the
// lower triangle is read), on exit the eigenvectors as columns; D (stride
// strideD) the eigenvalues in ascending order.  Returns 0, 3 (rocSOLVER not
// available), 1 (call failed) or 4 + i when the i-th matrix's iteration did
// not converge (read back from the device: one synchronisation).
template <typename T>
int eigh_batched(hipStream_t s, int n, T* A, int64_t lda, int64_t strideA, double* D, int64_t strideD, int batch,
                 char* ws) {
  if (n <= 0 || batch <= 0) return 0;
  const RocsolverApi& api = rocsolver_api();
  if (!api.ok) return 3;
  double* E = reinterpret_cast<double*>(ws);
  int* info = reinterpret_cast<int*>(ws + ((size_t)n * batch * sizeof(double) + 255) / 256 * 256);
  rocblas_handle h = nullptr;
  if (api.create(&h) != rocblas_status_success) return 1;
  int st = 0;
  if (api.set_stream(h, s) != rocblas_status_success) st = 1;
  if (st == 0) {
    rocblas_status r;
    if constexpr (ScalarTraits<T>::is_complex)
      r = api.zheevd(h, rocblas_evect_original, rocblas_fill_lower, n, reinterpret_cast<rocblas_double_complex*>(A),
                     (int)lda, strideA, D, strideD, E, n, info, batch);
    else
      r = api.dsyevd(h, rocblas_evect_original, rocblas_fill_lower, n, reinterpret_cast<double*>(A), (int)lda,
                     strideA, D, strideD, E, n, info, batch);
    if (r != rocblas_status_success) st = 1;
  }
  if (st == 0) {
    std::vector<int> hinfo((size_t)batch, 0);
    if (hipMemcpyAsync(hinfo.data(), info, sizeof(int) * batch, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
      st = 1;
    for (int b = 0; st == 0 && b < batch; ++b)
      if (hinfo[b] != 0) st = 4 + b;
  }
  api.destroy(h);
  return st;
}

template int eigh_batched<double>(hipStream_t, int, double*, int64_t, int64_t, double*, int64_t, int, char*);
template int eigh_batched<zd>(hipStream_t, int, zd*, int64_t, int64_t, double*, int64_t, int, char*);

}  // namespace mak
