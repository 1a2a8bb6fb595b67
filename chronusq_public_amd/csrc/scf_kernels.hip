// Elementwise kernels of the device-resident SCF loop; the dense steps
// (GEMM, eigensolve) go to rocBLAS / rocSOLVER.
#include "scf_kernels.hpp"

namespace cqx {

namespace {

inline unsigned nblk(size_t n) { return (unsigned)((n + 255) / 256); }

__global__ void scf_fock_kernel(const double *H, const double *J, const double *K, double *G,
                                double *F, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double g = 2.0 * J[i] - K[i];
  G[i] = g;
  F[i] = H[i] + g;
}

__global__ void scale_cols_kernel(double *M, int nrow, int ncol, const double *f) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)nrow * ncol) return;
  M[i] *= f[i / nrow];
}

__global__ void antisym_kernel(const double *A, double *E, int n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)n * n) return;
  const size_t r = i % n, c = i / n;
  E[i] = A[i] - A[c + r * n];
}

__global__ void sub_kernel(const double *A, const double *B, double *C, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) C[i] = A[i] - B[i];
}

// per block: how many of the n doubles are not finite
__global__ __launch_bounds__(256) void count_nonfinite_kernel(const double *a, size_t n, unsigned *part) {
  unsigned c = 0;
  for (size_t k = (size_t)blockIdx.x * 256 + threadIdx.x; k < n; k += (size_t)gridDim.x * 256)
    c += isfinite(a[k]) ? 0u : 1u;
  __shared__ unsigned red[256];
  red[threadIdx.x] = c;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

}  // namespace

hipError_t launch_count_nonfinite(const double *a, size_t n, unsigned *part, hipStream_t s) {
  hipLaunchKernelGGL(count_nonfinite_kernel, dim3(kNonfiniteBlocks), dim3(256), 0, s, a, n, part);
  return hipGetLastError();
}

hipError_t launch_scf_fock(const double *H, const double *J, const double *K, double *G, double *F,
                           size_t n, hipStream_t s) {
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(scf_fock_kernel, dim3(nblk(n)), dim3(256), 0, s, H, J, K, G, F, n);
  return hipGetLastError();
}

hipError_t launch_scale_cols(double *M, int nrow, int ncol, const double *f, hipStream_t s) {
  const size_t n = (size_t)nrow * ncol;
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(scale_cols_kernel, dim3(nblk(n)), dim3(256), 0, s, M, nrow, ncol, f);
  return hipGetLastError();
}

hipError_t launch_antisym(const double *A, double *E, int n, hipStream_t s) {
  const size_t m = (size_t)n * n;
  if (!m) return hipSuccess;
  hipLaunchKernelGGL(antisym_kernel, dim3(nblk(m)), dim3(256), 0, s, A, E, n);
  return hipGetLastError();
}

hipError_t launch_sub(const double *A, const double *B, double *C, size_t n, hipStream_t s) {
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(sub_kernel, dim3(nblk(n)), dim3(256), 0, s, A, B, C, n);
  return hipGetLastError();
}

}  // namespace cqx
