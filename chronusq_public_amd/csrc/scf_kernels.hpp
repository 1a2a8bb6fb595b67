// Elementwise device kernels of the SCF driver (scf_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

namespace cqx {
// G = 2 J - K, F = H + G (n elements)
hipError_t launch_scf_fock(const double *H, const double *J, const double *K, double *G, double *F,
                           size_t n, hipStream_t s);
// M[:, j] *= f[j] for j < ncol (column-major, ld = nrow); f on the device
hipError_t launch_scale_cols(double *M, int nrow, int ncol, const double *f, hipStream_t s);
// E = A - A^T (n x n)
hipError_t launch_antisym(const double *A, double *E, int n, hipStream_t s);
// part[b] = non-finite entries among the n doubles of a seen by block b
// (kNonfiniteBlocks partial counts)
constexpr int kNonfiniteBlocks = 256;
hipError_t launch_count_nonfinite(const double *a, size_t n, unsigned *part, hipStream_t s);
// C = A - B (n elements)
hipError_t launch_sub(const double *A, const double *B, double *C, size_t n, hipStream_t s);
}  // namespace cqx
