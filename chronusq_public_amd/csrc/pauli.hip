// Pauli-spinor component operations on the device: SpinScatter /
// SpinGather between a 2NB x 2NB spinor matrix and its S, Z, Y, X
// components (src/cqlinalg/blasutil.cxx:71-189), the Pauli energy traces
// of SingleSlater::computeEnergy (include/singleslater/quantum.hpp:147-212,
// OperatorTrace = Re dot, include/singleslater/quantum/base.hpp:46-51) and
// the NaN test of PauliSpinorMatrices::hasNaN
// (include/matrix/paulispinormatrices.hpp:383-388).
#include <hip/hip_runtime.h>

#include <cmath>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "cqx.h"
#include "ctx_access.hpp"

namespace {

// element types: real matrices as double, complex as (re, im) pairs
struct Cx {
  double re, im;
};
__device__ __forceinline__ Cx cx_add(Cx a, Cx b) { return {a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ Cx cx_sub(Cx a, Cx b) { return {a.re - b.re, a.im - b.im}; }
__device__ __forceinline__ Cx cx_scale(double s, Cx a) { return {s * a.re, s * a.im}; }

template <bool CPLX>
__device__ __forceinline__ Cx ld(const double *p, size_t i) {
  return CPLX ? Cx{p[2 * i], p[2 * i + 1]} : Cx{p[i], 0.0};
}
template <bool CPLX>
__device__ __forceinline__ void st(double *p, size_t i, Cx v) {
  if (CPLX) {
    p[2 * i] = v.re;
    p[2 * i + 1] = v.im;
  } else {
    p[i] = v.re;
  }
}

// y times "the Y factor": i for complex matrices, -1 for real ones
// (ComplexScale, blasutil.cxx:67-69)
template <bool CPLX>
__device__ __forceinline__ Cx yfact(Cx v) {
  return CPLX ? Cx{-v.im, v.re} : Cx{-v.re, 0.0};
}

// S = AA + BB, Z = AA - BB (S = 2 AA, Z = 0 when BB equals AA);
// Y = f (AB - BA), X = AB + BA (both zero when the off-diagonal blocks are)
template <bool CPLX>
__global__ void spin_scatter_kernel(int nb, const double *A, int lda, double *S, double *Z, double *Y, double *X,
                                    int zero_abba, int bb_eq_aa) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)nb * nb) return;
  const int i = (int)(t % nb), j = (int)(t / nb);
  const size_t aa = (size_t)i + (size_t)j * lda, ab = (size_t)i + (size_t)(j + nb) * lda;
  const size_t ba = (size_t)(i + nb) + (size_t)j * lda, bb = (size_t)(i + nb) + (size_t)(j + nb) * lda;
  const Cx vaa = ld<CPLX>(A, aa);
  if (bb_eq_aa) {
    if (S) st<CPLX>(S, t, cx_scale(2.0, vaa));
    if (Z) st<CPLX>(Z, t, Cx{0.0, 0.0});
  } else {
    const Cx vbb = ld<CPLX>(A, bb);
    if (S) st<CPLX>(S, t, cx_add(vaa, vbb));
    if (Z) st<CPLX>(Z, t, cx_sub(vaa, vbb));
  }
  if (zero_abba) {
    if (Y) st<CPLX>(Y, t, Cx{0.0, 0.0});
    if (X) st<CPLX>(X, t, Cx{0.0, 0.0});
  } else {
    const Cx vab = ld<CPLX>(A, ab), vba = ld<CPLX>(A, ba);
    if (Y) st<CPLX>(Y, t, yfact<CPLX>(cx_sub(vab, vba)));
    if (X) st<CPLX>(X, t, cx_add(vab, vba));
  }
}

// AA = (S + Z) / 2, BB = (S - Z) / 2 (S / 2 both when Z is zero);
// BA = (X + f Y) / 2, AB = (X - f Y) / 2 (zero when X and Y are)
template <bool CPLX>
__global__ void spin_gather_kernel(int nb, double *A, int lda, const double *S, const double *Z, const double *Y,
                                   const double *X, int zero_xy, int zero_z) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)nb * nb) return;
  const int i = (int)(t % nb), j = (int)(t / nb);
  const size_t aa = (size_t)i + (size_t)j * lda, ab = (size_t)i + (size_t)(j + nb) * lda;
  const size_t ba = (size_t)(i + nb) + (size_t)j * lda, bb = (size_t)(i + nb) + (size_t)(j + nb) * lda;
  const Cx vs = ld<CPLX>(S, t);
  if (zero_z) {
    st<CPLX>(A, aa, cx_scale(0.5, vs));
    st<CPLX>(A, bb, cx_scale(0.5, vs));
  } else {
    const Cx vz = ld<CPLX>(Z, t);
    st<CPLX>(A, aa, cx_scale(0.5, cx_add(vs, vz)));
    st<CPLX>(A, bb, cx_scale(0.5, cx_sub(vs, vz)));
  }
  if (zero_xy) {
    st<CPLX>(A, ba, Cx{0.0, 0.0});
    st<CPLX>(A, ab, Cx{0.0, 0.0});
  } else {
    const Cx vx = ld<CPLX>(X, t), fy = yfact<CPLX>(ld<CPLX>(Y, t));
    st<CPLX>(A, ba, cx_scale(0.5, cx_add(vx, fy)));
    st<CPLX>(A, ab, cx_scale(0.5, cx_sub(vx, fy)));
  }
}

constexpr int kRedThreads = 256;
constexpr int kRedBlocks = 1024;

// Partial sums of Re sum_k conj(a_k) b_k over n elements (2n doubles when
// complex), one per block: a fixed-order reduction, the same result on
// every run.
template <bool CPLX>
__global__ __launch_bounds__(kRedThreads) void re_dot_kernel(const double *a, const double *b, long long n,
                                                             double *part) {
  double s = 0.0;
  for (long long k = (long long)blockIdx.x * kRedThreads + threadIdx.x; k < n; k += (long long)gridDim.x * kRedThreads)
    s += CPLX ? a[2 * k] * b[2 * k] + a[2 * k + 1] * b[2 * k + 1] : a[k] * b[k];
  __shared__ double red[kRedThreads];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = kRedThreads / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// Per block: how many of the n doubles are not finite
__global__ __launch_bounds__(kRedThreads) void nonfinite_kernel(const double *a, long long n, unsigned *part) {
  unsigned c = 0;
  for (long long k = (long long)blockIdx.x * kRedThreads + threadIdx.x; k < n; k += (long long)gridDim.x * kRedThreads)
    c += isfinite(a[k]) ? 0u : 1u;
  __shared__ unsigned red[kRedThreads];
  red[threadIdx.x] = c;
  __syncthreads();
  for (int w = kRedThreads / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

struct Err : std::runtime_error {
  int code;
  Err(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define HIPDO(x)                                                                                       \
  do {                                                                                                 \
    hipError_t e_ = (x);                                                                               \
    if (e_ != hipSuccess) throw Err(CQX_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_));      \
  } while (0)

// A device buffer for the length of one call.
struct Tmp {
  void *p = nullptr;
  explicit Tmp(size_t bytes) {
    if (hipMalloc(&p, bytes ? bytes : 1) != hipSuccess) throw Err(CQX_ERR_MEM, "hipMalloc failed");
  }
  ~Tmp() {
    if (p) (void)hipFree(p);
  }
  double *d() const { return static_cast<double *>(p); }
};

// A matrix argument on the device: the caller's pointer (mem = CQX_DEVICE)
// or a staged copy of host memory.
struct DevArg {
  std::unique_ptr<Tmp> tmp;
  const double *p = nullptr;
  DevArg(const void *src, size_t bytes, int mem, hipStream_t s) {
    if (!src) return;
    if (mem == CQX_DEVICE) {
      p = static_cast<const double *>(src);
      return;
    }
    tmp = std::make_unique<Tmp>(bytes);
    HIPDO(hipMemcpyAsync(tmp->p, src, bytes, hipMemcpyHostToDevice, s));
    p = tmp->d();
  }
};

// An output matrix: written on the device, copied back for host memory.
struct DevOut {
  std::unique_ptr<Tmp> tmp;
  void *dst = nullptr;
  double *p = nullptr;
  size_t bytes = 0;
  int mem = CQX_DEVICE;
  DevOut(void *d, size_t b, int m) : dst(d), bytes(b), mem(m) {
    if (!d) return;
    if (m == CQX_DEVICE) {
      p = static_cast<double *>(d);
    } else {
      tmp = std::make_unique<Tmp>(b);
      p = tmp->d();
    }
  }
  void back(hipStream_t s) {
    if (dst && mem != CQX_DEVICE) HIPDO(hipMemcpyAsync(dst, p, bytes, hipMemcpyDeviceToHost, s));
  }
};

double re_dot(const double *a, const double *b, long long n, bool cplx, hipStream_t s) {
  Tmp part(sizeof(double) * kRedBlocks);
  if (cplx)
    hipLaunchKernelGGL(re_dot_kernel<true>, dim3(kRedBlocks), dim3(kRedThreads), 0, s, a, b, n, part.d());
  else
    hipLaunchKernelGGL(re_dot_kernel<false>, dim3(kRedBlocks), dim3(kRedThreads), 0, s, a, b, n, part.d());
  HIPDO(hipGetLastError());
  std::vector<double> h(kRedBlocks);
  HIPDO(hipMemcpyAsync(h.data(), part.p, sizeof(double) * kRedBlocks, hipMemcpyDeviceToHost, s));
  HIPDO(hipStreamSynchronize(s));
  double r = 0.0;
  for (double v : h) r += v;
  return r;
}

template <class F>
int guarded(cqx_ctx *ctx, F &&f) {
  try {
    HIPDO(hipSetDevice(cqx::ctx_device(ctx)));
    f(cqx::ctx_stream(ctx));
    return CQX_OK;
  } catch (const Err &e) {
    return cqx::ctx_fail(ctx, e.code, e.what());
  } catch (const std::exception &e) {
    return cqx::ctx_fail(ctx, CQX_ERR_HIP, e.what());
  }
}

}  // namespace

extern "C" {

int cqx_spin_scatter(cqx_ctx *ctx, int nb, const void *A, int lda, void *S, void *Z, void *Y, void *X,
                     int is_complex, int flags, int mem) {
  if (!ctx || nb <= 0 || !A || lda < 2 * nb || (mem != CQX_HOST && mem != CQX_DEVICE))
    return cqx::ctx_fail(ctx, CQX_ERR_ARG, "cqx_spin_scatter: bad arguments");
  return guarded(ctx, [&](hipStream_t s) {
    const size_t ew = is_complex ? 16 : 8, nb2 = (size_t)nb * nb;
    DevArg a(A, ew * (size_t)lda * 2 * nb, mem, s);
    DevOut o[4] = {{S, ew * nb2, mem}, {Z, ew * nb2, mem}, {Y, ew * nb2, mem}, {X, ew * nb2, mem}};
    const int zero_abba = (flags & CQX_SPIN_ZERO_ABBA) != 0, bbaa = (flags & CQX_SPIN_BB_EQ_AA) != 0;
    const unsigned grid = (unsigned)((nb2 + 255) / 256);
    if (is_complex)
      hipLaunchKernelGGL(spin_scatter_kernel<true>, dim3(grid), dim3(256), 0, s, nb, a.p, lda, o[0].p, o[1].p, o[2].p,
                         o[3].p, zero_abba, bbaa);
    else
      hipLaunchKernelGGL(spin_scatter_kernel<false>, dim3(grid), dim3(256), 0, s, nb, a.p, lda, o[0].p, o[1].p,
                         o[2].p, o[3].p, zero_abba, bbaa);
    HIPDO(hipGetLastError());
    for (auto &x : o) x.back(s);
    HIPDO(hipStreamSynchronize(s));
  });
}

int cqx_spin_gather(cqx_ctx *ctx, int nb, void *A, int lda, const void *S, const void *Z, const void *Y,
                    const void *X, int is_complex, int flags, int mem) {
  const int zero_xy = (flags & CQX_SPIN_ZERO_XY) != 0, zero_z = (flags & CQX_SPIN_ZERO_Z) != 0;
  if (!ctx || nb <= 0 || !A || lda < 2 * nb || !S || (!zero_z && !Z) || (!zero_xy && (!X || !Y)) ||
      (mem != CQX_HOST && mem != CQX_DEVICE))
    return cqx::ctx_fail(ctx, CQX_ERR_ARG, "cqx_spin_gather: bad arguments");
  return guarded(ctx, [&](hipStream_t s) {
    const size_t ew = is_complex ? 16 : 8, nb2 = (size_t)nb * nb, abytes = ew * (size_t)lda * 2 * nb;
    DevArg in[4] = {{S, ew * nb2, mem, s}, {zero_z ? nullptr : Z, ew * nb2, mem, s},
                    {zero_xy ? nullptr : Y, ew * nb2, mem, s}, {zero_xy ? nullptr : X, ew * nb2, mem, s}};
    DevOut a(A, abytes, mem);
    // rows beyond 2 NB of a padded leading dimension keep what they held
    if (mem != CQX_DEVICE && lda > 2 * nb) HIPDO(hipMemcpyAsync(a.p, A, abytes, hipMemcpyHostToDevice, s));
    const unsigned grid = (unsigned)((nb2 + 255) / 256);
    if (is_complex)
      hipLaunchKernelGGL(spin_gather_kernel<true>, dim3(grid), dim3(256), 0, s, nb, a.p, lda, in[0].p, in[1].p,
                         in[2].p, in[3].p, zero_xy, zero_z);
    else
      hipLaunchKernelGGL(spin_gather_kernel<false>, dim3(grid), dim3(256), 0, s, nb, a.p, lda, in[0].p, in[1].p,
                         in[2].p, in[3].p, zero_xy, zero_z);
    HIPDO(hipGetLastError());
    a.back(s);
    HIPDO(hipStreamSynchronize(s));
  });
}

int cqx_compute_energy(cqx_ctx *ctx, int nb, int ncomp, const void *const *D, const void *const *H,
                       const void *const *G, int is_complex, int mem, double e_nuc, double *out) {
  if (!ctx || nb <= 0 || (ncomp != 1 && ncomp != 2 && ncomp != 4) || !D || !out ||
      (mem != CQX_HOST && mem != CQX_DEVICE))
    return cqx::ctx_fail(ctx, CQX_ERR_ARG, "cqx_compute_energy: bad arguments");
  for (int c = 0; c < ncomp; ++c)
    if (!D[c]) return cqx::ctx_fail(ctx, CQX_ERR_ARG, "cqx_compute_energy: missing density component");
  return guarded(ctx, [&](hipStream_t s) {
    const size_t nb2 = (size_t)nb * nb, bytes = nb2 * (is_complex ? 16 : 8);
    double e1 = 0.0, e2 = 0.0;
    for (int c = 0; c < ncomp; ++c) {
      DevArg d(D[c], bytes, mem, s);
      if (H && H[c]) {
        DevArg h(H[c], bytes, mem, s);
        e1 += re_dot(d.p, h.p, (long long)nb2, is_complex != 0, s);
      }
      if (G && G[c]) {
        DevArg g(G[c], bytes, mem, s);
        e2 += re_dot(d.p, g.p, (long long)nb2, is_complex != 0, s);
      }
    }
    e1 *= 0.5;
    e2 *= 0.25;
    if (!std::isfinite(e1) || !std::isfinite(e2)) throw Err(CQX_ERR_ARG, "cqx_compute_energy: energy is not finite");
    out[0] = e1;
    out[1] = e2;
    out[2] = e1 + e2 + e_nuc;
  });
}

int cqx_count_nonfinite(cqx_ctx *ctx, const double *M, size_t n, int mem, size_t *count) {
  if (!ctx || (n && !M) || !count || (mem != CQX_HOST && mem != CQX_DEVICE))
    return cqx::ctx_fail(ctx, CQX_ERR_ARG, "cqx_count_nonfinite: bad arguments");
  if (n == 0) {
    *count = 0;
    return CQX_OK;
  }
  return guarded(ctx, [&](hipStream_t s) {
    DevArg m(M, sizeof(double) * n, mem, s);
    Tmp part(sizeof(unsigned) * kRedBlocks);
    hipLaunchKernelGGL(nonfinite_kernel, dim3(kRedBlocks), dim3(kRedThreads), 0, s, m.p, (long long)n,
                       static_cast<unsigned *>(part.p));
    HIPDO(hipGetLastError());
    std::vector<unsigned> h(kRedBlocks);
    HIPDO(hipMemcpyAsync(h.data(), part.p, sizeof(unsigned) * kRedBlocks, hipMemcpyDeviceToHost, s));
    HIPDO(hipStreamSynchronize(s));
    size_t c = 0;
    for (unsigned v : h) c += v;
    *count = c;
  });
}

}  // extern "C"
