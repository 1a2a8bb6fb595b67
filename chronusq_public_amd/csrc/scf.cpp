// Device-resident closed-shell SCF over the direct J/K engine.
//
// The loop is the reference's conventional SCF for a real RHF reference:
// form the Fock matrix from the current density, evaluate the energy and
// the convergence tests, then CDIIS-extrapolate the Fock matrix, diagonalise
// it in the Lowdin basis and form the next density
// (include/orbitalmodifiernew/orbitaloptimizernew/impl.hpp:59-103 and
// :266-330, conventionalSCFnew/extrap.hpp:36-123).  Matrices stay in HBM;
// dense products use rocBLAS GEMM (matrix cores), the eigensolves rocSOLVER
// syevd, and the Fock builds cqx_two_body_contract on the same stream.
//
// The superposition-of-atomic-densities guess solves each element once as
// a spin-averaged atom with fractional occupation of its open shell (the
// same loop with fractional occupations) and places the atomic densities
// on the block diagonal (guess.hpp:297-326 in the reference builds its SAD
// guess from atomic SCFs in the same way).
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "cqx.h"
#include "ctx_access.hpp"
#include "kernels.hpp"
#include "onebody.hpp"
#include "scf_access.hpp"
#include "scf_kernels.hpp"

namespace {

struct ScfError : std::runtime_error {
  int code;
  ScfError(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define HIPOK(x)                                                                             \
  do {                                                                                       \
    hipError_t e_ = (x);                                                                     \
    if (e_ != hipSuccess) throw ScfError(CQX_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define BLASOK(x)                                                                            \
  do {                                                                                       \
    rocblas_status s_ = (x);                                                                 \
    if (s_ != rocblas_status_success)                                                        \
      throw ScfError(CQX_ERR_HIP, std::string(#x) + ": rocBLAS status " + std::to_string((int)s_)); \
  } while (0)

#define CQXOK(c, x)                                                                          \
  do {                                                                                       \
    int r_ = (x);                                                                            \
    if (r_ != CQX_OK) throw ScfError(r_, std::string(#x) + ": " + cqx_last_error(c));       \
  } while (0)

template <class T>
struct Dev {
  T *p = nullptr;
  size_t n = 0;
  void alloc(size_t m) {
    if (m <= n) return;
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
    HIPOK(hipMalloc((void **)&p, sizeof(T) * std::max<size_t>(m, 1)));
    n = m;
  }
  ~Dev() {
    if (p) (void)hipFree(p);
  }
};

double now_ms() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// Solve the CDIIS system [B -1; -1 0][c; l] = [0; -1] by Gaussian
// elimination with partial pivoting; false when it is singular.
bool diis_solve(int n, const std::vector<double> &B, std::vector<double> &c) {
  const int m = n + 1;
  std::vector<double> A((size_t)m * (m + 1), 0.0);
  // scale by the largest diagonal (as Pulay extrapolations usually do)
  double sc = 0.0;
  for (int i = 0; i < n; ++i) sc = std::max(sc, std::fabs(B[(size_t)i * n + i]));
  if (!(sc > 0.0)) return false;
  for (int i = 0; i < n; ++i) {
    for (int j = 0; j < n; ++j) A[(size_t)i * (m + 1) + j] = B[(size_t)i * n + j] / sc;
    A[(size_t)i * (m + 1) + n] = -1.0;
    A[(size_t)n * (m + 1) + i] = -1.0;
  }
  A[(size_t)n * (m + 1) + m] = -1.0;
  for (int k = 0; k < m; ++k) {
    int piv = k;
    for (int i = k + 1; i < m; ++i)
      if (std::fabs(A[(size_t)i * (m + 1) + k]) > std::fabs(A[(size_t)piv * (m + 1) + k])) piv = i;
    if (std::fabs(A[(size_t)piv * (m + 1) + k]) < 1e-14) return false;
    if (piv != k)
      for (int j = 0; j <= m; ++j) std::swap(A[(size_t)k * (m + 1) + j], A[(size_t)piv * (m + 1) + j]);
    for (int i = 0; i < m; ++i) {
      if (i == k) continue;
      const double f = A[(size_t)i * (m + 1) + k] / A[(size_t)k * (m + 1) + k];
      if (f == 0.0) continue;
      for (int j = k; j <= m; ++j) A[(size_t)i * (m + 1) + j] -= f * A[(size_t)k * (m + 1) + j];
    }
  }
  c.resize(n);
  for (int i = 0; i < n; ++i) c[i] = A[(size_t)i * (m + 1) + m] / A[(size_t)i * (m + 1) + i];
  return true;
}

}  // namespace

struct cqx_scf {
  cqx_ctx *ctx = nullptr;
  hipStream_t stream = nullptr;
  rocblas_handle blas = nullptr;
  cqx_scf_opts o{};
  int nb = 0;
  double e_nuc = 0.0;
  bool frac = false;      // fractional occupation (spin-averaged atoms of the SAD guess)
  double nelec = 0.0;     // electrons, fractional mode
  Dev<double> S, H, X, D, Dp, J, K, G, F, Fx, W1, W2, W3, eps, ework, fscale, Fh, Eh;
  Dev<rocblas_int> info;
  Dev<unsigned> nonfinite;  // partial counts of the Fock matrix's check
  std::vector<double> B;  // CDIIS error overlaps, keep x keep
  std::vector<double> eps_host;
  int iter = 0, nstored = 0;
  bool converged = false;
  double E = 0, Eprev = 0, e1 = 0, e2 = 0, rmsd = 0, maxd = 0, dE = 0;
  double ms_setup = 0, ms_fock = 0, ms_total = 0;
  std::string err;

  ~cqx_scf() {
    if (blas) rocblas_destroy_handle(blas);
  }

  size_t n2() const { return (size_t)nb * nb; }

  void gemm(rocblas_operation ta, rocblas_operation tb, int m, int n, int k, double alpha, const double *A,
            const double *Bm, double beta, double *C) {
    BLASOK(rocblas_dgemm(blas, ta, tb, m, n, k, &alpha, A, nb, Bm, nb, &beta, C, nb));
  }

  double dot(const double *a, const double *b) {
    double r = 0.0;
    BLASOK(rocblas_ddot(blas, (rocblas_int)n2(), a, 1, b, 1, &r));
    return r;
  }

  // J = J[D], K = K[D] through the direct J/K engine on this SCF's context
  void fock_jk() {
    HIPOK(cqx::launch_zero(J.p, n2(), stream));
    HIPOK(cqx::launch_zero(K.p, n2(), stream));
    cqx_contraction c[2] = {{D.p, J.p, 1, CQX_COULOMB}, {D.p, K.p, 1, CQX_EXCHANGE}};
    CQXOK(ctx, cqx_two_body_contract(ctx, 1, 2, c, 0, CQX_DEVICE));
  }

  // G = 2 J - K and F_S = H_S + G; a Fock matrix with a non-finite entry
  // stops the SCF (PauliSpinorMatrices::hasNaN, paulispinormatrices.hpp:383-388)
  void fock_from_jk() {
    HIPOK(cqx::launch_scf_fock(H.p, J.p, K.p, G.p, F.p, n2(), stream));
    HIPOK(cqx::launch_count_nonfinite(F.p, n2(), nonfinite.p, stream));
    std::vector<unsigned> part(cqx::kNonfiniteBlocks);
    HIPOK(hipMemcpyAsync(part.data(), nonfinite.p, sizeof(unsigned) * part.size(), hipMemcpyDeviceToHost, stream));
    HIPOK(hipStreamSynchronize(stream));
    size_t bad = 0;
    for (unsigned v : part) bad += v;
    if (bad) throw ScfError(CQX_ERR_ARG, "the Fock matrix has " + std::to_string(bad) + " non-finite entries");
  }

  // eigenvectors of Fm in the Lowdin basis -> C = X V in W3, eigenvalues
  // of Fm / 2 in eps_host (Fm must not be one of the work buffers W1-W3)
  void diagonalise(const double *Fm) {
    if (Fm == W1.p || Fm == W2.p || Fm == W3.p) throw ScfError(CQX_ERR_ARG, "diagonalise: aliased input");
    gemm(rocblas_operation_none, rocblas_operation_none, nb, nb, nb, 1.0, Fm, X.p, 0.0, W1.p);
    gemm(rocblas_operation_transpose, rocblas_operation_none, nb, nb, nb, 1.0, X.p, W1.p, 0.0, W2.p);
    BLASOK(rocsolver_dsyevd(blas, rocblas_evect_original, rocblas_fill_upper, nb, W2.p, nb, eps.p, ework.p,
                            info.p));
    gemm(rocblas_operation_none, rocblas_operation_none, nb, nb, nb, 1.0, X.p, W2.p, 0.0, W3.p);
    eps_host.resize(nb);
    rocblas_int inf = 0;
    HIPOK(hipMemcpyAsync(eps_host.data(), eps.p, sizeof(double) * nb, hipMemcpyDeviceToHost, stream));
    HIPOK(hipMemcpyAsync(&inf, info.p, sizeof(inf), hipMemcpyDeviceToHost, stream));
    HIPOK(hipStreamSynchronize(stream));
    if (inf != 0) throw ScfError(CQX_ERR_HIP, "syevd failed to converge (info " + std::to_string(inf) + ")");
    for (double &e : eps_host) e *= 0.5;  // F_S = 2 F_alpha
  }

  // D_S = sum_i n_i C_i C_i^T: n_i = 2 for the nocc lowest orbitals, or the
  // aufbau fractional occupations of a spin-averaged atom
  void form_density() {
    std::vector<double> occ;
    if (!frac) {
      occ.assign(o.nocc, 2.0);
    } else {
      double left = nelec;
      for (int i = 0; i < nb && left > 1e-12;) {
        int j = i;
        while (j < nb && eps_host[j] - eps_host[i] < 1e-4 * std::max(1.0, std::fabs(eps_host[i]))) ++j;
        const int g = j - i;
        const double each = left >= 2.0 * g ? 2.0 : left / g;
        for (int k = i; k < j; ++k) occ.push_back(each);
        left -= each * g;
        i = j;
      }
    }
    const int nocc = (int)occ.size();
    if (nocc == 0) {
      HIPOK(cqx::launch_zero(D.p, n2(), stream));
      return;
    }
    if (nocc > nb) throw ScfError(CQX_ERR_ARG, "more occupied orbitals than basis functions");
    std::vector<double> f(nocc);
    for (int i = 0; i < nocc; ++i) f[i] = std::sqrt(occ[i]);
    HIPOK(hipMemcpyAsync(fscale.p, f.data(), sizeof(double) * nocc, hipMemcpyHostToDevice, stream));
    HIPOK(hipMemcpyAsync(W1.p, W3.p, sizeof(double) * nb * nocc, hipMemcpyDeviceToDevice, stream));
    HIPOK(cqx::launch_scale_cols(W1.p, nb, nocc, fscale.p, stream));
    gemm(rocblas_operation_none, rocblas_operation_transpose, nb, nb, nocc, 1.0, W1.p, W1.p, 0.0, D.p);
    HIPOK(hipStreamSynchronize(stream));  // f lives on this stack frame
  }

  // CDIIS: store F and the Lowdin-basis commutator X^T (F D S - S D F) X,
  // extrapolate F from the stored ones (nothing to combine on iteration 0);
  // the extrapolated matrix goes to Fx, which no later step writes before
  // the diagonalisation has read it
  const double *diis() {
    const int keep = o.diis_keep;
    if (keep <= 1) return F.p;
    const int slot = (iter - 1) % keep;
    double *Fs = Fh.p + (size_t)slot * n2(), *Es = Eh.p + (size_t)slot * n2();
    gemm(rocblas_operation_none, rocblas_operation_none, nb, nb, nb, 1.0, F.p, D.p, 0.0, W1.p);
    gemm(rocblas_operation_none, rocblas_operation_none, nb, nb, nb, 1.0, W1.p, S.p, 0.0, W2.p);
    HIPOK(cqx::launch_antisym(W2.p, W1.p, nb, stream));  // FDS - (FDS)^T = FDS - SDF
    gemm(rocblas_operation_transpose, rocblas_operation_none, nb, nb, nb, 1.0, X.p, W1.p, 0.0, W2.p);
    gemm(rocblas_operation_none, rocblas_operation_none, nb, nb, nb, 1.0, W2.p, X.p, 0.0, Es);
    HIPOK(hipMemcpyAsync(Fs, F.p, sizeof(double) * n2(), hipMemcpyDeviceToDevice, stream));
    nstored = std::min(nstored + 1, keep);
    for (int i = 0; i < nstored; ++i) {
      const double v = dot(Es, Eh.p + (size_t)i * n2());
      B[(size_t)slot * keep + i] = v;
      B[(size_t)i * keep + slot] = v;
    }
    if (iter == 1 || nstored < 2) return F.p;
    std::vector<double> Bn((size_t)nstored * nstored), c;
    for (int i = 0; i < nstored; ++i)
      for (int j = 0; j < nstored; ++j) Bn[(size_t)i * nstored + j] = B[(size_t)i * keep + j];
    if (!diis_solve(nstored, Bn, c)) return F.p;
    HIPOK(cqx::launch_zero(Fx.p, n2(), stream));
    for (int i = 0; i < nstored; ++i)
      BLASOK(rocblas_daxpy(blas, (rocblas_int)n2(), &c[i], Fh.p + (size_t)i * n2(), 1, Fx.p, 1));
    HIPOK(hipStreamSynchronize(stream));  // c lives on this stack frame
    return Fx.p;
  }

  // one iteration; returns true once converged
  bool iterate() {
    if (converged) return true;
    const double t0 = now_ms();
    fock_jk();
    ms_fock += now_ms() - t0;
    return after_jk(t0);
  }

  // the rest of an iteration once J and K hold J[D] and K[D] (t0: when the
  // iteration started); returns true once converged
  bool after_jk(double t0) {
    const double tf = now_ms();
    fock_from_jk();
    ms_fock += now_ms() - tf;
    e1 = 0.5 * dot(D.p, H.p);
    e2 = 0.25 * dot(D.p, G.p);
    E = e1 + e2 + e_nuc;
    ++iter;
    if (iter > 1) {
      dE = E - Eprev;
      HIPOK(cqx::launch_sub(D.p, Dp.p, W1.p, n2(), stream));
      double nrm = 0.0;
      BLASOK(rocblas_dnrm2(blas, (rocblas_int)n2(), W1.p, 1, &nrm));
      rmsd = nrm / nb;
      rocblas_int imax = 1;
      BLASOK(rocblas_idamax(blas, (rocblas_int)n2(), W1.p, 1, &imax));
      double v = 0.0;
      HIPOK(hipMemcpyAsync(&v, W1.p + (imax - 1), sizeof(double), hipMemcpyDeviceToHost, stream));
      HIPOK(hipStreamSynchronize(stream));
      maxd = std::fabs(v);
      // impl.hpp:286-321: energy not rising and changing less than eneConvTol,
      // density RMS and max changes below their tolerances
      const bool econv = dE < 1e-9 && std::fabs(dE) < o.ene_tol;
      const bool dconv = rmsd < o.rmsd_tol && maxd < o.maxd_tol;
      converged = econv && dconv;
    }
    Eprev = E;
    if (!converged) {
      HIPOK(hipMemcpyAsync(Dp.p, D.p, sizeof(double) * n2(), hipMemcpyDeviceToDevice, stream));
      diagonalise(diis());
      form_density();
    }
    ms_total += now_ms() - t0;
    return converged;
  }

  void alloc_all() {
    const size_t m = n2();
    for (auto *b : {&S, &H, &X, &D, &Dp, &J, &K, &G, &F, &Fx, &W1, &W2, &W3}) b->alloc(m);
    eps.alloc(nb);
    nonfinite.alloc(cqx::kNonfiniteBlocks);
    ework.alloc(nb);
    fscale.alloc(nb);
    info.alloc(1);
    const int keep = std::max(o.diis_keep, 1);
    Fh.alloc((size_t)keep * m);
    Eh.alloc((size_t)keep * m);
    B.assign((size_t)keep * keep, 0.0);
  }

  // S^-1/2 by the eigenvectors of S (orthogonalization/impl.hpp:53-80)
  void lowdin() {
    HIPOK(hipMemcpyAsync(W2.p, S.p, sizeof(double) * n2(), hipMemcpyDeviceToDevice, stream));
    BLASOK(rocsolver_dsyevd(blas, rocblas_evect_original, rocblas_fill_upper, nb, W2.p, nb, eps.p, ework.p,
                            info.p));
    std::vector<double> s(nb);
    rocblas_int inf = 0;
    HIPOK(hipMemcpyAsync(s.data(), eps.p, sizeof(double) * nb, hipMemcpyDeviceToHost, stream));
    HIPOK(hipMemcpyAsync(&inf, info.p, sizeof(inf), hipMemcpyDeviceToHost, stream));
    HIPOK(hipStreamSynchronize(stream));
    if (inf != 0) throw ScfError(CQX_ERR_HIP, "syevd of the overlap failed");
    if (std::fabs(s[0]) < 1e-10) throw ScfError(CQX_ERR_ARG, "Contracted Basis Set is Linearly Dependent!");
    for (double &v : s) v = 1.0 / std::sqrt(v);
    HIPOK(hipMemcpyAsync(fscale.p, s.data(), sizeof(double) * nb, hipMemcpyHostToDevice, stream));
    HIPOK(hipMemcpyAsync(W1.p, W2.p, sizeof(double) * n2(), hipMemcpyDeviceToDevice, stream));
    HIPOK(cqx::launch_scale_cols(W1.p, nb, nb, fscale.p, stream));
    gemm(rocblas_operation_none, rocblas_operation_transpose, nb, nb, nb, 1.0, W1.p, W2.p, 0.0, X.p);
    HIPOK(hipStreamSynchronize(stream));
  }
};

namespace {

thread_local std::string g_scf_error;

int scf_fail(cqx_scf *s, int code, const std::string &m) {
  if (s) s->err = m;
  g_scf_error = m;
  return code;
}

std::unique_ptr<cqx_scf> make_scf(cqx_ctx *ctx, int natom, const double *Z, const double *xyz,
                                  const cqx_scf_opts &o, bool frac, double nelec);

// Spin-averaged fractional-occupation SCF of one neutral atom of charge z
// carrying the shells of atom `atom` of the molecular basis; its density.
std::vector<double> atomic_density(const cqx::Basis &mol, int atom, double z, int device, double thresh) {
  std::vector<int> l, pure, nprim, at;
  std::vector<double> alpha, coef, center;
  for (const auto &sh : mol.shells) {
    if (sh.atom != atom) continue;
    l.push_back(sh.l);
    pure.push_back(sh.pure ? 1 : 0);
    nprim.push_back(sh.nprim());
    at.push_back(0);
    alpha.insert(alpha.end(), sh.alpha.begin(), sh.alpha.end());
    coef.insert(coef.end(), sh.coef.begin(), sh.coef.end());
    center.insert(center.end(), {0.0, 0.0, 0.0});
  }
  if (l.empty()) return {};
  cqx_basis *b = nullptr;
  CQXOK(nullptr, cqx_basis_create((int)l.size(), l.data(), pure.data(), nprim.data(), alpha.data(), coef.data(),
                                  center.data(), at.data(), &b));
  std::unique_ptr<cqx_basis, int (*)(cqx_basis *)> bg(b, cqx_basis_destroy);
  cqx_ctx *c = nullptr;
  CQXOK(nullptr, cqx_ctx_create(b, thresh, device, &c));
  std::unique_ptr<cqx_ctx, int (*)(cqx_ctx *)> cg(c, cqx_ctx_destroy);
  cqx_scf_opts o;
  cqx_scf_default_opts(&o);
  o.guess = CQX_GUESS_CORE;
  o.max_iter = 64;
  o.rmsd_tol = 1e-6;
  o.maxd_tol = 1e-4;
  o.ene_tol = 1e-6;
  const double Zc[1] = {z}, X0[3] = {0.0, 0.0, 0.0};
  auto s = make_scf(c, 1, Zc, X0, o, true, z);
  for (int k = 0; k < o.max_iter; ++k)
    if (s->iterate()) break;  // an unconverged atom still gives a usable guess
  std::vector<double> D(s->n2());
  HIPOK(hipMemcpy(D.data(), s->D.p, sizeof(double) * D.size(), hipMemcpyDeviceToHost));
  return D;
}

std::unique_ptr<cqx_scf> make_scf(cqx_ctx *ctx, int natom, const double *Z, const double *xyz,
                                  const cqx_scf_opts &o, bool frac, double nelec) {
  const double t0 = now_ms();
  auto s = std::make_unique<cqx_scf>();
  s->ctx = ctx;
  s->stream = cqx::ctx_stream(ctx);
  s->o = o;
  s->frac = frac;
  s->nelec = nelec;
  const cqx::Basis &bas = cqx::ctx_basis(ctx);
  s->nb = bas.nbasis;
  HIPOK(hipSetDevice(cqx::ctx_device(ctx)));
  BLASOK(rocblas_create_handle(&s->blas));
  BLASOK(rocblas_set_stream(s->blas, s->stream));
  s->alloc_all();
  // nuclear repulsion and one-electron integrals
  std::vector<double> Zc(Z, Z + natom), C(xyz, xyz + 3 * (size_t)natom);
  for (int i = 0; i < natom; ++i)
    for (int j = 0; j < i; ++j) {
      const double dx = C[3 * i] - C[3 * j], dy = C[3 * i + 1] - C[3 * j + 1], dz = C[3 * i + 2] - C[3 * j + 2];
      s->e_nuc += Zc[i] * Zc[j] / std::sqrt(dx * dx + dy * dy + dz * dz);
    }
  const auto Sm = cqx::one_body_matrix(bas, cqx::kOverlap, {}, {});
  const auto Tm = cqx::one_body_matrix(bas, cqx::kKinetic, {}, {});
  auto Hm = cqx::one_body_matrix(bas, cqx::kPotential, Zc, C);
  for (size_t i = 0; i < Hm.size(); ++i) Hm[i] = 2.0 * (Tm[i] + Hm[i]);
  HIPOK(hipMemcpyAsync(s->S.p, Sm.data(), sizeof(double) * Sm.size(), hipMemcpyHostToDevice, s->stream));
  HIPOK(hipMemcpyAsync(s->H.p, Hm.data(), sizeof(double) * Hm.size(), hipMemcpyHostToDevice, s->stream));
  HIPOK(hipStreamSynchronize(s->stream));
  s->lowdin();
  if (o.fp32_cut > 0.0) CQXOK(ctx, cqx_set_option(ctx, "fp32_cut", o.fp32_cut));
  // guess density
  bool done = false;
  if (o.guess == CQX_GUESS_SAD && !frac) {
    std::vector<double> Dg(s->n2(), 0.0);
    std::map<int, std::vector<double>> cache;  // element -> atomic density
    std::vector<int> first_bf(natom, -1);
    for (size_t k = 0; k < bas.shells.size(); ++k) {
      const int a = bas.shells[k].atom;
      if (a >= natom) throw ScfError(CQX_ERR_ARG, "basis atom label beyond the nuclei given");
      if (first_bf[a] < 0) first_bf[a] = bas.bf[k];
    }
    for (int a = 0; a < natom; ++a) {
      if (first_bf[a] < 0) continue;
      const int z = (int)std::lround(Zc[a]);
      auto it = cache.find(z);
      if (it == cache.end())
        it = cache.emplace(z, atomic_density(bas, a, Zc[a], cqx::ctx_device(ctx), cqx::ctx_thresh(ctx))).first;
      const auto &Da = it->second;
      const int na = (int)std::lround(std::sqrt((double)Da.size()));
      for (int j = 0; j < na; ++j)
        for (int i = 0; i < na; ++i) Dg[(first_bf[a] + i) + (size_t)(first_bf[a] + j) * s->nb] = Da[i + (size_t)j * na];
    }
    HIPOK(hipMemcpyAsync(s->D.p, Dg.data(), sizeof(double) * Dg.size(), hipMemcpyHostToDevice, s->stream));
    HIPOK(hipStreamSynchronize(s->stream));
    done = true;
  }
  if (!done) {  // core guess: orbitals of H
    s->diagonalise(s->H.p);
    s->form_density();
  }
  HIPOK(hipStreamSynchronize(s->stream));
  s->ms_setup = now_ms() - t0;
  return s;
}

}  // namespace

extern "C" {

void cqx_scf_default_opts(cqx_scf_opts *o) {
  if (!o) return;
  o->nocc = 0;
  o->max_iter = 128;
  o->rmsd_tol = 1e-8;
  o->maxd_tol = 1e-5;
  o->ene_tol = 1e-5;
  o->diis_keep = 8;
  o->guess = CQX_GUESS_SAD;
  o->fp32_cut = 0.0;
}

int cqx_scf_create(cqx_ctx *ctx, int natom, const double *Z, const double *xyz, const cqx_scf_opts *opts,
                   cqx_scf **out) {
  if (!ctx || natom <= 0 || !Z || !xyz || !opts || !out) return scf_fail(nullptr, CQX_ERR_ARG, "cqx_scf_create: bad arguments");
  if (opts->nocc <= 0 || opts->max_iter <= 0 || opts->diis_keep < 0 || opts->diis_keep > 64)
    return scf_fail(nullptr, CQX_ERR_ARG, "cqx_scf_create: bad options");
  if (opts->nocc > cqx::ctx_basis(ctx).nbasis)
    return scf_fail(nullptr, CQX_ERR_ARG, "cqx_scf_create: more occupied orbitals than basis functions");
  try {
    auto s = make_scf(ctx, natom, Z, xyz, *opts, false, 0.0);
    *out = s.release();
    return CQX_OK;
  } catch (const ScfError &e) {
    return scf_fail(nullptr, e.code, e.what());
  } catch (const std::exception &e) {
    return scf_fail(nullptr, CQX_ERR_HIP, e.what());
  }
}

int cqx_scf_iterate(cqx_scf *s, int *done) {
  if (!s) return scf_fail(nullptr, CQX_ERR_ARG, "cqx_scf_iterate: null handle");
  try {
    const bool d = s->iterate();
    if (done) *done = d ? 1 : 0;
    return CQX_OK;
  } catch (const ScfError &e) {
    return scf_fail(s, e.code, e.what());
  } catch (const std::exception &e) {
    return scf_fail(s, CQX_ERR_HIP, e.what());
  }
}

int cqx_scf_result_get(const cqx_scf *s, cqx_scf_result *r) {
  if (!s || !r) return scf_fail(nullptr, CQX_ERR_ARG, "cqx_scf_result_get: bad arguments");
  r->e_total = s->E;
  r->e_one = s->e1;
  r->e_two = s->e2;
  r->e_nuc = s->e_nuc;
  r->iterations = s->iter;
  r->converged = s->converged ? 1 : 0;
  r->rmsd = s->rmsd;
  r->maxd = s->maxd;
  r->delta_e = s->dE;
  r->ms_setup = s->ms_setup;
  r->ms_fock = s->ms_fock;
  r->ms_total = s->ms_setup + s->ms_total;
  return CQX_OK;
}

int cqx_scf_run(cqx_scf *s, cqx_scf_result *r) {
  if (!s) return scf_fail(nullptr, CQX_ERR_ARG, "cqx_scf_run: null handle");
  try {
    while (!s->converged && s->iter < s->o.max_iter) s->iterate();
  } catch (const ScfError &e) {
    return scf_fail(s, e.code, e.what());
  } catch (const std::exception &e) {
    return scf_fail(s, CQX_ERR_HIP, e.what());
  }
  if (r) cqx_scf_result_get(s, r);
  if (!s->converged)
    return scf_fail(s, CQX_ERR_NOT_CONVERGED,
                    "SCF Failed to converged within " + std::to_string(s->o.max_iter) + " iterations");
  return CQX_OK;
}

int cqx_scf_matrix(const cqx_scf *s, int which, double *out) {
  if (!s || !out) return scf_fail(nullptr, CQX_ERR_ARG, "cqx_scf_matrix: bad arguments");
  const double *src = which == CQX_SCF_DENSITY ? s->D.p
                      : which == CQX_SCF_FOCK  ? s->F.p
                      : which == CQX_SCF_HCORE ? s->H.p
                      : which == CQX_SCF_OVERLAP ? s->S.p
                                                 : nullptr;
  if (!src) return scf_fail(nullptr, CQX_ERR_ARG, "cqx_scf_matrix: unknown matrix");
  if (hipMemcpyAsync(out, src, sizeof(double) * s->n2(), hipMemcpyDeviceToHost, s->stream) != hipSuccess ||
      hipStreamSynchronize(s->stream) != hipSuccess)
    return scf_fail(nullptr, CQX_ERR_HIP, "cqx_scf_matrix: copy failed");
  return CQX_OK;
}

int cqx_scf_orbital_energies(const cqx_scf *s, double *eps) {
  if (!s || !eps) return scf_fail(nullptr, CQX_ERR_ARG, "cqx_scf_orbital_energies: bad arguments");
  std::copy(s->eps_host.begin(), s->eps_host.end(), eps);
  return CQX_OK;
}

int cqx_scf_destroy(cqx_scf *s) {
  delete s;
  return CQX_OK;
}

const char *cqx_scf_last_error(const cqx_scf *s) { return s ? s->err.c_str() : g_scf_error.c_str(); }

}  // extern "C"

namespace cqx {

int scf_fock_buffers(cqx_scf *s, const double **D, double **J, double **K, int *nb) {
  if (!s || !D || !J || !K || !nb) return scf_fail(s, CQX_ERR_ARG, "scf_fock_buffers: bad arguments");
  *D = s->D.p;
  *J = s->J.p;
  *K = s->K.p;
  *nb = s->nb;
  return CQX_OK;
}

int scf_after_jk(cqx_scf *s, double ms_jk, int *done) {
  if (!s) return scf_fail(nullptr, CQX_ERR_ARG, "scf_after_jk: null handle");
  try {
    s->ms_fock += ms_jk;
    const bool d = s->converged ? true : s->after_jk(now_ms() - ms_jk);
    if (done) *done = d ? 1 : 0;
    return CQX_OK;
  } catch (const ScfError &e) {
    return scf_fail(s, e.code, e.what());
  } catch (const std::exception &e) {
    return scf_fail(s, CQX_ERR_HIP, e.what());
  }
}

int scf_iterations(const cqx_scf *s, int *iter, int *max_iter) {
  if (!s) return scf_fail(nullptr, CQX_ERR_ARG, "scf_iterations: null handle");
  if (iter) *iter = s->iter;
  if (max_iter) *max_iter = s->o.max_iter;
  return CQX_OK;
}

}  // namespace cqx
