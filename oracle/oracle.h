/*
 * FP64 CPU oracle for the direct J/K (Fock) build.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * product path: only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline leg may load it.  The product library never links or calls it.
 *
 * It restates, in plain C, the algorithm of ChronusQ's
 *   GTODirectTPIContraction::directScaffoldNew
 *     (include/particleintegrals/contract/direct.hpp:1412-2123)
 * together with the pieces that feed it:
 *   G94 basis parsing            src/basisset/reference.cxx:127-300
 *   shell order / bf maps        src/basisset/basisset.cxx:98-196
 *   significant shell pairs      src/basisset/basisset.cxx:652-705
 *   Schwarz bounds               src/particleintegrals/aointegrals/aointegrals_utilities.cxx:53-162
 *   shell-block inf-norms        include/particleintegrals/contract/direct.hpp:73-105
 * The ERI arithmetic itself lives in the third-party libint2
 * (xsligroup/libint-cq @ 2.7.0-beta.6), which is not present; the oracle
 * evaluates ERIs with its own Obara-Saika vertical recursion plus
 * Head-Gordon-Pople horizontal recursion, and its own Boys function.
 *
 * Conventions (those of libint2 as used by ChronusQ):
 *   - contracted shells normalised so that the x^l Cartesian component has
 *     unit norm; pure (spherical) functions for l > 1 unless forced Cartesian;
 *   - Cartesian order xx,xy,xz,yy,yz,zz (lx descending, then ly descending);
 *   - real solid harmonics ordered m = -l..l;
 *   - matrices column-major, M(i,j) = M[i + j*NB];
 *   - shell-quartet buffers row-major ijkl.
 *
 * Parity unpinned: libint2's primitive-level screening (precision
 * max(eps, thresh/(maxShBlkNorm*maxPrim^4))) is not restated; the oracle
 * keeps every primitive.  Its effect is below ~1e-14 per quartet.
 */
#ifndef CQX_ORACLE_H
#define CQX_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_basis oracle_basis;

/* Build a basis for a molecule.  gbs_text: G94 basis text.  Z[natom],
 * xyz[3*natom] in bohr.  force_cart: 1 forces Cartesian d/f/...
 * Returns NULL on error (message via oracle_last_error()). */
oracle_basis *oracle_basis_create(const char *gbs_text, int natom,
                                  const int *Z, const double *xyz,
                                  int force_cart);
void oracle_basis_destroy(oracle_basis *b);
const char *oracle_last_error(void);

int oracle_nbasis(const oracle_basis *b);
int oracle_nshell(const oracle_basis *b);
int oracle_maxprim(const oracle_basis *b);
int oracle_maxl(const oracle_basis *b);
/* per-shell info: l, pure, nprim, first bf, atom index */
void oracle_shell_info(const oracle_basis *b, int s, int *l, int *pure,
                       int *nprim, int *bf, int *atom);
/* normalised primitive data of shell s (nprim entries each) */
void oracle_shell_prims(const oracle_basis *b, int s, double *alpha,
                        double *coef, double *origin);

/* Boys function F_m(T), m = 0..mmax */
void oracle_boys(int mmax, double T, double *F);

/* overlap matrix (NB x NB, column-major) */
void oracle_overlap(const oracle_basis *b, double *S);

/* contracted shell-quartet ERIs (s1 s2|s3 s4), row-major ijkl, returns count */
int oracle_eri_quartet(const oracle_basis *b, int s1, int s2, int s3, int s4,
                       double *out);

/* full NB^4 ERI tensor, (ij|kl) at [((i*NB+j)*NB+k)*NB+l]; tiny bases only */
void oracle_eri_full(const oracle_basis *b, double *out);

/* significant pair list (s2 <= s1) as flat arrays: returns #pairs;
 * if s1/s2 are non-NULL they receive the pairs in reference loop order */
int oracle_shell_pairs(const oracle_basis *b, int *s1, int *s2);

/* Schwarz matrix Q (NS x NS, column-major, symmetric) */
void oracle_schwarz(const oracle_basis *b, double *Q);

/* Contraction descriptors: one per matrix.  X/AX point at NB*NB doubles
 * (is_complex=0) or NB*NB interleaved complex doubles (is_complex=1).
 * type: 0 = COULOMB, 1 = EXCHANGE.  her: Hermitian flag.  AX must be
 * zeroed by the caller (as ChronusQ's callers do). */
typedef struct {
  const void *X;
  void *AX;
  int her;
  int type;
} oracle_contraction;

/* GTODirectTPIContraction::twoBodyContract semantics.  nthreads <= 0 uses
 * omp_get_max_threads().  Returns 0 on success. */
int oracle_two_body_contract(oracle_basis *b, double thresh_schwarz,
                             int screen, int nmat,
                             const oracle_contraction *c, int is_complex,
                             int nthreads);

/* DirectGradContraction::directScaffoldGrad semantics (real, Hermitian
 * only, no screening): c[g * nmat + m] for g = 3 * atom + xyz holds the
 * density X and the output AX of matrix m for nuclear coordinate g; every
 * AX receives 1/2 (A + A^T) of the derivative-integral digestion (callers
 * zero it). */
int oracle_grad_contract(oracle_basis *b, int nmat, const oracle_contraction *c, int nthreads);

/* compute and cache the Schwarz bounds (otherwise done by the first
 * contract call, as DirectTPI::computeSchwarz is in the reference) */
void oracle_prepare(oracle_basis *b);

/* quartet accounting of the last contract call on this basis */
void oracle_last_stats(const oracle_basis *b, long long *n_computed,
                       long long *n_skipped);

#ifdef __cplusplus
}
#endif
#endif
