/*
 * cqx -- MI355X-native direct Fock-build (J/K) engine, C ABI.
 *
 * This is the drop-in boundary for ChronusQ's direct two-body contraction:
 *   GTODirectTPIContraction<MatsT,IntsT>::twoBodyContract
 *     (include/particleintegrals/twopints/gtodirecttpi.hpp:199-206)
 *   -> directScaffoldNew
 *     (include/particleintegrals/contract/direct.hpp:1412-2123)
 * and the DirectTPI state it reads (gtodirecttpi.hpp:30-141).  A C++
 * adapter deriving from TPIContractions<MatsT,IntsT> (twopints.hpp:172-260)
 * calls these entry points; see include/cqx/cqx.hpp for the mirror of the
 * ChronusQ-side classes.
 *
 * All entry points return 0 on success and a non-zero cqx_status on error;
 * the message is available from cqx_last_error().  ChronusQ's CErr throws
 * std::runtime_error("FATAL") (include/cerr.hpp:33-51); the C++ adapter
 * converts a non-zero status into that exception.
 *
 * Matrices are column-major, M(i,j) = M[i + j*NB], as cqmatrix::Matrix
 * stores them (include/matrix/matrix.hpp:189).  Complex matrices are
 * interleaved (re, im) pairs, the layout of std::complex<double>.
 */
#ifndef CQX_H
#define CQX_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum cqx_status {
  CQX_OK = 0,
  CQX_ERR_ARG = 1,     /* invalid argument */
  CQX_ERR_HIP = 2,     /* HIP runtime error */
  CQX_ERR_MEM = 3,     /* device or host allocation failed */
  CQX_ERR_BASIS = 4,   /* basis-set parse or lookup failure */
  CQX_ERR_UNSUPPORTED = 5,
  CQX_ERR_NOT_CONVERGED = 6 /* SCF did not converge within max_iter */
};

enum cqx_contraction_type { /* TWOBODY_CONTRACTION_TYPE, twopints.hpp:31-41 */
  CQX_COULOMB = 0,
  CQX_EXCHANGE = 1
};

enum cqx_memory { CQX_HOST = 0, CQX_DEVICE = 1 };

typedef struct cqx_basis cqx_basis; /* host-side shell set */
typedef struct cqx_ctx cqx_ctx;     /* device-resident DirectTPI state */

/* ---------------------------------------------------------------------
 * Shell sets.  A basis is a list of contracted shells, each with its
 * angular momentum, pure/Cartesian flag, origin (bohr) and primitives whose
 * coefficients are already normalised in the libint2 convention that
 * ChronusQ's BasisSet holds (libint2::Shell after renorm()).  Replaces the
 * BasisSet the reference builds in src/basisset/basisset.cxx:98-196.
 * ------------------------------------------------------------------- */

/* Build from per-shell arrays.  alpha/coef are concatenated over shells
 * (nprim[s] entries each); center is 3*nshell doubles; atom[s] labels the
 * nuclear centre (shells with equal labels count as "same centre"). */
int cqx_basis_create(int nshell, const int *l, const int *pure,
                     const int *nprim, const double *alpha,
                     const double *coef, const double *center,
                     const int *atom, cqx_basis **out);

/* Build from G94 text for a molecule, with the parsing and normalisation
 * rules of ReferenceBasisSet (src/basisset/reference.cxx:127-357): SP
 * shells split into S and P, pure functions iff l > 1 unless force_cart,
 * shells ordered by atom then file order.  Z[natom], xyz[3*natom] bohr. */
int cqx_basis_from_g94(const char *g94_text, int natom, const int *Z,
                       const double *xyz, int force_cart, cqx_basis **out);

int cqx_basis_destroy(cqx_basis *b);
int cqx_basis_dims(const cqx_basis *b, int *nshell, int *nbasis,
                   int *max_prim, int *max_l);
/* per-shell l, pure, nprim, first basis function, atom label */
int cqx_basis_shell(const cqx_basis *b, int s, int *l, int *pure,
                    int *nprim, int *bf, int *atom);
/* normalised exponents/coefficients of shell s and its origin */
int cqx_basis_shell_prims(const cqx_basis *b, int s, double *alpha,
                          double *coef, double *origin);

/* Several molecules' shell sets side by side, for a batched context: a
 * context created on the result evaluates every molecule's quartets in the
 * same launches (one per class pair and precision), never pairing shells of
 * different molecules.  cqx_two_body_contract on it takes nmol x n
 * contractions, molecule-major (molecule m's n at [m n, (m + 1) n), NB_m x
 * NB_m each, the same types and HER flags for every molecule).  No
 * counterpart in the reference, which contracts one molecule per call. */
int cqx_basis_batch(int nmol, const cqx_basis *const *parts, cqx_basis **out);

/* One-electron integrals of a basis into out (column-major NB x NB, pure
 * functions where the shell is pure); host computation, no device needed.
 * Stands in for OnePInts::computeAOInts with libint2's overlap, kinetic
 * and point-nucleus operators (src/particleintegrals/aointegrals/
 * aointegrals_onee_drivers.cxx:762-784).  For CQX_POTENTIAL the operator
 * is sum_C -Z[C] / |r - C| over ncharge point charges Z at xyz (bohr). */
enum { CQX_OVERLAP = 0, CQX_KINETIC = 1, CQX_POTENTIAL = 2 };
int cqx_one_body(const cqx_basis *b, int kind, int ncharge, const double *Z,
                 const double *xyz, double *out);

/* ---------------------------------------------------------------------
 * Context: DirectTPI<double>(basis, basis, mol, threshSchwarz) plus the
 * device-resident shell-pair data.  Creating it computes the significant
 * shell-pair list (ShellPairData::computeShellPairs, basisset.cxx:652-705)
 * and the Schwarz bounds on the GPU (DirectTPI::computeSchwarz,
 * aointegrals_utilities.cxx:53-162).  device < 0 uses the current device.
 * ------------------------------------------------------------------- */
int cqx_ctx_create(const cqx_basis *b, double thresh_schwarz, int device,
                   cqx_ctx **out);
int cqx_ctx_destroy(cqx_ctx *ctx);
const char *cqx_last_error(const cqx_ctx *ctx); /* NULL ctx: global error */

/* Schwarz matrix (nshell x nshell, column-major, symmetric), copied to a
 * host buffer. */
int cqx_ctx_schwarz(cqx_ctx *ctx, double *Q);
/* number of significant shell pairs and the pairs themselves (s2 <= s1,
 * in the reference loop order); s1/s2 may be NULL to query the count. */
int cqx_ctx_shell_pairs(cqx_ctx *ctx, int *npairs, int *s1, int *s2);

/* One contraction: TwoBodyContraction<MatsT> (twopints.hpp:59-75) without
 * the ERI4/intTrans fields that only the in-core path reads. */
typedef struct {
  const void *X; /* NB x NB, double or interleaved complex */
  void *AX;      /* NB x NB, same type; accumulated into (callers zero it) */
  int her;       /* X is Hermitian */
  int type;      /* cqx_contraction_type */
} cqx_contraction;

/* GTODirectTPIContraction::twoBodyContract(comm, screen, list, pert):
 * AX += J[X] or K[X] for every entry, with the reference's Schwarz x
 * shell-block screening, 8-fold symmetry, digestion and Hermitian
 * symmetrisation.  is_complex selects double vs complex<double>; mem says
 * whether X/AX are host or device pointers.  Any number of contractions;
 * they are processed in groups of at most 8 that share the ERIs. */
int cqx_two_body_contract(cqx_ctx *ctx, int screen, int ncontract,
                          const cqx_contraction *list, int is_complex,
                          int mem);

/* GTODirectTPIContraction::directScaffoldNewSCRSize (direct.hpp:2125-2180),
 * which FockBuilder::formRawGDSCRSizePerBatch (fockbuilder/impl.hpp:252-269)
 * multiplies by the contractions of a batch: *bytes receives an upper bound
 * of the device memory one cqx_two_body_contract call with this list
 * allocates on the context (densities and outputs staged on the device,
 * shell-block norms, Cartesian planes and accumulators, candidate offsets,
 * quartet lists and their tables), under the context's current options
 * ("list_cap", "fp32_cut", "graph", "autotune").  Only the types and HER
 * flags of the list are read; X / AX may be NULL.  The workspace stays
 * with the context for later calls (cqx_stats.workspace_bytes). */
int cqx_scratch_size(cqx_ctx *ctx, int ncontract, const cqx_contraction *list,
                     int is_complex, size_t *bytes);

/* Allocation-free cqx_two_body_contract.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory during a call (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same contraction, screening and results as
 * cqx_two_body_contract with mem = CQX_DEVICE: X and AX are device pointers
 * and AX is accumulated in place (callers zero it).  Every per-call buffer
 * comes from one fixed device region of CQX_NOALLOC_SCRATCH_BYTES per GPU,
 * held by the library from the first context created on that GPU until the
 * process ends; nothing is allocated, recorded or tuned.  A list whose
 * fixed workspace (staged densities, shell-block norms, Cartesian planes
 * and accumulators, candidate offsets) does not fit the region is rejected
 * with CQX_ERR_UNSUPPORTED before any launch; the screened quartet lists
 * are built in as many batches as the rest of the region requires.  All
 * calls on one GPU share the region: calls that overlap in time on the same
 * GPU (from any context or thread) race and must be serialised by the
 * caller.  The call returns after its work on the context's stream is
 * complete. */
#define CQX_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
int cqx_two_body_contract_noalloc(cqx_ctx *ctx, int screen, int ncontract,
                                  const cqx_contraction *list, int is_complex);

/* Contracted shell-quartet ERIs for nq quartets (s1,s2,s3,s4 arrays), in
 * the libint2 layout the reference digests (row-major ijkl, pure
 * functions where the shell is pure).  out receives the quartets back to
 * back; offsets[q] (nq+1 entries, host) gives where each starts. */
int cqx_eri_quartets(cqx_ctx *ctx, int nq, const int *s1, const int *s2,
                     const int *s3, const int *s4, double *out,
                     size_t *offsets);

/* Statistics of the last cqx_two_body_contract call. */
typedef struct {
  long long n_candidates; /* quartets tested after the Schwarz pre-cut */
  long long n_quartets;   /* quartets evaluated */
  long long n_prim_quartets_max; /* sum over evaluated quartets of Ka Kb Kc Kd */
  double flops_model;     /* algorithmic flop model summed over evaluated quartets */
  double ms_total;        /* wall time of the call */
  double ms_eri;          /* device time of the ERI/digestion kernels */
  int n_launch_eri;       /* ERI kernel launches */
  long long n_quartets_f32; /* of n_quartets, those evaluated in FP32 */
  double flops_model_f32;   /* of flops_model, the part on FP32 quartets */
  int graph;                /* 0 launched eagerly, 1 eagerly and recorded into a graph,
                               2 a replay of the recorded graph (ms_eri then covers
                               the whole recorded build) */
  long long workspace_bytes; /* device workspace the context holds after the call
                                (cqx_two_body_contract_noalloc: scratch bytes it used) */
} cqx_stats;
int cqx_last_stats(const cqx_ctx *ctx, cqx_stats *st);

/* DirectGradContraction::gradTwoBodyContract (direct.hpp:2198-2580;
 * gradints/direct.hpp:31-76): derivative-integral J / K contractions for
 * every nuclear coordinate.  list[g * nmat + m] (g = 3 atom + xyz) holds
 * the density X of matrix m -- the same X and type for every g, as
 * FockBuilder::getGDGrad builds the list (fockbuilder/impl.hpp:371-446) --
 * and the NB x NB output AX of coordinate g, which receives
 * 1/2 (A + A^T) (callers zero it).  Real, Hermitian contractions only
 * (the reference instantiates <double, double> and stops on non-Hermitian
 * ones), no screening (as the reference, direct.hpp:2204-2205), s through g
 * shells, one-molecule contexts; mem as for cqx_two_body_contract. */
int cqx_grad_two_body_contract(cqx_ctx *ctx, int ngrad, int nmat, const cqx_contraction *list, int mem);

/* Tuning knobs: "prim_screen" (primitive-pair drop threshold, default
 * 1e-18; rebuilds the pair data), "fast_max_nc" (largest Cartesian quartet
 * size routed to the lane-group kernels without a tuned choice, default
 * 27), "use_wave" (0 routes
 * the larger s/p/d classes and the f and g classes of at most 1080
 * Cartesian components to the generic kernel instead of the templated
 * wave-per-quartet ones), "fast_lanes" (lanes per shell quartet of the
 * lane-group kernels: 1, 2, 4, 8 or 16; 0, the default, picks it per launch
 * from the mean primitive count), "autotune" (1: the next contraction
 * times every lane count of every lane-group launch and every number of
 * quartets a wave kernel evaluates before digesting them (1, 2, 4, 8) on
 * scratch accumulators (1, 2 or, up to 216 Cartesian components, 4 for
 * the f and g classes), each replicated on the build's streams at once;
 * for (ds|ds), (dd|ss), (dp|ps) and (ds|pp), built in both families, the
 * lane-group and wave shapes compete; keeps the fastest per class pair and
 * precision and uses it from then on; 0 forgets the choices), "fp32_cut"
 * (mixed precision: quartets whose
 * Schwarz product Q_ab Q_cd is below this value are evaluated in FP32 with
 * an FP32 Rys table, digested and accumulated in FP64; default 0 = every
 * quartet in FP64; classes with an f or g shell always stay FP64), "list_cap"
 * (screened quartet-list entries built per batch, default 2^28: the list
 * workspace is 8 B per entry, twice that with fp32_cut, whatever the
 * molecule size), "graph" (1: a build records its launches, from the
 * shell-block norms to the accumulation into AX, into a HIP graph, and the
 * next builds with the same contraction signature replay it with one host
 * synchronisation; the list and ERI kernels read their counts from the
 * device; a replay that finds more candidates than the recording build's
 * capacities is discarded and built again eagerly; one contraction group
 * (at most 8 contractions per molecule) and one list batch only, otherwise
 * every build runs eagerly), "graph_margin" (capacity of each class pair's
 * candidates and lists in a recorded graph, as a multiple of the recording
 * build's candidates; default 1.25), "streams" (1-4: a build, eager or
 * recorded, deals its ERI launches round-robin over this many streams,
 * forked from and joined back into the context's stream, so that launches
 * of different class pairs overlap; default 4; results differ from one
 * stream only in the order of the FP64 atomic additions), "tiles" (1: the
 * lane-group class pairs of a build whose contractions are all Hermitian
 * evaluate their quartets in tiles of up to 8 bra pairs sharing their first
 * shell against up to 8 ket pairs sharing theirs, and sum the shared J / K
 * targets of a tile in LDS before adding them to HBM once; 2: the same,
 * only for builds with at most one Hermitian J and one Hermitian K plane
 * per molecule (real RHF J / K); 0, the default: never; not used by
 * recorded or allocation-free builds), "tile_classes" (bit k: the k-th lane-group class
 * of (ss|ss), (ps|ss), (ps|ps), (pp|ss), (pp|ps), (ds|ss), (ds|ps), (dp|ss)
 * takes tiles; default 126, all but (ss|ss) and (dp|ss)), "rys_lds" (1, the
 * default: the lane-group kernels read the Rys table section of their root
 * count from a copy in LDS; 0: from global memory).  Setting any option
 * drops a recorded graph. */
int cqx_set_option(cqx_ctx *ctx, const char *name, double value);

/* ---------------------------------------------------------------------
 * Pauli-spinor components.  A two-component matrix A (2NB x 2NB, spin
 * blocks AA, AB / BA, BB) and its components S = AA + BB, Z = AA - BB,
 * Y = f (AB - BA), X = AB + BA, with f = i for complex matrices and -1 for
 * real ones (ComplexScale, src/cqlinalg/blasutil.cxx:67-69).  Each call runs
 * on the context's GPU; mem says whether every matrix argument is a host or
 * a device pointer.  Column-major throughout; lda >= 2 NB.
 * ------------------------------------------------------------------- */
enum {
  CQX_SPIN_ZERO_ABBA = 1, /* scatter: Y = X = 0 (zeroABBA) */
  CQX_SPIN_BB_EQ_AA = 2,  /* scatter: S = 2 AA, Z = 0 (BBeqAA) */
  CQX_SPIN_ZERO_XY = 4,   /* gather: AB = BA = 0 (zeroXY) */
  CQX_SPIN_ZERO_Z = 8     /* gather: AA = BB = S / 2 (zeroZ) */
};
/* SpinScatter(N, A, LDA, AS, AZ, AY, AX, zeroABBA, BBeqAA)
 * (blasutil.cxx:71-129): NB x NB components out; any of S, Z, Y, X may be
 * NULL (not written). */
int cqx_spin_scatter(cqx_ctx *ctx, int nb, const void *A, int lda, void *S, void *Z,
                     void *Y, void *X, int is_complex, int flags, int mem);
/* SpinGather(N, A, LDA, AS, AZ, AY, AX, zeroXY, zeroZ) (blasutil.cxx:131-189):
 * AA = (S + Z)/2, BB = (S - Z)/2, BA = (X + f Y)/2, AB = (X - f Y)/2.  Z may be
 * NULL with CQX_SPIN_ZERO_Z, Y and X with CQX_SPIN_ZERO_XY. */
int cqx_spin_gather(cqx_ctx *ctx, int nb, void *A, int lda, const void *S, const void *Z,
                    const void *Y, const void *X, int is_complex, int flags, int mem);
/* SingleSlater::computeEnergy (include/singleslater/quantum.hpp:147-212):
 * out[0] = E1 = 1/2 sum_c Re<D_c, H_c>, out[1] = E2 = 1/4 sum_c Re<D_c, G_c>,
 * out[2] = E1 + E2 + e_nuc, over ncomp (1: S; 2: S, Z; 4: S, Z, Y, X)
 * NB x NB components, Re<A, B> = Re sum conj(A) B (OperatorTrace,
 * quantum/base.hpp:46-51).  H or G (or single components of them) may be
 * NULL: that term is then zero.  A non-finite energy is CQX_ERR_ARG, as the
 * reference asserts. */
int cqx_compute_energy(cqx_ctx *ctx, int nb, int ncomp, const void *const *D,
                       const void *const *H, const void *const *G, int is_complex,
                       int mem, double e_nuc, double *out);
/* PauliSpinorMatrices::hasNaN (include/matrix/paulispinormatrices.hpp:383-388),
 * extended to infinities: *count receives how many of the n doubles at M
 * are not finite (a complex matrix is 2 NB^2 doubles). */
int cqx_count_nonfinite(cqx_ctx *ctx, const double *M, size_t n, int mem, size_t *count);

/* ---------------------------------------------------------------------
 * Device memory for callers that keep densities and results resident in
 * HBM (mem = CQX_DEVICE).  No counterpart in the reference, whose matrices
 * live in host memory (CQMemManager); these let a caller stage X / AX on
 * the GPU of a context without a second HIP runtime in the process.
 * ------------------------------------------------------------------- */
int cqx_device_alloc(cqx_ctx *ctx, size_t bytes, void **ptr);
int cqx_device_free(cqx_ctx *ctx, void *ptr);
/* kind: 0 host->device, 1 device->host, 2 device->device; synchronous */
int cqx_memcpy(cqx_ctx *ctx, void *dst, const void *src, size_t bytes, int kind);
/* zero device memory (fill kernel on the context stream) and wait */
int cqx_device_zero(cqx_ctx *ctx, void *ptr, size_t bytes);

/* ---------------------------------------------------------------------
 * Device-resident RHF SCF on a context: the SingleSlater / conventional
 * SCF loop of the reference for a closed-shell real reference
 * (include/orbitalmodifiernew/orbitaloptimizernew/impl.hpp:40-137 and
 * :266-330; Lowdin orthogonalisation include/orthogonalization/impl.hpp:
 * 40-80; CDIIS include/orbitalmodifiernew/conventionalSCFnew/extrap.hpp:
 * 36-123; formDensity / computeEnergy include/singleslater/quantum.hpp:
 * 44-212).  Pauli-scalar conventions: D = D_S = 2 C_occ C_occ^T,
 * H = 2 (T + V), G = 2 J[D] - K[D], F = H + G, E = 1/2 Tr(D H) +
 * 1/4 Tr(D G) + V_NN.  Fock builds run through cqx_two_body_contract on the
 * context's GPU; orthogonalisation, DIIS, eigensolve (rocSOLVER syevd) and
 * density formation (rocBLAS GEMM on the matrix cores) stay on the device.
 * ------------------------------------------------------------------- */
enum cqx_guess { CQX_GUESS_CORE = 0, CQX_GUESS_SAD = 1 };

typedef struct {
  int nocc;          /* doubly occupied orbitals */
  int max_iter;      /* default 128 */
  double rmsd_tol;   /* ||D - D_prev||_2 / NB, default 1e-8 */
  double maxd_tol;   /* max |D - D_prev|, default 1e-5 (SCFControls) */
  double ene_tol;    /* |E - E_prev|, default 1e-5 (SCFControls) */
  int diis_keep;     /* CDIIS subspace size, default 8 */
  int guess;         /* cqx_guess, default CQX_GUESS_SAD */
  double fp32_cut;   /* mixed-precision Fock builds (see "fp32_cut"); 0 = FP64 */
} cqx_scf_opts;
/* fill the defaults above (nocc = 0: the caller sets it) */
void cqx_scf_default_opts(cqx_scf_opts *o);

typedef struct {
  double e_total, e_one, e_two, e_nuc;
  int iterations;    /* Fock builds done */
  int converged;
  double rmsd, maxd, delta_e;
  double ms_setup;   /* one-electron integrals, orthogonaliser, guess */
  double ms_fock;    /* Fock builds */
  double ms_total;
} cqx_scf_result;

typedef struct cqx_scf cqx_scf;
/* natom nuclei Z (charges) at xyz (bohr); the context's basis must sit on
 * them.  Computes S, T, V, the Lowdin orthogonaliser and the guess. */
int cqx_scf_create(cqx_ctx *ctx, int natom, const double *Z, const double *xyz,
                   const cqx_scf_opts *opts, cqx_scf **out);
/* One SCF iteration (Fock build, energy, convergence test, CDIIS,
 * eigensolve, new density); *done = 1 once converged. */
int cqx_scf_iterate(cqx_scf *s, int *done);
/* Iterate to convergence; CQX_ERR_NOT_CONVERGED after max_iter. */
int cqx_scf_run(cqx_scf *s, cqx_scf_result *res);
int cqx_scf_result_get(const cqx_scf *s, cqx_scf_result *res);
/* host copies (column-major NB x NB) of the density D_S, the Fock matrix
 * F_S built from it, the core Hamiltonian H_S and the overlap S */
int cqx_scf_matrix(const cqx_scf *s, int which, double *out);
enum { CQX_SCF_DENSITY = 0, CQX_SCF_FOCK = 1, CQX_SCF_HCORE = 2, CQX_SCF_OVERLAP = 3 };
/* orbital energies (eigenvalues of F_S / 2) of the last diagonalisation */
int cqx_scf_orbital_energies(const cqx_scf *s, double *eps);
int cqx_scf_destroy(cqx_scf *s);
const char *cqx_scf_last_error(const cqx_scf *s);

/* ---------------------------------------------------------------------
 * Serving loop (no counterpart in the reference, which runs one SCF per
 * process): SCF jobs arrive over time and run concurrently on one GPU.
 * The oldest waiting job is admitted as soon as fewer than max_inflight
 * jobs run and its footprint estimate fits hbm_budget_bytes beside the
 * running ones; a job whose footprint alone exceeds the budget is refused
 * (status CQX_ERR_MEM).  Each admitted job sets up its own context and SCF
 * and keeps its state on the device until it converges.  Its iterations
 * run either on its own stream (batch_fock = 0: the jobs' Fock builds
 * overlap through their streams) or in rounds (batch_fock = 1): every
 * round builds the J and K of the jobs in flight of one size class (basis
 * sizes within a factor of two) in one batched contraction over a context
 * merged from their pair data, then each job finishes its iteration on its
 * own stream; each size class runs its rounds on a thread of its own; jobs
 * join at the next round after set-up and leave once converged.  A Fock matrix with a
 * non-finite entry stops its job (CQX_ERR_ARG).  Blocks until every job
 * has finished; returns the first non-zero job status, if any.
 * ------------------------------------------------------------------- */
typedef struct {
  const char *g94;     /* basis set (G94 text) */
  int natom;
  const int *Z;        /* atomic numbers (neutral nuclei) */
  const double *xyz;   /* 3 natom, bohr */
  int nocc;            /* doubly occupied orbitals */
  double arrival_s;    /* arrival time, seconds after the call starts */
  double *density_out; /* NULL, or NB x NB host buffer: the final density D_S */
  double *fock_out;    /* NULL, or NB x NB host buffer: the Fock matrix F_S of D_S */
} cqx_job;

typedef struct {
  int max_inflight;        /* jobs on the device at once, default 8 */
  double hbm_budget_bytes; /* admission budget, default 200e9 */
  int device;
  double thresh;           /* Schwarz threshold, default 1e-12 */
  long long list_cap;      /* per-job quartet-list batch (see "list_cap"), default 2^25 */
  cqx_scf_opts scf;        /* per-job SCF options (nocc taken from the job) */
  int verbose;             /* 1: one line per finished job on stderr */
  int graph;               /* 1: each job's context (batch_fock: each merged
                              context) records its Fock build into a HIP graph
                              (option "graph") and later iterations replay it;
                              default 0 */
  int batch_fock;          /* 1: cross-request batching of the Fock builds in
                              rounds (see above); default 0 */
} cqx_serve_opts;
void cqx_serve_default_opts(cqx_serve_opts *o);

typedef struct {
  int status;              /* cqx_status of the job */
  int converged, iterations;
  double e_total;
  double t_arrival, t_admit, t_done; /* seconds after the call started */
  double footprint_bytes;  /* admission estimate */
  double ms_fock;          /* the job's Fock-build time */
} cqx_job_result;
int cqx_serve(const cqx_job *jobs, int njobs, const cqx_serve_opts *opts,
              cqx_job_result *out);

#ifdef __cplusplus
}
#endif
#endif /* CQX_H */
