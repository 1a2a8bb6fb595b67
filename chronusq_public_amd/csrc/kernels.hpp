// Host-callable launchers of the HIP kernels (implemented in *.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

#include "device_types.hpp"

namespace cqx {

struct EriLaunch {
  int la, lb, lc, ld;
  const Quartet *quartets;
  long long nq;
  // a launch sized before its list is built (a recorded build): nq is the
  // capacity of the launch's part of the list, which the grid covers; the
  // list's device counter holds the quartet count of the whole list, whose
  // entries from nq_base on the launch takes
  const unsigned long long *nq_dev;
  long long nq_base;
  const PairRec *bra, *ket;
  const PrimPair *prims;
  const PrimC32 *primc32;        // the same pairs for the FP32 lists of the templated kernels
  const double *rys;
  const float *rys32;            // the same table rounded to FP32
  int fp32;                      // evaluate the quartets in FP32 (digestion stays FP64)
  double mean_prims;             // mean primitive quartets per listed shell quartet
  int force_g;                   // lanes per quartet of the lane-group kernels (0: chosen per launch)
  int force_qt;                  // quartets per wave of the wave kernels (0: the class default)
  int rys_lds;                   // lane-group kernels: stage the class's Rys table section in LDS
  int team_ca, team_rpp;         // generic kernel: first-shell components per pass, roots per table pass (set by its launcher)
  PlaneSet planes;
  double *store;                 // STORE mode: contracted Cartesian quartets
  const long long *store_off;    // per quartet offset into store
  int store_mode;
};

// Generic one-wave-per-quartet kernel (any class up to (gg|gg)).
hipError_t launch_eri_team(const EriLaunch &a, hipStream_t s);
// Passes of the generic kernel within lds_max_bytes of LDS: first-shell
// components per pass (ca) and roots per 2D-table pass (rpp); false when
// not even one of each fits
bool team_shape(int la, int lb, int lc, int ld, int lds_max_bytes, int &ca, int &rpp);
// Fast thread-per-quartet kernels for the low classes; returns
// hipErrorNotSupported when the class has no specialised kernel.
hipError_t launch_eri_fast(const EriLaunch &a, hipStream_t s);
bool eri_fast_available(int la, int lb, int lc, int ld);
// (ds|ds), (dd|ss), (dp|ps), (ds|pp): lane-group kernels beside their wave
// kernels (the autotuner picks the family); eri_lane_group_built is true for
// these and the eight classes of eri_fast_available (which alone take tiles)
bool eri_lane_group_dual(int la, int lb, int lc, int ld);
bool eri_lane_group_built(int la, int lb, int lc, int ld);
// position of a lane-group class in (ss|ss), (ps|ss), (ps|ps), (pp|ss),
// (pp|ps), (ds|ss), (ds|ps), (dp|ss); -1 for the other classes
int fast_class_index(int la, int lb, int lc, int ld);
// One-wave-per-quartet templated kernels for the larger s/p/d classes and
// the f and g classes of at most 1080 Cartesian components (FP64).
hipError_t launch_eri_wave(const EriLaunch &a, hipStream_t s);
bool eri_wave_available(int la, int lb, int lc, int ld);
// whether the wave kernel of a class is built with qt (1, 2, 4, 8) quartets
// per wave for the precision (its LDS per block within the limit; the f
// and g classes: FP64 with 1 or 2, and 4 up to 216 Cartesian components)
bool eri_wave_qt_built(int la, int lb, int lc, int ld, int fp32, int qt);

// Nuclear-gradient J/K contraction over every quartet of one class pair
// (no screening, as directScaffoldGrad): quartet q is (bra[i] | ket[j]) with
// j <= i in row-major lower-triangle order when same_class, else
// i = q / nket, j = q % nket.
struct GradLaunch {
  int la, lb, lc, ld;
  const PairRec *bra, *ket;
  int nbra, nket, same_class;
  long long nq;
  const PrimPair *prims;
  const double *prim_alpha;   // exponent of the pair's first (kernel-orientation) shell, per primitive pair
  const double *rys;
  const int *shell_atom;      // global shell -> atom
  const double *X;            // nmat Cartesian density planes, nbc x nbc each
  double *acc;                // [3 natom][nmat][nbc x nbc] Cartesian accumulators
  int nmat, nbc;
  unsigned kmask;             // bit m: matrix m is an exchange contraction
  int chunk;                  // Cartesian components of a quartet per pass (grad_chunk)
  int rpp;                    // Rys roots whose 2D tables are built at a time (grad_chunk)
};
hipError_t launch_eri_grad(const GradLaunch &a, hipStream_t s);
// LDS bytes of a gradient launch that builds chunk Cartesian components per
// pass with the 2D tables of rpp roots at a time, and the chunk (a multiple
// of 64, or the whole quartet) and rpp within lds_max bytes that take the
// fewest passes (then the most roots per table pass); 0 when even one
// root's tables do not fit
size_t grad_lds_bytes(int la, int lb, int lc, int ld, int chunk, int rpp);
int grad_chunk(int la, int lb, int lc, int ld, size_t lds_max, int *rpp);

struct ListLaunch {
  const PairScr *bra, *ket;   // screening records of the bra / ket classes
  int same_class;
  const long long *cand_off;  // nbra + 1 prefix of candidate counts (from cand_base)
  long long cand_base;        // cand_off[0]: the list covers candidates cand_base + [0, ncand)
  const int *ket_lo;          // per bra: first ket of its molecule in the ket class
  int nbra;
  long long ncand;
  // a list sized before the candidates are counted: ncand is its capacity
  // and the candidate total is read here (the scan's last entry, absolute)
  const long long *ncand_dev;
  // block norms of molecule m: Bg[boff + s1 + s2 ns], per matrix
  // Bm[(boff + s1 + s2 ns) nmat + k], boff = mol_boff[m], ns = mol_ns[m]
  const double *Bg;
  const double *Bm;
  const long long *mol_boff;
  const int *mol_ns;
  // per pair of the bra / ket class: its own block norm B(s1, s2), global
  // then per matrix (nmat + 1 doubles); their max bounds the six norms of a
  // quartet from below, which accepts most candidates without the four
  // cross-norm loads
  const double *bra_pn, *ket_pn;
  int nmat;
  double thresh;
  int screen;
  Quartet *out;
  unsigned long long *count;
  double *prim_sum;           // sum over survivors of Ka Kb Kc Kd
  long long cap;              // entries out (out32) holds; the counters keep
  long long cap32;            // counting past them
  // survivors with Schwarz product Q_ab Q_cd < cut32 go to the FP32 list
  // (out32 / count32 / prim_sum32) instead; cut32 = 0 keeps one list
  double cut32;
  Quartet *out32;
  unsigned long long *count32;
  double *prim_sum32;
};
// All lists of a batch in one launch: lists[k] (device memory) covers the
// 1024-candidate chunks [start[k], start[k + 1]) of a common chunk space
// (start: nlist + 1 entries on the device, nchunk = start[nlist]).
constexpr int kListChunk = 1024;
// dst[i] = src[idx[i]], i < n (one launch; reads back counters in one copy)
hipError_t launch_gather(const long long *src, const long long *idx, int n, long long *dst, hipStream_t s);
hipError_t launch_quartet_lists(const ListLaunch *lists, const long long *start, int nlist, long long nchunk,
                                hipStream_t s);
// Tile lists of the tile class pairs of a build: sub-block t of a list
// gives tile t of each precision (FP32 when Q_ab Q_cd < cut32), an empty
// one (bits = 0) when nothing survives.  The Schwarz pre-cut (qq x the
// molecule's largest block norm) drops a whole sub-block; the candidates of
// the rest pass the same screen as the quartet lists.  One block takes
// kTileChunk sub-blocks of one list and adds its counts once.
struct TileListLaunch {
  const SubBlock *sub;
  const PairScr *bra, *ket;   // screening records of the bra / ket classes
  const int *bmem, *kmem;     // group-ordered member arrays of the classes
  const double *bra_pn, *ket_pn;
  const double *Bg, *Bm;
  const long long *mol_boff;
  const int *mol_ns;
  const double *molmax;
  int nmat;
  double thresh;
  int screen;
  Tile *out;                   // one tile per sub-block
  unsigned long long *nquart;  // quartets of the list
  unsigned long long *ncand;   // candidates of the sub-blocks past the pre-cut
  double *prim_sum;            // sum over quartets of Ka Kb Kc Kd
  double cut32;
  Tile *out32;
  unsigned long long *nquart32;
  double *prim_sum32;
};
// lists[k] (device) covers sub-blocks [start[k], start[k + 1]) of a common
// sub-block space of nsub = start[nlist] entries (multiples of kTileChunk)
hipError_t launch_tile_lists(const TileListLaunch *lists, const long long *start, int nlist, long long nsub,
                             hipStream_t s);

// ERI + digestion of a tile list (lane-group classes, Hermitian J / K
// planes only): every quartet's shared targets -- J(a, b), J(d, c), K(a, c),
// K(d, a), K(b, c) -- are summed in LDS over the tile and added to HBM once
// per tile; K(d, b) is added per quartet.
struct TileEriLaunch {
  int la, lb, lc, ld;
  const Tile *tiles;
  long long ntile;    // tiles of the list, empty ones included
  const int *bmem, *kmem;
  const PairRec *bra, *ket;
  const PrimPair *prims;
  const PrimC32 *primc32;
  const double *rys;
  const float *rys32;
  int fp32;
  double mean_prims;  // mean primitive quartets per quartet
  int force_g;        // lanes per quartet (0: chosen per launch)
  int rys_lds;        // stage the class's Rys table section in LDS
  PlaneSet planes;
};
hipError_t launch_eri_tile(const TileEriLaunch &a, hipStream_t s);
// LDS accumulator doubles per wave of a tile launch (0: the class has no
// tile kernel)
int tile_acc_doubles(int la, int lb, int lc, int ld, const PlaneSet &ps);

// pn[i (nmat + 1) + k] = block norm of pair i's own shells (k = 0 global,
// k = 1 + m matrix m)
hipError_t launch_pair_norms(const PairScr *pairs, int npair, const double *Bg, const double *Bm,
                             const long long *mol_boff, const int *mol_ns, int nmat, double *pn,
                             hipStream_t s);

// Candidate kets of every bra pair of one class pair: off[i + 1] - off[i]
// kets starting at klo[i] (the bra's molecule's first ket), off an
// inclusive prefix (off[0] = 0).  tmp == nullptr queries tmp_bytes.
struct CandLaunch {
  const PairRec *bra, *ket;
  int nbra;
  int bra_base;            // index of bra[0] in its class array (same-class limit j <= i)
  const int *seg;          // ket class: molecule m's kets [seg[m], seg[m + 1])
  const double *molmax;    // per molecule: largest shell-block norm
  double thresh;
  int screen, same_class;
  long long *off;          // nbra + 1
  int *klo;                // nbra
};
hipError_t launch_candidates(const CandLaunch &a, void *tmp, size_t &tmp_bytes, hipStream_t s);

// Shell-block infinity norms (max absolute row sum) of nmat spherical
// matrices; symmetrised for non-Hermitian ones; global max over matrices.
hipError_t launch_block_norms(const double *X, int nmat, int is_complex,
                              const int *her, int nb, int ns,
                              const int *sh_bf, int bf0, const int *sh_nf,
                              double *Bm, double *Bg, hipStream_t s);
// out[m] = max of molecule m's global block norms
hipError_t launch_mol_max(const double *Bg, const long long *mol_boff, const int *mol_ns, int nmol,
                          double *out, hipStream_t s);

// Cartesian <-> pure transforms of whole matrices.
struct CartMap {
  const int *cart_shell;  // Cartesian function -> shell
  const int *cart_comp;   // Cartesian function -> component within shell
  const int *sh_bf;       // shell -> first pure-or-Cartesian function
  const int *sh_bfc;      // shell -> first Cartesian function
  const int *sh_l;
  const int *sh_pure;
  const double *sph;      // packed solid harmonics: l-blocks of (2l+1) x ncart(l)
  const int *bf_shell;    // pure-or-Cartesian function -> shell
  const int *bf_comp;     // pure-or-Cartesian function -> index within shell
  int nb, nbc, ns;
  // one molecule of a batch: the function tables above start at its first
  // function, and bf0 / bfc0 (its first function / Cartesian function) are
  // subtracted from the global sh_bf / sh_bfc values
  int bf0, bfc0;
};
// X (NB x NB, real or interleaved complex) -> real planes of T X T^T, written
// with interleave stride ps (plane pointers already offset to their slot)
hipError_t launch_to_cart(const CartMap &m, const double *X, int is_complex,
                          double *plane_re, double *plane_im, int ps, hipStream_t s);
// AX += finalize(T^T A T): her = 1 gives 1/2 (A + A^H), her = 0 gives 1/2 A;
// A_im may be null (real accumulator only); planes interleaved with stride ps.
hipError_t launch_from_cart(const CartMap &m, const double *A_re, const double *A_im, int ps,
                            int her, int is_complex, double *AX, hipStream_t s);

hipError_t launch_zero(double *p, size_t n, hipStream_t s);

}  // namespace cqx
