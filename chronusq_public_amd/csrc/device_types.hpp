// Data shared between host orchestration and the HIP kernels.
#pragma once

#include <cstdint>

namespace cqx {

constexpr int kMaxLDev = 4;    // device kernels cover s..g shells
constexpr int kMaxPlanes = 16; // real accumulation planes per launch
constexpr int kMaxMat = 8;     // contractions sharing one ERI evaluation

// A significant shell pair in kernel orientation: shell a carries the higher
// angular momentum (la >= lb).  s1 >= s2 is the same pair in the order the
// reference loop visits it (direct.hpp:1659-1664), used for screening.
struct PairRec {
  int a, b;        // kernel-orientation shells (global indices)
  int s1, s2;      // reference-orientation shells, numbered within the molecule
  int pp_off, npp; // primitive pairs [pp_off, pp_off + npp)
  int bfa, bfb;    // first Cartesian function of a and b within the molecule
  double Q;        // Schwarz bound of the pair
  double AB[3];    // A - B
  int kk;          // Ka * Kb (contraction lengths, before primitive screening)
  int mol;         // molecule of a batched context (0 for one molecule)
  double A[3];     // centre of shell a
};

// What the quartet screen reads of a shell pair, 16 bytes: shells s1, s2
// (numbered within the molecule) in the low and high 16 bits of s12, the
// contraction-length product Ka Kb, the molecule and the Schwarz bound.
struct PairScr {
  uint32_t s12;
  uint16_t kk;
  uint16_t mol;
  double Q;
};

// One surviving primitive pair of a shell pair (kernel orientation).
struct PrimPair {
  double p;        // alpha_a + alpha_b
  double P[3];     // Gaussian product centre
  double PA[3];    // P - A
  double K;        // c_a c_b exp(-alpha_a alpha_b / p |AB|^2)
};

// A primitive pair as the FP32 lists read it, in one 16-byte load: with b
// the exponent of shell b, P - A = -(b / p) AB follows from the shell pair's
// AB, and P - Q = (A - C) + (P - A) - (Q - C) from the quartet's A - C, so
// neither centre is stored (K stays FP64: products of tiny prefactors stay
// representable).  Same indexing as PrimPair.
struct alignas(16) PrimC32 {
  float p;         // alpha_a + alpha_b
  float bp;        // alpha_b / p
  double K;        // c_a c_b exp(-alpha_a alpha_b / p |AB|^2)
};

// Digestion target: one real plane of a (possibly complex) contraction.
enum PlaneKind : int { kJHer = 0, kKHer = 1, kJNonHer = 2, kKNonHer = 3 };

// Planes are interleaved: element (i, j) of plane p lives at
// base[(i + j * NBc) * PlaneSet::ps + p]; X / A below already include + p.
struct Plane {
  int kind;         // PlaneKind
  double sign;      // +1 real part; -1 imaginary part of a conjugated X
  int mat;          // contraction index (screening mask bit)
  int pad;
};

// The Cartesian density and accumulator planes of one molecule: plane
// kstart[k] + i of element (r, c) sits at kx[k][(r + c nbc) kps[k] + i]
// (accumulators ka[k]); r, c count the molecule's Cartesian functions.
struct MolPlanes {
  const double *kx[4];
  double *ka[4];
  int nbc;
  int pad;
};

struct PlaneSet {
  Plane p[kMaxPlanes];
  int n;
  int kstart[4], kcount[4];  // planes are ordered by kind: [kstart, kstart + kcount)
  // The planes of one kind are interleaved in an array of their own, stride
  // kps[k] = kcount[k] rounded up to a power of two: a lone plane keeps
  // adjacent rows adjacent in memory, a group of planes shares one 64-byte
  // segment per element.  Per molecule: mols[PairRec::mol].
  const MolPlanes *mols;
  MolPlanes m0;  // mols[0], held in the kernel arguments for one-molecule launches
  int nmol;
  int kps[4];
  unsigned long long mats;   // contraction index of slot s in bits [4s, 4s + 4)
  unsigned negs;             // bit s: the slot's K coefficient is negated (conj of Im X)
};

// Tiles of quartets that share shells.  The pairs of a class are grouped by
// (molecule, first shell in kernel orientation); a class's member array
// lists its pair indices group by group.  A sub-block is up to kTileDim
// consecutive members of one bra group against up to kTileDim consecutive
// members of one ket group, so every quartet of it shares the bra's first
// shell a and the ket's first shell c; its surviving candidates make a Tile.
constexpr int kTileDim = 8;
// a tile class pair's sub-blocks are padded to a multiple of this many (empty
// sub-blocks with nb = nk = 0), so that one block of the list kernel takes a
// chunk of one list
constexpr int kTileChunk = 64;
struct SubBlock {
  int bra_m0, ket_m0;     // first bra / ket member of the chunk
  unsigned short nb, nk;  // chunk sizes, <= kTileDim
  int diag;               // same class, group and chunk on both sides: only i >= j
  double qq;              // max Q of the bra chunk x max Q of the ket chunk
};
// Candidate (i, j) = bra member bra_m0 + i against ket member ket_m0 + j is
// bit i kTileDim + j of bits; mask[i kTileDim + j] holds the contractions
// that passed their separated screen.
struct Tile {
  int bra_m0, ket_m0;
  unsigned short nb, nk;
  unsigned pad;
  unsigned long long bits;
  unsigned char mask[kTileDim * kTileDim];
  unsigned long long pad2;
};

// A quartet: bra / ket pair indices within their class arrays plus the
// mask of contractions whose separated screen it passed.
struct Quartet {
  uint32_t bra;
  uint32_t ket_mask;  // ket index (low 24 bits) | mask << 24
};

}  // namespace cqx
