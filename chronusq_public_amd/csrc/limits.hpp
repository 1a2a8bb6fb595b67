// Index-packing and launch-size limits of the device kernels, checked on the
// host before any launch.  Pure host code (no HIP), so that the sizing rules
// can be exercised without a GPU.
#pragma once

#include <cstddef>
#include <cstdint>

namespace cqx {

// Quartet::ket_mask carries the ket's index within its class array in the
// low 24 bits (the contraction mask in the high 8), so a class array may
// hold at most 2^24 shell pairs.
constexpr long long kKetIndexLimit = 1ll << 24;
inline bool ket_index_fits(size_t class_pairs) { return (long long)class_pairs <= kKetIndexLimit; }

// Kernel families of the ERI launches.
enum EriFamily : int { kFamilyLaneGroup = 0, kFamilyWave = 1, kFamilyTeam = 2 };

// Quartets one ERI launch takes at most: a longer list is launched in pieces.
constexpr long long kLaunchPiece = 1ll << 25;
// Work-items of one launch must stay below 2^32.
constexpr long long kMaxWorkItems = 1ll << 32;
// Blocks of the generic (team) kernel, which loops over its quartets.
constexpr long long kTeamGridMax = 65536;

// Threads per block and quartets per block of a family; shape = lanes per
// quartet (lane-group kernels: 1, 2, 4, 8, 16) or quartets per wave (wave
// kernels: 1, 2, 4, 8).
inline int family_threads(int family) { return family == kFamilyTeam ? 64 : 256; }
inline long long quartets_per_block(int family, int shape) {
  if (family == kFamilyLaneGroup) return 4ll * (64 / (shape > 0 ? shape : 1));
  if (family == kFamilyWave) return 4ll * (shape > 0 ? shape : 1);
  return 1;  // team kernel: one wave per quartet, grid-stride
}

// A lane-group block takes a run of consecutive batches (a batch: the
// quartets its 4 waves take at once), so that its copy of the Rys table
// section into LDS serves more quartets: the batches are split into runs of
// at most kLaneGroupRun (lengths differing by at most one), and a launch
// keeps kLaneGroupMinBlocks blocks when it has that many batches.
constexpr long long kLaneGroupRun = 16;
constexpr long long kLaneGroupMinBlocks = 2048;
inline long long lane_group_blocks(long long nbatch) {
  if (nbatch <= kLaneGroupMinBlocks) return nbatch;
  const long long by_run = (nbatch + kLaneGroupRun - 1) / kLaneGroupRun;
  return by_run > kLaneGroupMinBlocks ? by_run : kLaneGroupMinBlocks;
}

// Blocks of one launch of nq quartets (nq <= kLaunchPiece).
inline long long launch_blocks(long long nq, int family, int shape) {
  if (nq <= 0) return 0;
  if (family == kFamilyTeam) return nq < kTeamGridMax ? nq : kTeamGridMax;
  const long long per = quartets_per_block(family, shape);
  const long long nbatch = (nq + per - 1) / per;
  return family == kFamilyLaneGroup ? lane_group_blocks(nbatch) : nbatch;
}

// Pieces a list of nq quartets is launched in, and the largest launch's
// work-items; every piece covers at most kLaunchPiece quartets.
struct LaunchPlan {
  long long pieces;
  long long max_items;
};
inline LaunchPlan plan_launches(long long nq, int family, int shape) {
  LaunchPlan p{0, 0};
  if (nq <= 0) return p;
  p.pieces = (nq + kLaunchPiece - 1) / kLaunchPiece;
  const long long first = nq < kLaunchPiece ? nq : kLaunchPiece;
  p.max_items = launch_blocks(first, family, shape) * family_threads(family);
  return p;
}
inline bool launch_fits(long long nq, int family, int shape) {
  return plan_launches(nq, family, shape).max_items < kMaxWorkItems;
}

}  // namespace cqx
