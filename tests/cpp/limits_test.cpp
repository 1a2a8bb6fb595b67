// CPU unit test of the launch-size and index-packing rules of the device
// kernels (chronusq_public_amd/csrc/limits.hpp): prints one line per case,
// "case ok" or "case FAIL ...", and exits non-zero on a failure.
#include <cstdio>
#include <initializer_list>

#include "limits.hpp"

static int failures = 0;
static void expect(bool ok, const char *name) {
  std::printf("%s %s\n", name, ok ? "ok" : "FAIL");
  if (!ok) ++failures;
}

int main() {
  using namespace cqx;
  // a quartet list of 2^28 entries on a wave class with one quartet per wave
  const long long n28 = 1ll << 28;
  const LaunchPlan w1 = plan_launches(n28, kFamilyWave, 1);
  expect(w1.pieces == 8, "wave_qt1_2^28_pieces");
  expect(w1.max_items < kMaxWorkItems, "wave_qt1_2^28_items_below_2^32");
  expect(w1.max_items == (kLaunchPiece / 4) * 256, "wave_qt1_piece_items");
  expect(launch_fits(n28, kFamilyWave, 1), "wave_qt1_2^28_fits");
  // the same list unpieced would need 2^34 work-items
  expect(launch_blocks(n28, kFamilyWave, 1) * 256 >= (1ll << 34), "wave_qt1_unpieced_2^34");
  // lane-group kernels, every lane count, and the grid-stride team kernel
  for (int g : {1, 2, 4, 8, 16}) expect(launch_fits(n28 * 8, kFamilyLaneGroup, g), "lanegroup_2^31_fits");
  expect(launch_blocks(n28, kFamilyTeam, 1) == kTeamGridMax, "team_grid_clamped");
  expect(plan_launches(0, kFamilyWave, 1).pieces == 0, "empty_list_no_launch");
  expect(plan_launches(1, kFamilyLaneGroup, 16).max_items == 256, "one_quartet_one_block");
  // lane-group runs: short lists keep a block per batch, long ones give a
  // block up to kLaneGroupRun batches and keep kLaneGroupMinBlocks blocks
  expect(lane_group_blocks(100) == 100, "lanegroup_short_list_block_per_batch");
  expect(lane_group_blocks(kLaneGroupMinBlocks + 1) == kLaneGroupMinBlocks, "lanegroup_min_blocks");
  expect(lane_group_blocks(kLaneGroupMinBlocks * kLaneGroupRun * 4) == kLaneGroupMinBlocks * 4, "lanegroup_run_cap");
  // the kernel's split: block i takes batches [i nb / b, (i + 1) nb / b)
  for (long long nb : {1ll, 5000ll, 8193ll, 100000ll, 131073ll, 1ll << 21}) {
    const long long b = lane_group_blocks(nb);
    bool ok = b >= 1 && b <= nb;
    long long next = 0;
    for (long long i = 0; i < b && ok; ++i) {
      const long long b0 = i * nb / b, b1 = (i + 1) * nb / b;
      ok = b0 == next && b1 - b0 >= 1 && b1 - b0 <= kLaneGroupRun;
      next = b1;
    }
    expect(ok && next == nb, "lanegroup_runs_cover_batches");
  }
  // ket indices: a class of 2^24 pairs fits, one more does not
  expect(ket_index_fits((size_t)1 << 24), "ket_2^24_fits");
  expect(!ket_index_fits(((size_t)1 << 24) + 1), "ket_2^24+1_rejected");
  return failures ? 1 : 0;
}
