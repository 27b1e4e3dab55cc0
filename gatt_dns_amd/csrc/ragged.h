// Ragged launches: one kernel launch that serves several contexts whose grids differ.
//
// Every kernel of the time step is written as a body that takes its arguments and its
// block position.  The plain kernel passes its own argument block and the hardware
// block indices; its ragged twin takes an array of per-member argument blocks and a
// block table (the first block of each member in one flat grid), finds the member that
// owns the block and passes that member's arguments and its block within the member's
// own grid.  The bodies read only what the plain launch would give them, so a ragged
// launch computes exactly what the members' own launches compute.
//
// On the host, a LaunchRecorder installed on the thread turns the launch functions into
// recorders: they append the ragged twin, the member's argument block and its grid
// instead of launching.  A group of contexts records one step sequence per member and
// merges the i-th launches of all members into one ragged launch.
#pragma once
#include <hip/hip_runtime.h>

#include <cstring>
#include <vector>

namespace gatt {

// A block's position within its member's grid: x, y and the grid's x extent.
struct BlockPos {
  unsigned x, y, gx;
};

struct RaggedMap {
  const int* start;  // first flat block of each member; nmem + 1 entries, the last the total
  const int* gx;     // x extent of each member's grid
  int nmem;
};

__device__ __forceinline__ BlockPos grid_pos() { return BlockPos{blockIdx.x, blockIdx.y, gridDim.x}; }

// The member owning flat block blockIdx.x: the largest m with start[m] <= block (a member
// with no blocks shares its start with the next member, the larger index), and the
// block's position in that member's grid.
__device__ __forceinline__ int ragged_pos(const RaggedMap& map, BlockPos& bp) {
  const int b = (int)blockIdx.x;
  int lo = 0, hi = map.nmem - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (map.start[mid] <= b) lo = mid;
    else hi = mid - 1;
  }
  const int lb = b - map.start[lo];
  const int gx = map.gx[lo];
  bp = BlockPos{(unsigned)(lb % gx), (unsigned)(lb / gx), (unsigned)gx};
  return lo;
}

struct LaunchRecord {
  const void* fn = nullptr;  // the ragged twin; null marks the start of a boundary application
  std::vector<char> args;    // the member's argument block
  unsigned gx = 0, gy = 0;   // the member's grid
  dim3 block;
  size_t shmem = 0;
  int bc = -1;               // the boundary application the launch belongs to, -1 outside one
};

struct LaunchRecorder {
  std::vector<LaunchRecord> recs;
  int bc = -1;               // set by the boundary program while it records its waves
  int nbc = 0;               // boundary applications recorded so far
};

// installed by the group code while it records a member's step sequence
extern thread_local LaunchRecorder* g_recorder;

// Record instead of launching when a recorder is installed on this thread.
template <class A>
inline bool record_launch(const void* ragged_fn, const A& a, dim3 grid, dim3 block, size_t shmem) {
  if (!g_recorder) return false;
  LaunchRecord r;
  r.fn = ragged_fn;
  r.args.resize(sizeof(A));
  std::memcpy(r.args.data(), &a, sizeof(A));
  r.gx = grid.x;
  r.gy = grid.y;
  r.block = block;
  r.shmem = shmem;
  r.bc = g_recorder->bc;
  g_recorder->recs.push_back(std::move(r));
  return true;
}

// A launch with no ragged twin: recording it marks the member's step sequence as one a
// ragged group cannot merge.
inline bool record_unsupported() {
  if (!g_recorder) return false;
  LaunchRecord r;
  r.bc = -2;
  g_recorder->recs.push_back(std::move(r));
  return true;
}

// The argument blocks of the small kernels that take plain parameters.
struct ClearArgs {
  unsigned long long* p;
  int n;
};
struct ClockArgs {
  double* clock;
  double dt;
};
struct CopyArgs {
  double* dst;
  const double* src;
  long long n;
};
struct RecordArgs {
  unsigned long long* red;
  double* rec;
  int* flag;
  int nmembers;
  double* clock;
  double dt;
};

}  // namespace gatt
