// The lane-group line solves of two or three directions in one launch.
//
// On small grids a direction solve is a few microseconds of work behind a launch of about
// the same cost, and the x, y and z solves of one RHS pass are independent.  This kernel
// runs them as one launch: the flat grid holds the workgroups of the first direction's
// launch, then the second's, then the third's, and each workgroup runs the lane-group body
// (line_solve_impl.h) of its direction on its position in that direction's own grid, so
// every line is solved exactly as its own launch would solve it.
#include "line_solve_impl.h"

namespace gatt {

template <int NR, int M>
__global__ __launch_bounds__(256) void line_solve_multi_kernel(SolveMulti t) {
  const unsigned b = blockIdx.x;
  int k = 0;
  if (t.n > 1 && b >= t.first[1]) k = (t.n > 2 && b >= t.first[2]) ? 2 : 1;
  const unsigned local = b - t.first[k];
  const BlockPos bp{local % t.gx[k], local / t.gx[k], t.gx[k]};
  switch (t.a[k].dir) {
    case 0: line_solve_lg_body<0, NR, M>(t.a[k], bp); break;
    case 1: line_solve_lg_body<1, NR, M>(t.a[k], bp); break;
    default: line_solve_lg_body<2, NR, M>(t.a[k], bp); break;
  }
}

template <int NR>
static bool launch_multi_nr(const SolveMulti& t, int M, dim3 grid, dim3 block, size_t lds, hipStream_t st) {
  switch (M) {
    case 4: hipLaunchKernelGGL((line_solve_multi_kernel<NR, 4>), grid, block, lds, st, t); return true;
    case 8: hipLaunchKernelGGL((line_solve_multi_kernel<NR, 8>), grid, block, lds, st, t); return true;
    case 16: hipLaunchKernelGGL((line_solve_multi_kernel<NR, 16>), grid, block, lds, st, t); return true;
    case 32: hipLaunchKernelGGL((line_solve_multi_kernel<NR, 32>), grid, block, lds, st, t); return true;
  }
  return false;
}

bool multi_solve_supported(int nR, int M) {
  return nR >= 2 && nR <= 6 && (M == 4 || M == 8 || M == 16 || M == 32);
}

bool launch_line_solve_multi(const SolveArgs* parts, int n, const LineShape& s, int nmembers, hipStream_t st) {
  if (n < 1 || n > 3 || s.variant != 0 || !multi_solve_supported(parts[0].op.nR, s.M)) return false;
  SolveMulti t{};
  t.n = n;
  unsigned total = 0;
  for (int k = 0; k < n; ++k) {
    if (parts[k].op.nR != parts[0].op.nR) return false;
    SolveArgs a = parts[k];
    // as launch_nr_m sets them for a launch of its own (a.G: the part's lanes per line); the
    // LDS row, TL * G + 1 = 64 * waves + 1, is the same for every part
    a.TL = (64 / a.G) * s.waves_per_block;
    a.lds_row = a.TL * a.G + 1;
    a.inv_n = 1.0f / (float)a.op.N;
    const unsigned tiles = (unsigned)((a.nlines + a.TL - 1) / a.TL);
    t.a[k] = a;
    t.first[k] = total;
    t.gx[k] = tiles;
    total += tiles * (unsigned)(a.nf * nmembers);
  }
  t.first[n] = total;
  const size_t lds = (size_t)s.M * t.a[0].lds_row * sizeof(double);
  const dim3 grid(total), block(64 * s.waves_per_block);
  switch (parts[0].op.nR) {
    case 2: return launch_multi_nr<2>(t, s.M, grid, block, lds, st);
    case 3: return launch_multi_nr<3>(t, s.M, grid, block, lds, st);
    case 4: return launch_multi_nr<4>(t, s.M, grid, block, lds, st);
    case 5: return launch_multi_nr<5>(t, s.M, grid, block, lds, st);
    case 6: return launch_multi_nr<6>(t, s.M, grid, block, lds, st);
  }
  return false;
}

}  // namespace gatt
