// One-wavefront-per-line solves along x (line_solve_xg.h).
#include "line_solve_xg.h"

namespace gatt {
void launch_line_solve_xg(const SolveArgs& a, const LineShape& s, int nmembers, hipStream_t st) {
  launch_line_solve_xg_impl(a, s, nmembers, st);
}
}  // namespace gatt
