// One-wavefront-per-line solves along x and their transposes (line_solve_xg.h).
#include "line_solve_xg.h"

namespace gatt {
void launch_line_solve_xg(const SolveArgs& a, const LineShape& s, int nmembers, hipStream_t st) {
  launch_line_solve_xg_impl(a, s, nmembers, st);
}
bool launch_line_solve_xgt(const BandTArgs& a, int nmembers, hipStream_t st) {
  return launch_line_solve_xgt_impl(a, nmembers, st);
}
}  // namespace gatt
