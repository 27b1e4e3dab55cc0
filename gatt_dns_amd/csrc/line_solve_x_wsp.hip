// Fused wave-segment line solves along x (line_solve_wsp.h).
#include "line_solve_wsp.h"

namespace gatt {
template void launch_line_solve_wsp_dir<0>(const SolveArgs&, const LineShape&, int, hipStream_t);
}  // namespace gatt
