// Fused wave-segment line solves along z (line_solve_wsp.h).
#include "line_solve_wsp.h"

namespace gatt {
template void launch_line_solve_wsp_dir<2>(const SolveArgs&, const LineShape&, int, hipStream_t);
}  // namespace gatt
