// Wave-segment line solves along x (line_solve_ws.h).
#include "line_solve_ws.h"

namespace gatt {
template void launch_line_solve_ws_dir<0>(const SolveArgs&, const LineShape&, int, hipStream_t);
}  // namespace gatt
