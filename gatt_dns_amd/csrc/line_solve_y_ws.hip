// Wave-segment line solves along y (line_solve_ws.h).
#include "line_solve_ws.h"

namespace gatt {
template void launch_line_solve_ws_dir<1>(const SolveArgs&, const LineShape&, int, hipStream_t);
}  // namespace gatt
