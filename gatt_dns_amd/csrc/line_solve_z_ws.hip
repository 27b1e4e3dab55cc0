// Wave-segment line solves along z (line_solve_ws.h).
#include "line_solve_ws.h"

namespace gatt {
template void launch_line_solve_ws_dir<2>(const SolveArgs&, const LineShape&, int, hipStream_t);
}  // namespace gatt
