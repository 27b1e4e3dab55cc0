// Lane-group line solves along y (line_solve_impl.h).
#include "line_solve_impl.h"

namespace gatt {
template void launch_line_solve_dir<1>(const SolveArgs&, const LineShape&, int, hipStream_t);
}  // namespace gatt
