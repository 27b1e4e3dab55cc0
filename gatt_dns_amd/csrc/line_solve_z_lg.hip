// Lane-group line solves along z (line_solve_impl.h).
#include "line_solve_impl.h"

namespace gatt {
template void launch_line_solve_dir<2>(const SolveArgs&, const LineShape&, int, hipStream_t);
}  // namespace gatt
