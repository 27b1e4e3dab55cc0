"""Launch-shape autotuner of the compact line solves.

For every direction and operator (derivative, filter) of a :class:`Solver`,
times the line-solve kernel class over the candidate shapes -- the fused and the
plain wave-segment kernels with 16 or 32 points per lane, and the lane-group kernel with G lanes per
line (powers of two, at most 32 points per lane) and 1, 2 or 4 wavefronts per
workgroup -- with HIP events on the library stream, and keeps the fastest.
Filter timing overwrites the state, so call :func:`tune` before setting it (or
set it again afterwards).
"""
from .abi import GATT_DERIV, GATT_FILTER
from .runtime import GattError

KERNEL_CLASS = {(0, GATT_DERIV): 0, (1, GATT_DERIV): 1, (2, GATT_DERIV): 2,
                (0, GATT_FILTER): 3, (1, GATT_FILTER): 4, (2, GATT_FILTER): 5}


def candidates(n):
    """(variant, param, waves per block) triples; the library rejects shapes that do not fit."""
    out = [(4, M, 0) for M in (4, 8, 16)]
    out += [(v, M, 0) for v in (2, 3, 1) for M in (16, 32)]
    out += [(3, M, 2) for M in (16, 32)]   # x tile of 16 lines
    out += [(0, G, w) for G in (2, 4, 8, 16, 32, 64) if (n + G - 1) // G <= 32 for w in (1, 2, 4)]
    return out


def tune(solver, reps=5, log=None):
    """Pick the fastest shape per (direction, kind); returns {(dir, kind): (variant, param, wpb, ms)}."""
    case = solver.case
    dims = (case.nx, case.ny, case.nz)
    ndir = 3 if case.is3d else 2
    result = {}
    for d in range(ndir):
        for kind in (GATT_DERIV, GATT_FILTER):
            best = None
            for v, p, w in candidates(dims[d]):
                try:
                    solver.set_line_shape(d, kind, v, p, max(w, 1))
                except GattError:
                    continue
                ms, _ = solver.time_kernel(KERNEL_CLASS[(d, kind)], reps)
                if log:
                    log("  tune dir %d kind %d variant %d param %2d wpb %d: %.4f ms" % (d, kind, v, p, w, ms))
                if best is None or ms < best[3]:
                    best = (v, p, max(w, 1), ms)
            solver.set_line_shape(d, kind, best[0], best[1], best[2])
            result[(d, kind)] = best
    return result
