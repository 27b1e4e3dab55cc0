"""Finite-difference and filter stencils of the compact schemes.

Restates ``source/finiteDifferenceCoefficients.m`` (SL4, SL6, SL6O3, EX2, EX4) and
``source/spatialFilterCoefficients.m`` (10th-order Gaitonde-Visbal filter with the
'decentered', 'reducedOrder' and 'off' border closures).

Conventions (same as the reference):
  * centred stencils start at the centre point and list the right-hand side only
    (``[c0 c1 c2 ...]`` multiplies ``f(i), f(i+-1), f(i+-2)``);
  * decentred stencils are written as they appear in the top-left corner of the
    matrix, row by row.
"""
import math

import numpy as np


def finite_difference_coefficients(method):
    """Return ``(centLHS, centRHS, decLHS, decRHS)`` for a derivative scheme.

    Follows ``source/finiteDifferenceCoefficients.m:8-167``.
    """
    if method == "SL6":  # finiteDifferenceCoefficients.m:10-98
        w = 1.8
        ma = np.array([[1, 1, 1, -2],
                       [1, 4, 9, -6],
                       [1, 16, 81, -10],
                       [math.sin(w), math.sin(2 * w) / 2, math.sin(3 * w) / 3, -2 * w * math.cos(w)]])
        co = np.linalg.solve(ma, np.array([1.0, 0, 0, w]))
        a, b, c, alpha = co[0] / 2, co[1] / 4, co[2] / 6, co[3]
        cent_lhs = np.array([1.0, alpha])
        cent_rhs = np.array([0.0, a, b, c])
        # first boundary row
        ma = np.array([[1, 1, 1, 1, 1, 1, 0],
                       [0, 1, 2, 3, 4, 5, -1],
                       [0, 1, 4, 9, 16, 25, -2],
                       [0, 1, 8, 27, 64, 125, -3],
                       [0, 1, 16, 81, 256, 625, -4],
                       [0, 1, 32, 243, 1024, 3125, -5],
                       [0, 1, 64, 729, 4096, 15625, -6]], dtype=float)
        co = np.linalg.solve(ma, np.array([0.0, 1, 0, 0, 0, 0, 0]))
        lhs1 = [1.0, co[6], 0.0, 0.0]
        rhs1 = list(co[:6])
        # second row
        co = np.linalg.solve(np.array([[-1.0, 2], [-1, 6]]), np.array([-1.0, 0]))
        a, alpha = co[0] / 2, co[1]
        lhs2 = [alpha, 1.0, alpha, 0.0]
        rhs2 = [-a, 0.0, a, 0.0, 0.0, 0.0]
        # third row = centred SL4
        a, b, alpha = _sl4_centre(w)
        lhs3 = [0.0, alpha, 1.0, alpha]
        rhs3 = [-b, -a, 0.0, a, b, 0.0]
        return cent_lhs, cent_rhs, np.array([lhs1, lhs2, lhs3]), np.array([rhs1, rhs2, rhs3])

    if method == "SL6O3":  # finiteDifferenceCoefficients.m:100-119
        alpha, a, b, c = 0.392465753424658, 1.565410958904110, 0.237260273972603, -0.017739726027397
        cent_lhs = np.array([1.0, alpha])
        cent_rhs = np.array([0.0, a / 2, b / 4, c / 6])
        alpha3, a3, b3 = 0.350978473581213, 1.567318982387476, 0.134637964774951
        pi = math.pi
        dec_lhs = np.array([[1.0, (3 * pi + 40) / (3 * pi + 8), 0, 0],
                            [1 / 4, 1, 1 / 4, 0],
                            [0, alpha3, 1, alpha3]])
        den = 2 * (3 * pi + 8)
        dec_rhs = np.array([[-(13 * pi + 56) / den, (15 * pi + 8) / den, -(3 * pi - 56) / den, (pi - 8) / den, 0],
                            [-3 / 4, 0, 3 / 4, 0, 0],
                            [-b3 / 4, -a3 / 2, 0, a3 / 2, b3 / 4]])
        return cent_lhs, cent_rhs, dec_lhs, dec_rhs

    if method == "SL4":  # finiteDifferenceCoefficients.m:121-145
        a, b, alpha = _sl4_centre(1.8)
        cent_lhs = np.array([1.0, alpha])
        cent_rhs = np.array([0.0, a, b])
        dec_lhs = np.array([[1.0, 3.0, 0.0], [0.25, 1.0, 0.25]])
        dec_rhs = np.array([[-17 / 6, 3 / 2, 3 / 2, -1 / 6], [-3 / 4, 0.0, 3 / 4, 0.0]])
        return cent_lhs, cent_rhs, dec_lhs, dec_rhs

    if method == "EX2":  # finiteDifferenceCoefficients.m:148-154
        return (np.array([1.0]), np.array([0.0, 0.5]), np.array([[1.0]]),
                np.array([[-3 / 2, 2.0, -1 / 2]]))

    if method == "EX4":  # finiteDifferenceCoefficients.m:156-163
        return (np.array([1.0]), np.array([0.0, 2 / 3, -1 / 12]), np.array([[1.0]]),
                np.array([[-25 / 12, 4.0, -3.0, 4 / 3, -1 / 4], [-1 / 2, 0.0, 1 / 2, 0.0, 0.0]]))

    raise ValueError("Finite differences method not implemented: %s" % method)


def _sl4_centre(w):
    """Centred SL4 coefficients ``(a/2, b/4, alpha)`` from the 3x3 system of
    ``finiteDifferenceCoefficients.m:125-135`` (spectral-like fit at wavenumber w)."""
    ma = np.array([[1.0, 0.0, -2 / 3],
                   [0.0, 1.0, -4 / 3],
                   [math.sin(w), math.sin(2 * w) / 2, -2 * w * math.cos(w)]])
    co = np.linalg.solve(ma, np.array([4 / 3, -1 / 3, w]))
    return co[0] / 2, co[1] / 4, co[2]


def spatial_filter_coefficients(alpha, borders):
    """Return ``(LHS, RHS, decLHS, decRHS)`` of the 10th-order compact filter.

    Follows ``source/spatialFilterCoefficients.m:1-104``.  ``borders`` is one of
    'decentered', 'reducedOrder', 'off' or a bool (True -> 'decentered').
    """
    al = float(alpha)
    lhs = np.array([1.0, al])
    rhs = np.array([(193 + 126 * al) / 256, (105 + 302 * al) / 512, (-15 + 30 * al) / 128,
                    (45 - 90 * al) / 1024, (-5 + 10 * al) / 512, (1 - 2 * al) / 1024])
    if isinstance(borders, (bool, np.bool_)):
        borders = "decentered" if borders else "off"
    if borders == "decentered":  # spatialFilterCoefficients.m:18-81
        ab = np.zeros((11, 5))
        ab[:, 4] = [(-1 + 2 * al) / 1024, (5 - 10 * al) / 512, (-45 + 90 * al) / 1024, (15 + 98 * al) / 128,
                    (407 + 210 * al) / 512, (63 + 130 * al) / 256, (-105 + 210 * al) / 512, (15 - 30 * al) / 128,
                    (-45 + 90 * al) / 1024, (5 - 10 * al) / 512, (-1 + 2 * al) / 1024]
        ab[:, 3] = [(1 - 2 * al) / 1024, (-5 + 10 * al) / 512, (45 + 934 * al) / 1024, (113 + 30 * al) / 128,
                    (105 + 302 * al) / 512, (-63 + 126 * al) / 256, (105 - 210 * al) / 512, (-15 + 30 * al) / 128,
                    (45 - 90 * al) / 1024, (-5 + 10 * al) / 512, (1 - 2 * al) / 1024]
        ab[:, 2] = [(-1 + 2 * al) / 1024, (5 + 502 * al) / 512, (979 + 90 * al) / 1024, (15 + 98 * al) / 128,
                    (-105 + 210 * al) / 512, (63 - 126 * al) / 256, (-105 + 210 * al) / 512, (15 - 30 * al) / 128,
                    (-45 + 90 * al) / 1024, (5 - 10 * al) / 512, (-1 + 2 * al) / 1024]
        ab[:, 1] = [(1 + 1022 * al) / 1024, (507 + 10 * al) / 512, (45 + 934 * al) / 1024, (-15 + 30 * al) / 128,
                    (105 - 210 * al) / 512, (-63 + 126 * al) / 256, (105 - 210 * al) / 512, (-15 + 30 * al) / 128,
                    (45 - 90 * al) / 1024, (-5 + 10 * al) / 512, (1 - 2 * al) / 1024]
        ab[:, 0] = [(1023 + 1 * al) / 1024, (5 + 507 * al) / 512, (-45 + 45 * al) / 1024, (15 - 15 * al) / 128,
                    (-105 + 105 * al) / 512, (63 - 63 * al) / 256, (-105 + 105 * al) / 512, (15 - 15 * al) / 128,
                    (-45 + 45 * al) / 1024, (5 - 5 * al) / 512, (-1 + 1 * al) / 1024]
        return lhs, rhs, np.array([[1.0, al]]), ab.T.copy()
    if borders == "reducedOrder":  # spatialFilterCoefficients.m:83-97
        f2 = np.array([1 / 2 + al, 1 / 2 + al]) / np.array([1, 2])
        f4 = np.array([5 / 8 + 3 / 4 * al, 1 / 2 + al, -1 / 8 + 1 / 4 * al]) / np.array([1, 2, 2])
        f6 = np.array([11 / 16 + 5 / 8 * al, 15 / 32 + 17 / 16 * al, -3 / 16 + 3 / 8 * al,
                       1 / 32 - 1 / 16 * al]) / np.array([1, 2, 2, 2])
        f8 = np.array([93 / 128 + 70 / 128 * al, 7 / 16 + 18 / 16 * al, -7 / 32 + 14 / 32 * al,
                       1 / 16 - 1 / 8 * al, -1 / 128 + 1 / 64 * al]) / np.array([1, 2, 2, 2, 2])
        dec = np.zeros((5, 9))
        dec[0, 0] = 1.0
        dec[1, 0:3] = f2[[1, 0, 1]]
        dec[2, 0:5] = f4[[2, 1, 0, 1, 2]]
        dec[3, 0:7] = f6[[3, 2, 1, 0, 1, 2, 3]]
        dec[4, 0:9] = f8[[4, 3, 2, 1, 0, 1, 2, 3, 4]]
        return lhs, rhs, np.array([[1.0]]), dec
    if borders == "off":  # spatialFilterCoefficients.m:99-101
        return lhs, rhs, np.eye(5), np.eye(5)
    raise ValueError("unknown filter border closure %r" % (borders,))


# Extrapolation coefficients of the one-sided Neumann / null-second-derivative closures,
# getBoundaryConditions.m:75-80.
_NC1 = [[-1, 1], [-3 / 2, 2, -1 / 2], [-11 / 6, 3, -3 / 2, 1 / 3], [-25 / 12, 4, -3, 4 / 3, -1 / 4],
        [-137 / 60, 5, -5, 10 / 3, -5 / 4, 1 / 5], [-49 / 20, 6, -15 / 2, 20 / 3, -15 / 4, 6 / 5, -1 / 6]]
_NC2 = [[1, -2, 1], [2, -5, 4, -1], [35 / 12, -26 / 3, 19 / 2, -14 / 3, 11 / 12],
        [15 / 4, -77 / 6, 107 / 6, -13, 61 / 12, -5 / 6],
        [203 / 45, -87 / 5, 117 / 4, -254 / 9, 33 / 2, -27 / 5, 137 / 180],
        [469 / 90, -223 / 10, 879 / 20, -949 / 18, 41, -201 / 10, 1019 / 180, -7 / 10]]


def neumann_coefficients(order1, order2):
    """``(neumannCoeffs, neumann2Coeffs)``: ``f_b = sum_k c_k f_{b+k}`` (getBoundaryConditions.m:79-80)."""
    n1 = np.array(_NC1[order1 - 1], dtype=float)
    n2 = np.array(_NC2[order2 - 1], dtype=float)
    return -n1[1:] / n1[0], -n2[1:] / n2[0]
