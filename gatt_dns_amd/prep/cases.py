"""Problem descriptions for the time-advance path.

A :class:`Case` bundles everything the reference bakes into its generated
Fortran include files (``parameters.F90``, ``matrices.F90``,
``boundaryInfo.F90``, ``runDisturbances.F90`` / ``runForcings*.F90``) plus
the initial state.  Scalars and coefficient arrays are passed through
:func:`fortran_round` exactly as the reference's ``%.20fd0`` literals are.

Builders:
  * :func:`lid_driven`      -- ``examples/lidDriven.m`` on a uniform mesh
  * :func:`turbulence_box`  -- ``examples/turbulenceBox.m`` (periodic box + forcing)
  * :func:`poiseuille`      -- ``examples/poiseuille.m`` (channel + pressureGradient forcing)
"""
from dataclasses import dataclass, field

import numpy as np

from .boundaries import (BoundaryInfo, boundary_layer_free_slip, boundary_layer_isothermal, calc_sfd_region, disturbance_index, init_boundaries,
                         lid_driven_flow, periodic_box, poiseuille_flow)
from .matrices import BufferSide, DirMesh, LineCoeffs, NumMethods, fortran_round, make_matrices, prepare_thomas
from .compressible_bl import compressible_bl_flow
from .mesh import MeshSpec, generate_mesh

FIELDS = ("U", "V", "W", "R", "E")


@dataclass
class Disturbance:
    """One entry of ``runDisturbances.F90`` (kind 'lidDrivenMovement', 'holdInlet',
    'periodicSine' or 'packet_2d'); ``var`` is the flow variable passed as ``u``."""
    kind: str
    box: list
    X: np.ndarray
    Y: np.ndarray
    Z: np.ndarray
    par: list
    var: str = ""


@dataclass
class Forcing:
    """One entry of ``runForcings{2,3}D.F90`` (kind 'turbulenceGenerator' or
    'pressureGradient'); ``var`` names the variables passed ('RU' for pressureGradient)."""
    kind: str
    box: list
    par: list
    var: str = ""


@dataclass
class Case:
    name: str
    nx: int
    ny: int
    nz: int
    X: np.ndarray
    Y: np.ndarray
    Z: np.ndarray
    Re: float
    Ma: float
    Pr: float
    gamma: float
    T0: float
    deriv: list            # [LineCoeffs x, y, z]  (z is None in 2-D)
    filt: list
    types: list            # [typeMapX (ny,nz), typeMapY (nx,nz), typeMapZ (nx,ny)] 1-based
    filter_dirs: tuple
    filter_char_time: float  # FilterCharTime as written to parameters.F90 (-1 => full replacement)
    bi: BoundaryInfo
    disturb: list = field(default_factory=list)
    forcing: list = field(default_factory=list)
    sfd_type: int = 0
    sfd_delta: float = 0.0
    sfd_x_val: float = 0.0
    sfd_x: np.ndarray = None
    dxmin: tuple = (0.0, 0.0, 0.0)
    time_stepping: int = 1           # timeStepping: 1 RK4, 2 Euler, 3 SSPRK3 (writeFortranParameters.m:50-58)
    dt: float = 0.0
    flow_region: np.ndarray = None   # True at flow points (False inside walls)
    state0: dict = None

    @property
    def is3d(self):
        return self.nz > 1

    @property
    def npoints(self):
        return self.nx * self.ny * self.nz

    @property
    def E0(self):
        return 1.0 / ((self.gamma ** 2 - self.gamma) * self.Ma ** 2)


def _coeffs(meshes, flow_region, nm):
    mats = make_matrices(meshes, flow_region, nm)
    deriv, filt, types = [], [], []
    for key, dm in zip("xyz", meshes):
        types.append(mats[key].types.astype(np.int32))
        if key == "z" and dm.n == 1:
            deriv.append(None)
            filt.append(None)
            continue
        d, f = prepare_thomas(mats[key])
        deriv.append(d)
        filt.append(f)
    return deriv, filt, types


def _dxmin(X, Y, Z, nz):
    """``dxmin`` of writeFortranParameters.m:79-89."""
    dx = [1.0 / np.min(np.diff(X)), 1.0 / np.min(np.diff(Y)), 0.0 if nz == 1 else 1.0 / np.min(np.diff(Z))]
    return tuple(float(v) for v in fortran_round(np.array(dx)))


def cfl_dt(state, Ma, dxmin, nz):
    """CFL number divided by dt (main.F90:196-214)."""
    um = np.max(np.abs(state["U"]))
    vm = np.max(np.abs(state["V"]))
    c = (1 / Ma + um) * dxmin[0] + (1 / Ma + vm) * dxmin[1]
    if nz > 1:
        c += (1 / Ma + np.max(np.abs(state["W"]))) * dxmin[2]
    return c


def lid_driven(n=129, Re=1000.0, Ma=0.5, Pr=0.71, gamma=1.4, T0=300.0, alpha_f=0.49,
               spatial="SL4", cfl=0.5, nx=None, ny=None, lid_u0=1.0, roughness=()):
    """2-D lid-driven cavity on [0,1]^2 with a uniform mesh (examples/lidDriven.m, uniform variant).
    ``roughness``: blocks (dicts with x, y, z intervals) closed off as walls inside the cavity
    (lidDrivenFlow.m:16-25).

    Filter borders 'reducedOrder' with the x start/end and y end closures switched
    off (``examples/lidDriven.m:119-122``).  dt = cfl / CFLdt on the initial state
    after the first boundary application.
    """
    nx = nx or n
    ny = ny or n
    X = np.linspace(0.0, 1.0, nx)
    Y = np.linspace(0.0, 1.0, ny)
    Z = np.array([0.0])
    meshes = [DirMesh(X), DirMesh(Y), DirMesh(Z, periodic=True, fix_periodic_domain_size=True, xf=1.0)]
    E0 = 1.0 / ((gamma ** 2 - gamma) * Ma ** 2)
    P0 = (gamma - 1) * E0
    fr, walls, corners, bc = lid_driven_flow(nx, ny, 1, E0, P0, X=X, Y=Y, Z=Z, roughness=roughness)
    nm = NumMethods(spatial_derivs=spatial, spatial_derivs_buffer="EX4", filter_strength=alpha_f,
                    filter_borders="reducedOrder", filter_borders_start=(False, True, True),
                    filter_borders_end=(False, False, True))
    deriv, filt, types = _coeffs(meshes, fr, nm)
    bi = init_boundaries(bc, corners, nm.neumann_order, nm.neumann2_order)
    box = disturbance_index(X, Y, Z, (0.0, 1.0), (1.0, 1.0), (-np.inf, np.inf))
    dist = Disturbance("lidDrivenMovement", box, fortran_round(X[box[0] - 1:box[1]]),
                       fortran_round(Y[box[2] - 1:box[3]]), fortran_round(Z[box[4] - 1:box[5]]),
                       [float(fortran_round(lid_u0))])
    rnd = lambda v: float(fortran_round(v))
    case = Case("lid_driven_%dx%d" % (nx, ny) + ("_rug%d" % len(roughness) if roughness else ""), nx, ny, 1, X, Y, Z, rnd(Re), rnd(Ma), rnd(Pr), rnd(gamma),
                rnd(T0), deriv, filt, types, (1, 1, 1), -1.0, bi, [dist], [],
                dxmin=_dxmin(X, Y, Z, 1))
    case.flow_region = fr
    st = {k: np.zeros((nx, ny, 1), order="F") for k in FIELDS}
    st["R"][:] = 1.0
    st["E"][:] = E0
    case.state0 = st
    # dt from the CFL of the state after the first boundary application (lid at U0)
    umax = max(abs(lid_u0), 0.0)
    cfl_dt_val = (1 / case.Ma + umax) * case.dxmin[0] + (1 / case.Ma) * case.dxmin[1]
    case.dt = rnd(cfl / cfl_dt_val)
    return case


def solenoidal_field(n, seed, k0=4.0, urms=0.5):
    """Random-phase solenoidal velocity with E(k) ~ k^4 exp(-2 (k/k0)^2) on an n^3 periodic box."""
    rng = np.random.Generator(np.random.PCG64(seed))
    k1 = np.fft.fftfreq(n, 1.0 / n)
    kx, ky, kz = np.meshgrid(k1, k1, k1, indexing="ij")
    k2 = kx ** 2 + ky ** 2 + kz ** 2
    kk = np.sqrt(k2)
    amp = np.zeros_like(kk)
    nz_ = kk > 0
    # shell energy E(k) spread over 4 pi k^2 modes -> |u_hat| ~ sqrt(E(k)/k^2)
    amp[nz_] = np.sqrt(kk[nz_] ** 4 * np.exp(-2 * (kk[nz_] / k0) ** 2) / kk[nz_] ** 2)
    u = []
    for _ in range(3):
        ph = rng.uniform(0, 2 * np.pi, kk.shape)
        u.append(amp * np.exp(1j * ph))
    # project out the k-parallel part
    with np.errstate(invalid="ignore", divide="ignore"):
        dot = (kx * u[0] + ky * u[1] + kz * u[2]) / np.where(nz_, k2, 1.0)
    u = [u[0] - kx * dot, u[1] - ky * dot, u[2] - kz * dot]
    vel = [np.real(np.fft.ifftn(c)) for c in u]
    rms = np.sqrt(np.mean(vel[0] ** 2 + vel[1] ** 2 + vel[2] ** 2) / 3)
    return [np.asfortranarray(v * (urms / rms)) for v in vel]


def turbulence_box(n=256, Re=1000.0, Ma=0.3, Pr=0.71, gamma=1.4, T0=300.0, alpha_f=0.49,
                   spatial="SL4", cfl=0.5, seed=20260, urms=0.5, epsilon=1.0, forcing=True):
    """3-D periodic turbulence box (examples/turbulenceBox.m with the sizes of the 256^3 config)."""
    X = np.arange(n) / n
    meshes = [DirMesh(X.copy(), periodic=True, fix_periodic_domain_size=True, xf=1.0) for _ in range(3)]
    E0 = 1.0 / ((gamma ** 2 - gamma) * Ma ** 2)
    fr, walls, corners, bc = periodic_box(n, n, n)
    nm = NumMethods(spatial_derivs=spatial, spatial_derivs_buffer="EX4", filter_strength=alpha_f,
                    filter_borders="reducedOrder")
    deriv, filt, types = _coeffs(meshes, fr, nm)
    bi = init_boundaries(bc, corners, nm.neumann_order, nm.neumann2_order)
    rnd = lambda v: float(fortran_round(v))
    forc = []
    if forcing:
        forc.append(Forcing("turbulenceGenerator", [1, n, 1, n, 1, n], [rnd(epsilon), rnd(E0)]))
    case = Case("turbulence_box_%d" % n, n, n, n, X, X.copy(), X.copy(), rnd(Re), rnd(Ma), rnd(Pr),
                rnd(gamma), rnd(T0), deriv, filt, types, (1, 1, 1), -1.0, bi, [], forc,
                dxmin=_dxmin(X, X, X, n))
    case.flow_region = fr
    u, v, w = solenoidal_field(n, seed, urms=urms)
    st = {"U": u, "V": v, "W": w, "R": np.ones((n, n, n), order="F"), "E": np.full((n, n, n), E0, order="F")}
    case.state0 = st
    case.dt = rnd(cfl / cfl_dt(st, case.Ma, case.dxmin, n))
    return case


def periodic_rest(nx, ny, nz=1, Re=1000.0, Ma=0.3, Pr=0.71, gamma=1.4, T0=300.0, alpha_f=0.49,
                  spatial="SL4", cfl=0.5, stretch=0.0):
    """Fluid at rest in a periodic box (periodicBox.m) on uniform meshes, no forcing.

    The linearised step about this state is diagonal in Fourier modes, so the
    spectrum of the time-T propagator is known in closed form; used to check
    the stability analysis.  ``stretch`` = a in [0, 1) makes the meshes
    non-uniform, x_i = (i + a n / (2 pi) sin(2 pi (i + 1/2) / n)) / n shifted to x_0 = 0,
    still periodic on [0, 1) (dX/deta then differs between the first and the last node, where the
    periodic solve's index quirks of derivs.F90:786-802 show).
    """
    meshes = []
    for n in (nx, ny, nz):
        i = np.arange(n)
        X = (i + stretch * n / (2 * np.pi) * np.sin(2 * np.pi * (i + 0.5) / n)) / n if n > 1 else np.array([0.0])
        if n > 1 and stretch:
            X = X - X[0]
        meshes.append(DirMesh(X, periodic=True, fix_periodic_domain_size=True, xf=1.0))
    X, Y, Z = (m.X for m in meshes)
    E0 = 1.0 / ((gamma ** 2 - gamma) * Ma ** 2)
    fr, walls, corners, bc = periodic_box(nx, ny, nz)
    nm = NumMethods(spatial_derivs=spatial, spatial_derivs_buffer="EX4", filter_strength=alpha_f,
                    filter_borders="reducedOrder")
    deriv, filt, types = _coeffs(meshes, fr, nm)
    bi = init_boundaries(bc, corners, nm.neumann_order, nm.neumann2_order)
    rnd = lambda v: float(fortran_round(v))
    case = Case("periodic_rest_%dx%dx%d" % (nx, ny, nz), nx, ny, nz, X, Y, Z, rnd(Re), rnd(Ma), rnd(Pr),
                rnd(gamma), rnd(T0), deriv, filt, types, (1, 1, 1), -1.0, bi, [], [],
                dxmin=_dxmin(X, Y, Z, nz))
    case.flow_region = fr
    shape = (nx, ny, nz)
    st = {k: np.zeros(shape, order="F") for k in FIELDS}
    st["R"][:] = 1.0
    st["E"][:] = E0
    case.state0 = st
    case.dt = rnd(cfl / cfl_dt(st, case.Ma, case.dxmin, nz))
    return case


def poiseuille(nx=64, ny=48, nz=1, Re=2000.0, Ma=0.5, Pr=0.71, gamma=1.4, T0=300.0, L=4.0, H=1.0,
               dpdx=None, alpha_f=0.49, spatial="SL6", cfl=0.5, profile=True, disturb=(), y_mesh="tanh"):
    """Plane channel of ``examples/poiseuille.m``: x periodic (uniform, length L), y on
    [-H/2, H/2] with the tanh mesh refined at both walls, z periodic in 3-D, and the
    pressureGradient forcing ``dU += dpdx/R`` everywhere (``dpdx = 16/Re``).

    The x mesh uses a fixed periodic domain size (the example's ``fixPeriodicDomainSize =
    false`` mesh needs ``generateMesh``).  The initial state is the laminar parabola
    ``U = dpdx Re / (2 mu) (H^2/4 - y^2)`` at rest temperature (``profile=True``) or
    ``U = 1`` with no-slip walls (``profile=False``).  ``disturb`` adds entries of
    ``runDisturbances.F90``: tuples ``(kind, x range, y range, var, par)``.
    """
    dpdx = 16.0 / Re if dpdx is None else dpdx
    X = np.arange(nx) / nx * L
    Y = generate_mesh(-H / 2, H / 2, MeshSpec(n=ny, type="tanh", local="b", par=2.0)) if y_mesh == "tanh" \
        else np.linspace(-H / 2, H / 2, ny)
    Z = np.arange(nz) / nz if nz > 1 else np.array([0.0])
    meshes = [DirMesh(X, periodic=True, fix_periodic_domain_size=True, xf=L), DirMesh(Y),
              DirMesh(Z, periodic=True, fix_periodic_domain_size=True, xf=1.0)]
    E0 = 1.0 / ((gamma ** 2 - gamma) * Ma ** 2)
    P0 = (gamma - 1) * E0
    fr, walls, corners, bc = poiseuille_flow(nx, ny, nz, E0, P0)
    nm = NumMethods(spatial_derivs=spatial, spatial_derivs_buffer="EX4", filter_strength=alpha_f,
                    filter_borders="reducedOrder", filter_borders_start=(False, True, True),
                    filter_borders_end=(False, False, True))
    deriv, filt, types = _coeffs(meshes, fr, nm)
    bi = init_boundaries(bc, corners, nm.neumann_order, nm.neumann2_order)
    rnd = lambda v: float(fortran_round(v))
    whole = disturbance_index(X, Y, Z, (-np.inf, np.inf), (-np.inf, np.inf), (-np.inf, np.inf))
    forc = [Forcing("pressureGradient", whole, [rnd(dpdx)], var="RU")]
    dist = []
    for kind, xr, yr, var, par in disturb:
        box = disturbance_index(X, Y, Z, xr, yr, (-np.inf, np.inf))
        dist.append(Disturbance(kind, box, fortran_round(X[box[0] - 1:box[1]]), fortran_round(Y[box[2] - 1:box[3]]),
                                fortran_round(Z[box[4] - 1:box[5]]), [rnd(v) for v in par], var=var))
    case = Case("poiseuille_%dx%dx%d" % (nx, ny, nz), nx, ny, nz, X, Y, Z, rnd(Re), rnd(Ma), rnd(Pr), rnd(gamma),
                rnd(T0), deriv, filt, types, (1, 1, 1), -1.0, bi, dist, forc, dxmin=_dxmin(X, Y, Z, nz))
    case.flow_region = fr
    shape = (nx, ny, nz)
    st = {k: np.zeros(shape, order="F") for k in FIELDS}
    st["R"][:] = 1.0
    st["E"][:] = E0
    if profile:
        # mu = 1 at the reference temperature (T = 1 when E = E0)
        u = dpdx * Re / 2.0 * (H * H / 4.0 - Y ** 2)
        st["U"][:] = u[None, :, None]
    else:
        st["U"][:] = 1.0
        st["U"][:, [0, -1], :] = 0.0
    case.state0 = st
    case.dt = rnd(cfl / cfl_dt(st, case.Ma, case.dxmin, nz))
    return case


def cavity_sweep(count=16, nx_range=(256, 768), ny_range=(128, 256), re_range=(500.0, 2000.0), Ma=0.5):
    """Ragged parameter sweep of lid-driven cavities: grid sizes and Reynolds numbers spread linearly."""
    nxs = np.round(np.linspace(nx_range[0], nx_range[1], count)).astype(int)
    nys = np.round(np.linspace(ny_range[0], ny_range[1], count)).astype(int)
    res = np.linspace(re_range[0], re_range[1], count)
    return [lid_driven(nx=int(a), ny=int(b), Re=float(r), Ma=Ma) for a, b, r in zip(nxs, nys, res)]


def _buffer_points(x_edge, h0, n, growth, sign):
    """n buffer points beyond x_edge whose spacing rises from h0 to growth*h0 along a sigmoid."""
    k = np.arange(1, n + 1)
    ramp = 1.0 / (1.0 + np.exp(-(k - 0.5 * n) / (0.12 * n)))
    ramp = (ramp - ramp[0]) / (ramp[-1] - ramp[0])
    h = h0 * (1.0 + (growth - 1.0) * ramp)
    return x_edge + sign * np.cumsum(h)


def blasius(eta_max=20.0, n=4001):
    """Blasius solution of f''' + f f''/2 = 0, f(0) = f'(0) = 0, f'(inf) = 1 (RK4 shooting on f''(0))."""
    h = eta_max / (n - 1)
    rhs = lambda y: np.array([y[1], y[2], -0.5 * y[0] * y[2]])

    def integrate(s):
        y = np.array([0.0, 0.0, s])
        out = np.zeros((n, 3))
        out[0] = y
        for i in range(1, n):
            k1 = rhs(y)
            k2 = rhs(y + h / 2 * k1)
            k3 = rhs(y + h / 2 * k2)
            k4 = rhs(y + h * k3)
            y = y + h / 6 * (k1 + 2 * k2 + 2 * k3 + k4)
            out[i] = y
        return out

    lo, hi = 0.2, 0.5
    for _ in range(50):
        mid = 0.5 * (lo + hi)
        if integrate(mid)[-1, 1] > 1.0:
            hi = mid
        else:
            lo = mid
    return np.linspace(0.0, eta_max, n), integrate(0.5 * (lo + hi))


def boundary_layer(nx=1024, ny=256, nbuf=40, x_range=(-1.0, 30.0), y_top=8.0, wall_ratio=1.02, Re=1000.0,
                   Ma=0.5, Pr=0.71, gamma=1.4, T0=300.0, alpha_f=0.49, spatial="SL4", cfl=0.5, sfd_chi=0.02,
                   sfd_delta=10.0, noise=1e-5, seed=20261, buffer_growth=8.0, nz=1, z_length=4.0,
                   adiabatic=False, z_symmetry=False, pressure_inlet=False, free_slip=False,
                   initial="compressible", sfd_type=2, cavities=(), roughness=(), wall_depth=0.0, n_below=0):
    """Isothermal-wall boundary layer (boundaryLayerIsothermal.m; ``adiabatic=True``: the walls of
    boundaryLayerAdiabatic.m; ``free_slip=True``: the slip plate of boundaryLayerFreeSlip.m, with
    zero-gradient u and w on the wall and no x break point) on a stretched mesh with buffers.

    With nz > 1 the 2-D configuration is extruded (``z_symmetry``: between the symmetry planes of
    boundaryLayerIsothermalSymmZ.m on [0, z_length]; otherwise) over a spanwise-periodic uniform z of length
    z_length (the flow region, boundary boxes and initial profiles are the same in every plane).

    The mesh is built here (a 'file' mesh for the reference): x uniform on x_range with nbuf-point
    buffers at both ends, y geometric from the wall with ratio ``wall_ratio`` up to ``y_top`` and an
    nbuf-point buffer above; buffer spacings grow along a sigmoid to ``buffer_growth`` times the edge
    spacing.  Numerics as parameters.m: SL4 with EX4 in the buffers, 'reducedOrder' filter borders with
    the x-start, x-end and y-end closures off, SFD type 2 (chi in the x buffers only, applyY = false;
    ``sfd_type=1``: chi everywhere, main.F90:390, as a steady-base computation uses it).
    Initial flow (``initial``): 'compressible' (the default) is calcCompressibleBL.m's similarity
    solution for the isothermal (or, with ``adiabatic``, the adiabatic) wall, uniform upstream of
    the plate; 'blasius' is the incompressible Blasius velocity scaled so that the displacement
    thickness is 1 at x = 0 with uniform R and E.  Either gets seeded N(0, noise^2) on u and v in
    the flow region.

    Cavities and roughness elements (``flowType.cav`` / ``flowType.rug``: dicts with x, y, z
    intervals) are cut into the plate as the flow-type scripts do; a cavity below the plate needs
    mesh rows below y = 0: ``n_below`` uniform rows from y = -``wall_depth`` (``domain.yi < 0`` of
    the reference's cavity configuration, parameters.m:28-44), which come on top of the ``ny``
    rows from the wall up.
    """
    npx, npy = nx - 2 * nbuf, ny - nbuf
    Xp = np.linspace(x_range[0], x_range[1], npx)
    dx = Xp[1] - Xp[0]
    X = np.concatenate([_buffer_points(Xp[0], dx, nbuf, buffer_growth, -1.0)[::-1], Xp,
                        _buffer_points(Xp[-1], dx, nbuf, buffer_growth, 1.0)])
    dy0 = y_top * (wall_ratio - 1.0) / (wall_ratio ** (npy - 1) - 1.0)
    Yp = np.concatenate([[0.0], np.cumsum(dy0 * wall_ratio ** np.arange(npy - 1))])
    Yp[-1] = y_top
    Y = np.concatenate([Yp, _buffer_points(Yp[-1], Yp[-1] - Yp[-2], nbuf, buffer_growth, 1.0)])
    if n_below:
        Y = np.concatenate([-wall_depth + (wall_depth / n_below) * np.arange(n_below), Y])
        ny = len(Y)
    z_sym = z_symmetry and nz > 1
    if z_sym:      # symmetry planes at both ends of a uniform z on [0, z_length]
        Z = np.linspace(0.0, z_length, nz)
    else:
        Z = np.arange(nz) * (z_length / nz) if nz > 1 else np.array([0.0])
    E0 = 1.0 / ((gamma ** 2 - gamma) * Ma ** 2)
    P0 = (gamma - 1) * E0
    if free_slip:
        if adiabatic or z_sym or pressure_inlet:
            raise ValueError("free_slip takes none of adiabatic, z_symmetry, pressure_inlet")
        fr, walls, corners, bc, bp, hold = boundary_layer_free_slip(X, Y, Z, E0, P0, cavities=cavities,
                                                                    roughness=roughness)
    else:
        fr, walls, corners, bc, bp, hold = boundary_layer_isothermal(X, Y, Z, E0, P0, adiabatic=adiabatic,
                                                                     z_periodic=not z_sym, z_symmetry=z_sym,
                                                                     pressure_inlet=pressure_inlet,
                                                                     cavities=cavities, roughness=roughness)
    buf = BufferSide(n=nbuf, upwind=False, transition=0.2)
    meshes = [DirMesh(X, buffer_i=buf, buffer_f=buf, break_points=[bp] if bp else []),
              DirMesh(Y, buffer_f=buf),
              DirMesh(Z) if z_sym else
              DirMesh(Z, periodic=True, fix_periodic_domain_size=True, xf=z_length if nz > 1 else 1.0)]
    nm = NumMethods(spatial_derivs=spatial, spatial_derivs_buffer="EX4", filter_strength=alpha_f,
                    filter_borders="reducedOrder", filter_borders_start=(False, True, True),
                    filter_borders_end=(False, False, True))
    deriv, filt, types = _coeffs(meshes, fr, nm)
    bi = init_boundaries(bc, corners, nm.neumann_order, nm.neumann2_order)
    rnd = lambda v: float(fortran_round(v))
    chi = calc_sfd_region(nx, ny, nz, ((nbuf, nbuf), (0, nbuf), (0, 0)), sfd_chi, apply=(2, 0, 2))
    name = ("boundary_layer_free_slip_%dx%d" if free_slip else "boundary_layer_%dx%d") % (nx, ny) + \
        ("x%d" % nz if nz > 1 else "") + ("_cav%d" % len(cavities) if cavities else "") + \
        ("_rug%d" % len(roughness) if roughness else "")
    dist = []
    if hold:
        # a domain that starts past the leading edge holds its inflow plane at the initial
        # flow (boundaryLayerIsothermal.m:53-67: holdInlet on x = X(1), density not held)
        box = [1, 1, 1, ny, 1, nz]
        dist.append(Disturbance("holdInlet", box, fortran_round(X[:1]), fortran_round(Y), fortran_round(Z), [0.0]))
    case = Case(name, nx, ny, nz, X, Y, Z, rnd(Re), rnd(Ma), rnd(Pr), rnd(gamma),
                rnd(T0), deriv, filt, types, (1, 1, 1), -1.0, bi, dist, [],
                sfd_type=sfd_type, sfd_delta=rnd(sfd_delta), sfd_x_val=rnd(sfd_chi),
                sfd_x=np.asfortranarray(fortran_round(chi)) if sfd_type == 2 else None, dxmin=_dxmin(X, Y, Z, nz))
    case.flow_region = fr
    case.buffers = ((nbuf, nbuf), (0, nbuf), (0, 0))
    shape = (nx, ny, nz)
    if initial == "compressible":
        # calcCompressibleBL.m ('compressibleBL_isothermal' / '_adiabatic' of generateInitialFlow.m:110-126)
        fl = compressible_bl_flow(case.Re, case.Ma, case.Pr, case.T0, case.gamma, X, Y, Z, adiabatic=adiabatic)
        U, V = fl["U"][:, :, 0], fl["V"][:, :, 0]
        R0, Ei = fl["R"], fl["E"]
    else:
        # incompressible Blasius scaled so that the displacement thickness is 1 at x = 0
        eta, sol = blasius()
        x0 = Re / 1.7207876573 ** 2
        xe = np.maximum(X, 0.0)[:, None] + x0
        et = Y[None, :] * np.sqrt(Re / xe)
        fp = np.interp(et, eta, sol[:, 1], right=1.0)
        f = np.where(et <= eta[-1], np.interp(et, eta, sol[:, 0]), sol[-1, 0] + (et - eta[-1]))
        U = fp.copy()
        V = 0.5 * np.sqrt(1.0 / (Re * xe)) * (et * fp - f)
        upstream = X < 0
        U[upstream, :] = 1.0
        V[upstream, :] = 0.0
        R0, Ei = np.ones(shape), np.full(shape, E0)
    rng = np.random.Generator(np.random.PCG64(seed))
    U = U[:, :, None] + noise * rng.standard_normal(shape)
    V = V[:, :, None] + noise * rng.standard_normal(shape)
    st = {"U": np.asfortranarray(U), "V": np.asfortranarray(V),
          "W": np.zeros(shape, order="F"), "R": np.asfortranarray(R0),
          "E": np.asfortranarray(Ei)}
    wall = ~fr
    for k in ("U", "V"):
        st[k][wall] = 0.0
    case.state0 = st
    case.dt = rnd(cfl / cfl_dt(st, case.Ma, case.dxmin, nz))
    return case
