// The point-wise right-hand side and stage update shared by assemble_kernel
// (pointwise.hip) and the fused line kernel (rhs_line.hip), so both evaluate the
// same expressions in the same order.
#pragma once
#include "gatt_internal.h"

namespace gatt {

// nseq.F90:280-318 (3-D), 106-134 (2-D) at point n of the stage state S, with the
// turbulenceGenerator forcing (turbulenceGenerator.F90:57-61, global sums of member
// `member`) and pressureGradient (pressureGradient.F90:13) after it.  D holds the 15
// first derivatives (3f + d) and g the flux derivatives, g[4d + q]:
//   x: txx, txy, txz, qx | y: txy, tyy, tyz, qy | z: txz, tyz, tzz, qz
// (2-D: entries 2, 6 and 8-11 unused).  d receives dU, dV, dW, dR, dE and sv the stage
// state U, V, W, R, E at n.  The stresses are recomputed from the derivatives and E with
// the expressions of visc_kernel, so they equal the values that kernel stores.
__device__ __forceinline__ void rhs_point(const AssembleArgs& a, const double* S, const double* D, long long n,
                                          const double* g, int member, double* d, double* sv) {
  const long long np = a.np;
  const double U = ld_nt(S + 0 * np + n), V = ld_nt(S + 1 * np + n), R = ld_nt(S + 3 * np + n), E = ld_nt(S + 4 * np + n);
  const FlowConst c = a.c;
  const double DUx = ld_nt(D + 0 * np + n), DUy = ld_nt(D + 1 * np + n);
  const double DVx = ld_nt(D + 3 * np + n), DVy = ld_nt(D + 4 * np + n);
  const double DRx = ld_nt(D + 9 * np + n), DRy = ld_nt(D + 10 * np + n);
  const double DEx = ld_nt(D + 12 * np + n), DEy = ld_nt(D + 13 * np + n);
  const double P = c.gm1 * R * E;
  const double DPx = c.gm1 * (R * DEx + DRx * E);
  const double DPy = c.gm1 * (R * DEy + DRy * E);
  const double mu1 = c.muA / (c.muC * E + c.muB);
  const double mu2 = c.muC * E;
  const double mu = mu1 * mu2 * sqrt(mu2);
  const double rmu = c.invRe * mu;
  double W = 0.0;
  if (a.is3d) {
    W = ld_nt(S + 2 * np + n);
    const double DUz = ld_nt(D + 2 * np + n), DVz = ld_nt(D + 5 * np + n);
    const double DWx = ld_nt(D + 6 * np + n), DWy = ld_nt(D + 7 * np + n), DWz = ld_nt(D + 8 * np + n);
    const double DRz = ld_nt(D + 11 * np + n), DEz = ld_nt(D + 14 * np + n);
    const double DPz = c.gm1 * (R * DEz + DRz * E);
    const double tz = -(2.0 / 3.0 * (DUx + DVy + DWz));
    const double txx = rmu * (2.0 * DUx + tz), tyy = rmu * (2.0 * DVy + tz), tzz = rmu * (2.0 * DWz + tz);
    const double txy = rmu * (DUy + DVx), txz = rmu * (DUz + DWx), tyz = rmu * (DVz + DWy);
    const double dtxxdx = g[0], dtxydx = g[1], dtxzdx = g[2], htx = g[3];
    const double dtxydy = g[4], dtyydy = g[5], dtyzdy = g[6], hty = g[7];
    const double dtxzdz = g[8], dtyzdz = g[9], dtzzdz = g[10], htz = g[11];
    const double vtx = dtxxdx + dtxydy + dtxzdz;
    const double vty = dtyydy + dtxydx + dtyzdz;
    const double vtz = dtzzdz + dtxzdx + dtyzdy;
    const double div = DUx + DVy + DWz;
    d[3] = -(R * div) - DRx * U - DRy * V - DRz * W;
    d[0] = -(DUx * U + DUy * V + DUz * W) + (vtx - DPx) / R;
    d[1] = -(DVx * U + DVy * V + DVz * W) + (vty - DPy) / R;
    d[2] = -(DWx * U + DWy * V + DWz * W) + (vtz - DPz) / R;
    d[4] = -(DEx * U + DEy * V + DEz * W) +
           (-(P * div) + txx * DUx + tyy * DVy + tzz * DWz + txy * (DVx + DUy) + txz * (DWx + DUz) +
            tyz * (DWy + DVz) - htx - hty - htz) / R;
    if (a.sums) {
      // turbulenceGenerator.F90:57-61 with the global sums of this member
      const double* sm = a.sums + 8 * member;
      const double ET = sm[0], Us = sm[1], Vs = sm[2], Ws = sm[3], Es = sm[4], Nd = sm[5];
      const double gam = 10 * a.forcing_eps;
      const double fk = 0.25 * a.forcing_eps * Nd / ET;
      d[0] = d[0] + fk * U - gam * Us / Nd;
      d[1] = d[1] + fk * V - gam * Vs / Nd;
      d[2] = d[2] + fk * W - gam * Ws / Nd;
      d[4] = d[4] - gam * (Es / Nd - a.forcing_e0);
    }
  } else {
    const double dv = 2.0 / 3.0 * (DUx + DVy);
    const double txx = rmu * (2.0 * DUx - dv), tyy = rmu * (2.0 * DVy - dv), txy = rmu * (DUy + DVx);
    const double dtxxdx = g[0], dtxydx = g[1], htx = g[3];
    const double dtxydy = g[4], dtyydy = g[5], hty = g[7];
    const double vtx = dtxxdx + dtxydy;
    const double vty = dtyydy + dtxydx;
    const double div = DUx + DVy;
    d[3] = -(R * div) - DRx * U - DRy * V;
    d[0] = -(DUx * U + DUy * V) + (vtx - DPx) / R;
    d[1] = -(DVx * U + DVy * V) + (vty - DPy) / R;
    d[2] = 0.0;
    d[4] = -(DEx * U + DEy * V) + (-(P * div) + txx * DUx + tyy * DVy + txy * (DVx + DUy) - htx - hty) / R;
  }
  if (a.pg) {
    // pressureGradient.F90:13 (du = du + dpdx/r) on its box, after the other forcings
    const int i = (int)(n % a.nx);
    const int j = (int)((n / a.nx) % a.ny);
    const int k = (int)(n / ((long long)a.nx * a.ny));
    if (i >= a.pg_box[0] && i <= a.pg_box[1] && j >= a.pg_box[2] && j <= a.pg_box[3] && k >= a.pg_box[4] &&
        k <= a.pg_box[5]) {
      const double r = ld_nt(S + (long long)a.pg_r * np + n);
#pragma unroll
      for (int f = 0; f < 5; ++f)
        if (f == a.pg_u) d[f] = d[f] + a.pg_dpdx / r;
    }
  }
  sv[0] = U;
  sv[1] = V;
  sv[2] = W;
  sv[3] = R;
  sv[4] = E;
}

// One stage update at point n of member offset mo (state arrays Q, Out, Delta of AssembleArgs):
// stage 0 stores the right-hand side, 1-4 are the RK4 stages (timeIntegration.F90:393-487),
// 5 the Euler update (553-557) and 6-8 the SSPRK3 stages (623-627, 647-651, 672-676).  The
// first RK4 stage updates the state it differentiates (Q == S): its values come from sv.  The
// third leaves q + Delta in the accumulator, which the fourth reads in place of q and Delta
// (the sum (q + Delta) + dt/6 k4 of timeIntegration.F90:482-487 is unchanged).
// ov receives the values written to Out.
__device__ __forceinline__ void stage_point(const AssembleArgs& a, long long mo, long long n, const double* d,
                                            const double* sv, double* ov) {
  const long long np = a.np;
  const double* Q = a.Q + mo;
  double* O = a.Out + mo;
  double* Dl = a.Delta + mo;
  const double dt = a.dt;
  const bool q_is_s = a.Q == a.S;
#pragma unroll
  for (int f = 0; f < 5; ++f) {
    ov[f] = 0.0;
    if (!a.is3d && f == 2) continue;
    const long long o = f * np + n;
    double v;
    if (q_is_s && a.stage == 1) {
      st_nt(Dl + o, dt / 6.0 * d[f]);
      v = sv[f] + dt / 2.0 * d[f];
    } else {
      switch (a.stage) {
        case 0: v = d[f]; break;
        case 1: st_nt(Dl + o, dt / 6.0 * d[f]); v = ld_nt(Q + o) + dt / 2.0 * d[f]; break;
        case 2: st_nt(Dl + o, ld_nt(Dl + o) + dt / 3.0 * d[f]); v = ld_nt(Q + o) + dt / 2.0 * d[f]; break;
        case 3: {
          // the last stage's q + Delta, stored in the accumulator: stage 4 reads one array
          // where it would read two, and adds the same two values it would add
          const double q = ld_nt(Q + o);
          st_nt(Dl + o, q + (ld_nt(Dl + o) + dt / 3.0 * d[f]));
          v = q + dt * d[f];
          break;
        }
        case 4: v = ld_nt(Dl + o) + dt / 6.0 * d[f]; break;
        case 5: v = ld_nt(Q + o) + dt * d[f]; break;
        case 6: st_nt(Dl + o, d[f]); v = ld_nt(Q + o) + dt * d[f]; break;
        case 7: {
          const double k = ld_nt(Dl + o) + d[f];
          st_nt(Dl + o, k);
          v = ld_nt(Q + o) + dt * k / 4.0;
          break;
        }
        default: v = ld_nt(Q + o) + dt * (ld_nt(Dl + o) + 4.0 * d[f]) / 6.0; break;
      }
    }
    st_nt(O + o, v);
    ov[f] = v;
  }
}

}  // namespace gatt
