// Boundary conditions (applyBoundCond, source/Fortran/nseq.F90:327-599) as a
// compiled program of box operations.
//
// The host lowers the reference's ordered lists (Dirichlet U,V,W,E,P; Neumann
// U,V,W,E,P; corners; null second derivative U,V,W,E,P; disturbances) into
// BcOps and groups consecutive ops whose read and write sets do not intersect
// into one "wave".  A wave is one launch: every thread owns one point of one
// op, so a wave is equivalent to running its ops one after another.
#include "gatt_internal.h"

namespace gatt {

__device__ __forceinline__ long long pidx(int i, int j, int k, int nx, int ny) {
  return (long long)i + (long long)nx * ((long long)j + (long long)ny * k);
}

// One-sided extrapolation of applyNeumann (nseq.F90:561-597) at point (i,j,k)
// along direction dir (1..6), from a field given by an accessor.
template <typename F>
__device__ __forceinline__ double extrap(F&& val, int i, int j, int k, int dir, int L, const double* c) {
  int di = 0, dj = 0, dk = 0;
  switch (dir) {
    case 1: di = 1; break;
    case 2: di = -1; break;
    case 3: dj = 1; break;
    case 4: dj = -1; break;
    case 5: dk = 1; break;
    default: dk = -1; break;
  }
  double v = c[0] * val(i + di, j + dj, k + dk);
  for (int q = 2; q <= L; ++q) v = v + c[q - 1] * val(i + q * di, j + q * dj, k + q * dk);
  return v;
}

__device__ __forceinline__ double powi18(double x) {
  const double x2 = x * x;
  const double x4 = x2 * x2;
  const double x8 = x4 * x4;
  const double x16 = x8 * x8;
  return x16 * x2;
}

__device__ __forceinline__ void bc_wave_body(const BcArgs& a, BlockPos bp) {
  // find this block's op
  const long long blk = bp.x;
  int o = a.op_begin;
  while (o + 1 < a.op_end && a.ops[o + 1].first_block <= blk) ++o;
  const BcOp& op = a.ops[o];
  const long long p = (blk - op.first_block) * blockDim.x + threadIdx.x;
  if (p >= op.npts) return;
  const int bi = op.box[1] - op.box[0] + 1;
  const int bj = op.box[3] - op.box[2] + 1;
  const int i = op.box[0] + (int)(p % bi);
  const int j = op.box[2] + (int)((p / bi) % bj);
  const int k = op.box[4] + (int)(p / ((long long)bi * bj));
  const int nx = a.nx, ny = a.ny;
  const long long mo = (long long)bp.y * a.member_stride;
  double* U = a.q[0] + mo;
  double* R = a.q[3] + mo;
  double* E = a.q[4] + mo;
  const long long n = pidx(i, j, k, nx, ny);
  if (a.undo) {
    // a point is written by one op of a wave only, and read by no other op of it
    double* u = a.undo + (long long)bp.y * a.undo_member_stride + op.undo;
#pragma unroll
    for (int f = 0; f < 5; ++f) u[(long long)f * op.npts + p] = (a.q[f] + mo)[n];
  }
  switch (op.kind) {
    case BC_SET: {
      (a.q[op.field] + mo)[n] = op.value;
      break;
    }
    case BC_SET_R_FROM_P: {
      R[n] = op.value / (a.gamma - 1.0) / E[n];
      break;
    }
    case BC_NEU: {
      double* F = a.q[op.field] + mo;
      auto val = [&](int ii, int jj, int kk) { return F[pidx(ii, jj, kk, nx, ny)]; };
      F[n] = extrap(val, i, j, k, op.dir, op.L, a.coef[op.coef]);
      break;
    }
    case BC_NEU_P: {
      auto val = [&](int ii, int jj, int kk) {
        const long long m = pidx(ii, jj, kk, nx, ny);
        return R[m] * E[m];
      };
      const double Pb = extrap(val, i, j, k, op.dir, op.L, a.coef[op.coef]);
      R[n] = Pb / E[n];
      break;
    }
    case BC_CORNER: {
      const double nsum = (double)(abs(op.cd[0]) + abs(op.cd[1]) + abs(op.cd[2]));
      const double* c = a.coef[0];
      if (op.adiabatic) {
        auto valE = [&](int ii, int jj, int kk) { return E[pidx(ii, jj, kk, nx, ny)]; };
        double acc = 0.0;
        for (int ax = 0; ax < 3; ++ax) {
          if (op.cd[ax] == 0) continue;
          const int d = op.cd[ax] == 1 ? 2 * ax + 1 : 2 * ax + 2;
          acc = acc + extrap(valE, i, j, k, d, op.L, c);
        }
        E[n] = acc / nsum;
      }
      auto valP = [&](int ii, int jj, int kk) {
        const long long m = pidx(ii, jj, kk, nx, ny);
        return R[m] * E[m];
      };
      double acc = 0.0;
      for (int ax = 0; ax < 3; ++ax) {
        if (op.cd[ax] == 0) continue;
        const int d = op.cd[ax] == 1 ? 2 * ax + 1 : 2 * ax + 2;
        acc = acc + extrap(valP, i, j, k, d, op.L, c);
      }
      R[n] = acc / E[n] / nsum;
      break;
    }
    case BC_LID: {
      // lidDrivenMovement.F90:15-17
      const double* x = a.aux + op.aux;
      const int nxb = bi;
      const double s = 2 * (x[i - op.box[0]] - x[0]) / (x[nxb - 1] - x[0]) - 1;
      const double v = 1 - powi18(s);
      U[n] = op.value * (v * v);
      break;
    }
    case BC_HOLD: {
      // holdInlet.F90:18-48: capture on the first call, restore afterwards
      double* save = a.hold + (long long)bp.y * a.hold_member_stride + op.aux;
      const int vars[5] = {0, 1, 2, 4, 3};
      const int nv = op.hold_r ? 5 : 4;
      for (int v = 0; v < nv; ++v) {
        double* F = a.q[vars[v]] + mo;
        if (a.capture) save[(long long)v * op.npts + p] = F[n];
        else F[n] = save[(long long)v * op.npts + p];
      }
      break;
    }
    case BC_SINE: {
      // periodicSine.F90:19-21; aux = x(1:nx), alpha (single-precision value), omega
      const double t = *a.clock + a.toff;
      const double* x = a.aux + op.aux;
      const double alpha = x[bi], omega = x[bi + 1];
      double* F = a.q[op.field] + mo;
      F[n] = F[n] + op.value * sin(alpha * (x[i - op.box[0]] - x[0])) * sin(omega * t);
      break;
    }
    case BC_PACKET: {
      // packet_2d.F90:27-39; aux = x(1:nx), x0, Tend, Tp, sigmaT, sigmaX (single-precision
      // values), omega, nmodes; op.L = int(nmodes)
      const double t = *a.clock + a.toff;
      const double* x = a.aux + op.aux;
      const double* k = x + bi;
      const double x0 = k[0], Tend = k[1], Tp = k[2], sT = k[3], sX = k[4], omega = k[5], nmodes = k[6];
      if (t > Tend) break;
      double* F = a.q[op.field] + mo;
      const double dT = t - Tp;
      const double dX = x[i - op.box[0]] - x0;
      double acc = F[n];
      for (int m = 1; m <= op.L; ++m) {
        const double th = (double)m * omega * dT;
        const float re = (float)cos(th);   // aux is a single-precision complex
        acc = acc + exp(-(dT * dT) / (2.0 * (sT * sT))) * exp(-(dX * dX) / (2.0 * (sX * sX))) * (double)re / nmodes;
      }
      F[n] = op.value * acc;
      break;
    }
  }
}

__global__ __launch_bounds__(256) void bc_wave_kernel(BcArgs a) { bc_wave_body(a, grid_pos()); }
__global__ __launch_bounds__(256) void bc_wave_ragged(const BcArgs* __restrict__ args, RaggedMap map) {
  BlockPos bp;
  const int m = ragged_pos(map, bp);
  bc_wave_body(args[m], bp);
}

__global__ void clock_set_kernel(double* clock, double t) {
  if (threadIdx.x == 0) clock[0] = t;
}

__global__ void clock_add_kernel(double* clock, double dt) {
  if (threadIdx.x == 0) clock[0] = clock[0] + dt;
}
__global__ void clock_add_ragged(const ClockArgs* __restrict__ args, RaggedMap map) {
  BlockPos bp;
  const ClockArgs& a = args[ragged_pos(map, bp)];
  if (threadIdx.x == 0) a.clock[0] = a.clock[0] + a.dt;
}

void launch_clock_set(double* clock, double t, hipStream_t st) {
  hipLaunchKernelGGL(clock_set_kernel, dim3(1), dim3(64), 0, st, clock, t);
}

void launch_clock_add(double* clock, double dt, hipStream_t st) {
  if (record_launch((const void*)clock_add_ragged, ClockArgs{clock, dt}, dim3(1), dim3(64), 0)) return;
  hipLaunchKernelGGL(clock_add_kernel, dim3(1), dim3(64), 0, st, clock, dt);
}

void launch_bc_wave(const BcArgs& a, long long nblocks, int nmembers, hipStream_t st) {
  if (nblocks <= 0) return;
  const dim3 grid((unsigned)nblocks, nmembers);
  if (record_launch((const void*)bc_wave_ragged, a, grid, dim3(256), 0)) return;
  hipLaunchKernelGGL(bc_wave_kernel, grid, dim3(256), 0, st, a);
}

}  // namespace gatt

namespace gatt {

// ------------------------------------------------------------------------
// Tangent and adjoint of the boundary program.  For one wave, bc_wave_tl_kernel
// applies the tangent of every op to dq (using the base state before the op) and
// then the op itself to the base; bc_wave_adj_kernel applies the transpose of the
// wave's tangent to an adjoint lam, the base being the state before the wave.
// Ops of one wave neither write what another reads nor write the same point, so
// the transposed stencils may only meet on read points: they accumulate with
// atomics.
// ------------------------------------------------------------------------

struct OpPoint {
  int i, j, k;
  long long n, p;
};

__device__ __forceinline__ bool op_point(const BcLinArgs& a, const BcOp*& op_out, OpPoint& pt) {
  const long long blk = blockIdx.x;
  int o = a.op_begin;
  while (o + 1 < a.op_end && a.ops[o + 1].first_block <= blk) ++o;
  const BcOp& op = a.ops[o];
  op_out = &op;
  pt.p = (blk - op.first_block) * blockDim.x + threadIdx.x;
  if (pt.p >= op.npts) return false;
  const int bi = op.box[1] - op.box[0] + 1;
  const int bj = op.box[3] - op.box[2] + 1;
  pt.i = op.box[0] + (int)(pt.p % bi);
  pt.j = op.box[2] + (int)((pt.p / bi) % bj);
  pt.k = op.box[4] + (int)(pt.p / ((long long)bi * bj));
  pt.n = pidx(pt.i, pt.j, pt.k, a.nx, a.ny);
  return true;
}

__device__ __forceinline__ void dir_step(int dir, int& di, int& dj, int& dk) {
  di = dj = dk = 0;
  switch (dir) {
    case 1: di = 1; break;
    case 2: di = -1; break;
    case 3: dj = 1; break;
    case 4: dj = -1; break;
    case 5: dk = 1; break;
    default: dk = -1; break;
  }
}

__global__ __launch_bounds__(256) void bc_wave_tl_kernel(BcLinArgs a) {
  const BcOp* opp;
  OpPoint pt;
  if (!op_point(a, opp, pt)) return;
  const BcOp& op = *opp;
  const int nx = a.nx, ny = a.ny;
  const long long mo = (long long)blockIdx.y * a.member_stride, lo = (long long)blockIdx.y * a.lin_stride;
  double* R = a.q[3] + mo;
  double* E = a.q[4] + mo;
  double* tR = a.dq[3] + lo;
  double* tE = a.dq[4] + lo;
  const long long n = pt.n;
  switch (op.kind) {
    case BC_SET:
      (a.dq[op.field] + lo)[n] = 0.0;
      (a.q[op.field] + mo)[n] = op.value;
      break;
    case BC_SET_R_FROM_P: {
      const double Rn = op.value / (a.gamma - 1.0) / E[n];
      tR[n] = -(Rn / E[n] * tE[n]);
      R[n] = Rn;
      break;
    }
    case BC_NEU: {
      double* F = a.q[op.field] + mo;
      double* tF = a.dq[op.field] + lo;
      auto val = [&](int ii, int jj, int kk) { return F[pidx(ii, jj, kk, nx, ny)]; };
      auto tval = [&](int ii, int jj, int kk) { return tF[pidx(ii, jj, kk, nx, ny)]; };
      tF[n] = extrap(tval, pt.i, pt.j, pt.k, op.dir, op.L, a.coef[op.coef]);
      F[n] = extrap(val, pt.i, pt.j, pt.k, op.dir, op.L, a.coef[op.coef]);
      break;
    }
    case BC_NEU_P: {
      auto valP = [&](int ii, int jj, int kk) {
        const long long m = pidx(ii, jj, kk, nx, ny);
        return R[m] * E[m];
      };
      auto tvalP = [&](int ii, int jj, int kk) {
        const long long m = pidx(ii, jj, kk, nx, ny);
        return tR[m] * E[m] + R[m] * tE[m];
      };
      const double Pb = extrap(valP, pt.i, pt.j, pt.k, op.dir, op.L, a.coef[op.coef]);
      const double tPb = extrap(tvalP, pt.i, pt.j, pt.k, op.dir, op.L, a.coef[op.coef]);
      const double Rn = Pb / E[n];
      tR[n] = (tPb - Rn * tE[n]) / E[n];
      R[n] = Rn;
      break;
    }
    case BC_CORNER: {
      const double nsum = (double)(abs(op.cd[0]) + abs(op.cd[1]) + abs(op.cd[2]));
      const double* c = a.coef[0];
      if (op.adiabatic) {
        auto valE = [&](int ii, int jj, int kk) { return E[pidx(ii, jj, kk, nx, ny)]; };
        auto tvalE = [&](int ii, int jj, int kk) { return tE[pidx(ii, jj, kk, nx, ny)]; };
        double acc = 0.0, tacc = 0.0;
        for (int ax = 0; ax < 3; ++ax) {
          if (op.cd[ax] == 0) continue;
          const int d = op.cd[ax] == 1 ? 2 * ax + 1 : 2 * ax + 2;
          acc = acc + extrap(valE, pt.i, pt.j, pt.k, d, op.L, c);
          tacc = tacc + extrap(tvalE, pt.i, pt.j, pt.k, d, op.L, c);
        }
        tE[n] = tacc / nsum;
        E[n] = acc / nsum;
      }
      auto valP = [&](int ii, int jj, int kk) {
        const long long m = pidx(ii, jj, kk, nx, ny);
        return R[m] * E[m];
      };
      auto tvalP = [&](int ii, int jj, int kk) {
        const long long m = pidx(ii, jj, kk, nx, ny);
        return tR[m] * E[m] + R[m] * tE[m];
      };
      double acc = 0.0, tacc = 0.0;
      for (int ax = 0; ax < 3; ++ax) {
        if (op.cd[ax] == 0) continue;
        const int d = op.cd[ax] == 1 ? 2 * ax + 1 : 2 * ax + 2;
        acc = acc + extrap(valP, pt.i, pt.j, pt.k, d, op.L, c);
        tacc = tacc + extrap(tvalP, pt.i, pt.j, pt.k, d, op.L, c);
      }
      const double Rn = acc / E[n] / nsum;
      tR[n] = (tacc / nsum - Rn * tE[n]) / E[n];
      R[n] = Rn;
      break;
    }
    case BC_LID: {
      const double* x = a.aux + op.aux;
      const int nxb = op.box[1] - op.box[0] + 1;
      const double s = 2 * (x[pt.i - op.box[0]] - x[0]) / (x[nxb - 1] - x[0]) - 1;
      const double v = 1 - powi18(s);
      (a.dq[0] + lo)[n] = 0.0;
      (a.q[0] + mo)[n] = op.value * (v * v);
      break;
    }
    case BC_HOLD: {
      const double* save = a.hold + (long long)blockIdx.y * a.hold_member_stride + op.aux;
      const int vars[5] = {0, 1, 2, 4, 3};
      const int nv = op.hold_r ? 5 : 4;
      for (int v = 0; v < nv; ++v) {
        (a.dq[vars[v]] + lo)[n] = 0.0;
        (a.q[vars[v]] + mo)[n] = save[(long long)v * op.npts + pt.p];
      }
      break;
    }
  }
}

// lam[m] += c_k * w at the stencil points of an extrapolation from (i,j,k) along dir
template <typename F>
__device__ __forceinline__ void extrap_t(F&& add, int i, int j, int k, int dir, int L, const double* c) {
  int di, dj, dk;
  dir_step(dir, di, dj, dk);
  for (int q = 1; q <= L; ++q) add(i + q * di, j + q * dj, k + q * dk, c[q - 1]);
}

__global__ __launch_bounds__(256) void bc_wave_adj_kernel(BcLinArgs a) {
  const BcOp* opp;
  OpPoint pt;
  if (!op_point(a, opp, pt)) return;
  const BcOp& op = *opp;
  const int nx = a.nx, ny = a.ny;
  const long long mo = (long long)blockIdx.y * a.member_stride, lo = (long long)blockIdx.y * a.lin_stride;
  if (a.undo) {
    // the wave wrote only its own points, each from one op: putting this point back
    // makes the base the state before the wave for every read of this thread and of
    // the other ops (which read no point the wave writes)
    const double* u = a.undo + (long long)blockIdx.y * a.undo_member_stride + op.undo;
#pragma unroll
    for (int f = 0; f < 5; ++f) (a.q[f] + mo)[pt.n] = u[(long long)f * op.npts + pt.p];
  }
  const double* R = a.q[3] + mo;
  const double* E = a.q[4] + mo;
  double* bR = a.dq[3] + lo;
  double* bE = a.dq[4] + lo;
  const long long n = pt.n;
  switch (op.kind) {
    case BC_SET:
      (a.dq[op.field] + lo)[n] = 0.0;
      break;
    case BC_SET_R_FROM_P: {
      const double Rn = op.value / (a.gamma - 1.0) / E[n];
      const double t = bR[n];
      bR[n] = 0.0;
      bE[n] += -(Rn / E[n] * t);
      break;
    }
    case BC_NEU: {
      double* bF = a.dq[op.field] + lo;
      const double t = bF[n];
      bF[n] = 0.0;
      extrap_t([&](int ii, int jj, int kk, double ck) { atomicAdd(&bF[pidx(ii, jj, kk, nx, ny)], ck * t); },
               pt.i, pt.j, pt.k, op.dir, op.L, a.coef[op.coef]);
      break;
    }
    case BC_NEU_P: {
      auto valP = [&](int ii, int jj, int kk) {
        const long long m = pidx(ii, jj, kk, nx, ny);
        return R[m] * E[m];
      };
      const double Pb = extrap(valP, pt.i, pt.j, pt.k, op.dir, op.L, a.coef[op.coef]);
      const double Rn = Pb / E[n];
      const double t = bR[n];
      bR[n] = 0.0;
      bE[n] += -(Rn / E[n] * t);
      const double Pbar = t / E[n];
      extrap_t([&](int ii, int jj, int kk, double ck) {
                 const long long m = pidx(ii, jj, kk, nx, ny);
                 atomicAdd(&bR[m], ck * E[m] * Pbar);
                 atomicAdd(&bE[m], ck * R[m] * Pbar);
               },
               pt.i, pt.j, pt.k, op.dir, op.L, a.coef[op.coef]);
      break;
    }
    case BC_CORNER: {
      const double nsum = (double)(abs(op.cd[0]) + abs(op.cd[1]) + abs(op.cd[2]));
      const double* c = a.coef[0];
      // base values after the E part of the op
      double En = E[n];
      if (op.adiabatic) {
        auto valE = [&](int ii, int jj, int kk) { return E[pidx(ii, jj, kk, nx, ny)]; };
        double acc = 0.0;
        for (int ax = 0; ax < 3; ++ax) {
          if (op.cd[ax] == 0) continue;
          const int d = op.cd[ax] == 1 ? 2 * ax + 1 : 2 * ax + 2;
          acc = acc + extrap(valE, pt.i, pt.j, pt.k, d, op.L, c);
        }
        En = acc / nsum;
      }
      auto valP = [&](int ii, int jj, int kk) {
        const long long m = pidx(ii, jj, kk, nx, ny);
        return R[m] * E[m];
      };
      double acc = 0.0;
      for (int ax = 0; ax < 3; ++ax) {
        if (op.cd[ax] == 0) continue;
        const int d = op.cd[ax] == 1 ? 2 * ax + 1 : 2 * ax + 2;
        acc = acc + extrap(valP, pt.i, pt.j, pt.k, d, op.L, c);
      }
      const double Rn = acc / En / nsum;
      // transpose of R(b) = acc / E(b) / nsum
      const double t = bR[n];
      bR[n] = 0.0;
      double bEn = bE[n] - Rn / En * t;
      const double bacc = t / (En * nsum);
      for (int ax = 0; ax < 3; ++ax) {
        if (op.cd[ax] == 0) continue;
        const int d = op.cd[ax] == 1 ? 2 * ax + 1 : 2 * ax + 2;
        extrap_t([&](int ii, int jj, int kk, double ck) {
                   const long long m = pidx(ii, jj, kk, nx, ny);
                   atomicAdd(&bR[m], ck * E[m] * bacc);
                   atomicAdd(&bE[m], ck * R[m] * bacc);
                 },
                 pt.i, pt.j, pt.k, d, op.L, c);
      }
      if (op.adiabatic) {
        // transpose of E(b) = (sum of extrapolations of E) / nsum: E(b) itself is overwritten
        const double w = bEn / nsum;
        bEn = 0.0;
        for (int ax = 0; ax < 3; ++ax) {
          if (op.cd[ax] == 0) continue;
          const int d = op.cd[ax] == 1 ? 2 * ax + 1 : 2 * ax + 2;
          extrap_t([&](int ii, int jj, int kk, double ck) { atomicAdd(&bE[pidx(ii, jj, kk, nx, ny)], ck * w); },
                   pt.i, pt.j, pt.k, d, op.L, c);
        }
      }
      bE[n] = bEn;
      break;
    }
    case BC_LID:
      (a.dq[0] + lo)[n] = 0.0;
      break;
    case BC_HOLD: {
      const int vars[5] = {0, 1, 2, 4, 3};
      const int nv = op.hold_r ? 5 : 4;
      for (int v = 0; v < nv; ++v) (a.dq[vars[v]] + lo)[n] = 0.0;
      break;
    }
  }
}

void launch_bc_wave_tl(const BcLinArgs& a, long long nblocks, int nmembers, hipStream_t st) {
  if (nblocks <= 0) return;
  hipLaunchKernelGGL(bc_wave_tl_kernel, dim3((unsigned)nblocks, nmembers), dim3(256), 0, st, a);
}

void launch_bc_wave_adj(const BcLinArgs& a, long long nblocks, int nmembers, hipStream_t st) {
  if (nblocks <= 0) return;
  hipLaunchKernelGGL(bc_wave_adj_kernel, dim3((unsigned)nblocks, nmembers), dim3(256), 0, st, a);
}

}  // namespace gatt
