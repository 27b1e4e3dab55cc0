// Tangent-linear and adjoint kernels of the time step (no reference
// counterpart: differentiation of nseq.F90, derivs.F90 and the RK4/filter/SFD
// sequence as implemented in pointwise.hip and the line solves).
//
// Point-wise parts: the tangent of the viscous stresses / heat fluxes and of
// the RHS assembly, and their transposes, derived from the forward formulas of
// visc_kernel and assemble_kernel (nseq.F90:218-318, 2-D 72-134).  The linear
// parts (line solves, RK combinations, filter, SFD) reuse the forward kernels
// for the tangent; the adjoint uses the transposed solves (line_solve_t.hip)
// and the transposed RHS band below.
#include <algorithm>

#include "gatt_internal.h"

namespace gatt {

static inline unsigned grid_for(long long n, int threads) { return (unsigned)((n + threads - 1) / threads); }

// -------------------------------------------------------------- band^T
// out = beta*src + (accumulate ? out : 0) + alpha * sum_l R((e-l) mod N, l) in(e-l)
// Grid: x over i in blocks of 256 threads, y over j, z over (k, field, member),
// so every index comes from the grid without integer division per point.
template <int DIR>
__global__ __launch_bounds__(256) void band_t_kernel(BandTArgs a, int m0) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= a.nx) return;
  const int j = blockIdx.y;
  const int k = (int)blockIdx.z % a.nz;
  const int fm = (int)blockIdx.z / a.nz;
  const int f = fm % a.nf;
  const int mem = m0 + fm / a.nf;
  const long long nxy = (long long)a.nx * a.ny;
  const long long p = (long long)i + (long long)a.nx * j + nxy * k;
  const int e = DIR == 0 ? i : (DIR == 1 ? j : k);
  const int line = DIR == 0 ? j + a.ny * k : (DIR == 1 ? i + a.nx * k : i + a.nx * j);
  const long long stride = DIR == 0 ? 1 : (DIR == 1 ? (long long)a.nx : nxy);
  const int N = a.N, H = a.nR - 1, NW = 2 * a.nR - 1;
  // the reference's band layout R(point, diagonal, type): along x consecutive
  // threads read consecutive words of one diagonal
  const double* Rt = a.R + (size_t)a.types[line] * N * NW;
  const double* in = a.in[f] + (long long)mem * a.in_stride;
  double acc = 0.0;
  for (int l = -H; l <= H; ++l) {
    int se = e - l;
    se = se < 0 ? se + N : (se >= N ? se - N : se);
    acc = acc + Rt[(size_t)(H + l) * N + se] * in[p + (long long)(se - e) * stride];
  }
  double* out = a.out[f] + (long long)mem * a.out_stride;
  double v = a.alpha * acc;
  if (a.src[f]) v = a.beta * ld_nt(a.src[f] + (long long)mem * a.out_stride + p) + v;
  if ((a.accumulate >> f) & 1) v = ld_nt(out + p) + v;
  st_nt(out + p, v);
}

// y and z: lane = line (64 consecutive lines, every point a 512-byte row) and
// wavefront w = the M points [w*M, w*M+M) of the lines; each lane holds the
// window of its segment plus H points either side in registers, so every input
// row is read once per wavefront.  Same sums, in the same order, as
// band_t_kernel.
// UNI: the 64 lines of a block share one type, so the coefficient addresses are
// wave-uniform and go through the scalar cache.
template <int DIR, int NR, int M, bool UNI>
__global__ __launch_bounds__(1024) void band_t_seg_kernel(BandTArgs a) {
  constexpr int H = NR - 1;
  constexpr int W = M + 2 * H;
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const int N = a.N;
  const int nlines = DIR == 1 ? a.nx * a.nz : a.nx * a.ny;
  const int line0 = blockIdx.x * 64;
  const int nact = min(64, nlines - line0);
  const int line = lane < nact ? line0 + lane : line0;
  const int f = (int)blockIdx.y % a.nf;
  const int mem = (int)blockIdx.y / a.nf;
  const long long nxy = (long long)a.nx * a.ny;
  long long base;
  if (DIR == 1) {
    const int k = line / a.nx;
    base = (long long)(line - k * a.nx) + nxy * k;
  } else {
    base = line;
  }
  const long long stride = DIR == 1 ? (long long)a.nx : nxy;
  const int s = w * M;
  int cnt = N - s;
  cnt = cnt < 0 ? 0 : (cnt > M ? M : cnt);
  const double* in = a.in[f] + (long long)mem * a.in_stride;
  double fw[W];
#pragma unroll
  for (int u = 0; u < W; ++u) {
    int e = s - H + u;
    e = e < 0 ? e + N : (e >= N ? e - N : e);
    e = e >= N ? N - 1 : e;
    fw[u] = in[base + (long long)e * stride];
  }
  const int ty = UNI ? __builtin_amdgcn_readfirstlane(a.types[line0]) : a.types[line];
  const double* Rt = a.R + (size_t)ty * N * (2 * NR - 1);
  double* out = a.out[f] + (long long)mem * a.out_stride;
  const double* src = a.src[f] ? a.src[f] + (long long)mem * a.out_stride : nullptr;
  if (lane >= nact) return;
#pragma unroll
  for (int tt = 0; tt < M; ++tt) {
    if (tt < cnt) {
      const int e = s + tt;
      double acc = 0.0;
#pragma unroll
      for (int l = -H; l <= H; ++l) {
        int se = e - l;
        se = se < 0 ? se + N : (se >= N ? se - N : se);
        acc = acc + Rt[(size_t)(H + l) * N + se] * fw[tt + H - l];
      }
      const long long p = base + (long long)e * stride;
      double v = a.alpha * acc;
      if (src) v = a.beta * ld_nt(src + p) + v;
      if ((a.accumulate >> f) & 1) v = ld_nt(out + p) + v;
      st_nt(out + p, v);
    }
  }
}

template <int DIR, int M>
static bool launch_band_t_seg_m(const BandTArgs& a, int nmembers, hipStream_t st) {
  const int nlines = DIR == 1 ? a.nx * a.nz : a.nx * a.ny;
  const int G = (a.N + M - 1) / M;
  dim3 grid((unsigned)((nlines + 63) / 64), (unsigned)(a.nf * nmembers));
  switch (a.nR) {
#define GATT_BAND_T_NR(v)                                                                   \
  case v:                                                                                   \
    if (a.uniform) hipLaunchKernelGGL((band_t_seg_kernel<DIR, v, M, true>), grid, dim3(64 * G), 0, st, a); \
    else hipLaunchKernelGGL((band_t_seg_kernel<DIR, v, M, false>), grid, dim3(64 * G), 0, st, a);        \
    return true;
    GATT_BAND_T_NR(2) GATT_BAND_T_NR(3) GATT_BAND_T_NR(4) GATT_BAND_T_NR(5)
#undef GATT_BAND_T_NR
  }
  return false;
}

// Segment kernel along y and z for lines of up to 256 points (16 wavefronts of
// 16 points) and bands of up to 9 diagonals, where the windows stay in
// registers; otherwise the point kernel.
template <int DIR>
static bool launch_band_t_seg(const BandTArgs& a, int nmembers, hipStream_t st) {
  if ((long long)a.nf * nmembers > 65535) return false;
  const int H = a.nR - 1;
  if (a.N <= 256 && 16 + H <= a.N) return launch_band_t_seg_m<DIR, 16>(a, nmembers, st);
  return false;
}

void launch_band_t(const BandTArgs& a, int nmembers, hipStream_t st) {
  if (a.dir == 1 && launch_band_t_seg<1>(a, nmembers, st)) return;
  if (a.dir == 2 && launch_band_t_seg<2>(a, nmembers, st)) return;
  // members in chunks that keep the z extent of the grid within its limit
  const int per_chunk = std::max(1, 65535 / (a.nz * a.nf));
  for (int m0 = 0; m0 < nmembers; m0 += per_chunk) {
    const int nm = std::min(per_chunk, nmembers - m0);
    dim3 grid((unsigned)((a.nx + 255) / 256), (unsigned)a.ny, (unsigned)(a.nz * a.nf * nm));
    if (a.dir == 0) hipLaunchKernelGGL(band_t_kernel<0>, grid, dim3(256), 0, st, a, m0);
    else if (a.dir == 1) hipLaunchKernelGGL(band_t_kernel<1>, grid, dim3(256), 0, st, a, m0);
    else hipLaunchKernelGGL(band_t_kernel<2>, grid, dim3(256), 0, st, a, m0);
  }
}

// -------------------------------------------------------------- axpby
__global__ __launch_bounds__(256) void axpby_kernel(AxpbyArgs a) {
  const long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (p >= a.np) return;
  const int f = (int)blockIdx.y % a.nf;
  const int mem = (int)blockIdx.y / a.nf;
  const double x = ld_nt(a.x[f] + (long long)mem * a.x_stride + p);
  double* y = a.y[f] + (long long)mem * a.y_stride + p;
  // b == 0 overwrites y (and a == b == 0 clears it, whatever y held)
  st_nt(y, a.b == 0.0 ? (a.a == 0.0 ? 0.0 : a.a * x) : a.a * x + a.b * ld_nt(y));
}

void launch_axpby(const AxpbyArgs& a, int nmembers, hipStream_t st) {
  hipLaunchKernelGGL(axpby_kernel, dim3(grid_for(a.np, 256), (unsigned)(a.nf * nmembers)), dim3(256), 0, st, a);
}

__global__ __launch_bounds__(256) void adj_combine_kernel(AdjCombineArgs a) {
  const long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (p >= a.np) return;
  const int f = (int)blockIdx.y % a.nf;
  const long long o = (long long)((int)blockIdx.y / a.nf) * a.stride + p;
  const double sb = ld_nt(a.sb[f] + o);
  if (a.mode == 0) {
    double* acc = a.acc[f] + o;
    st_nt(acc, a.first ? 1.0 * sb : 1.0 * sb + 1.0 * ld_nt(acc));
    const double t = a.a1 * ld_nt(a.qb[f] + o);
    st_nt(a.k[f] + o, a.a2 * sb + 1.0 * t);
  } else {
    double* qb = a.qb[f] + o;
    const double t = 1.0 * ld_nt(a.acc[f] + o) + 1.0 * ld_nt(qb);
    st_nt(qb, 1.0 * sb + 1.0 * t);
  }
}

void launch_adj_combine(const AdjCombineArgs& a, int nmembers, hipStream_t st) {
  hipLaunchKernelGGL(adj_combine_kernel, dim3(grid_for(a.np, 256), (unsigned)(a.nf * nmembers)), dim3(256), 0, st,
                     a);
}

// -------------------------------------------------------------- viscous terms
struct Mu {
  double mu, dmu;   // mu(E) and d mu / d E
};

__device__ __forceinline__ Mu sutherland(const FlowConst& c, double E) {
  const double den = c.muC * E + c.muB;
  const double mu1 = c.muA / den;
  const double mu2 = c.muC * E;
  const double mu = mu1 * mu2 * sqrt(mu2);
  // d ln(mu)/dE = -muC/(muC E + muB) + 3/(2E)
  return {mu, mu * (1.5 / E - c.muC / den)};
}

__global__ __launch_bounds__(256) void visc_tl_kernel(ViscLinArgs a) {
  const long long n = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (n >= a.np) return;
  const long long np = a.np;
  const long long mo = (long long)blockIdx.y * a.member_stride;
  const long long wo = (long long)blockIdx.y * a.work_stride;
  const long long lo = (long long)blockIdx.y * a.lin_stride;
  const double* D = a.D + wo;
  const double* dD = a.dD + lo;
  double* dT = a.dT + lo;
  const FlowConst c = a.c;
  const Mu m = sutherland(c, ld_nt(a.E + mo + n));
  const double dE = ld_nt(a.dE + lo + n);
  const double rmu = c.invRe * m.mu;
  const double drmu = c.invRe * (m.dmu * dE);
  const double dmu = m.dmu * dE;
  const double DUx = ld_nt(D + 0 * np + n), DUy = ld_nt(D + 1 * np + n), DVx = ld_nt(D + 3 * np + n), DVy = ld_nt(D + 4 * np + n);
  const double DEx = ld_nt(D + 12 * np + n), DEy = ld_nt(D + 13 * np + n);
  const double tDUx = ld_nt(dD + 0 * np + n), tDUy = ld_nt(dD + 1 * np + n), tDVx = ld_nt(dD + 3 * np + n), tDVy = ld_nt(dD + 4 * np + n);
  const double tDEx = ld_nt(dD + 12 * np + n), tDEy = ld_nt(dD + 13 * np + n);
  if (a.is3d) {
    const double DUz = ld_nt(D + 2 * np + n), DVz = ld_nt(D + 5 * np + n), DWx = ld_nt(D + 6 * np + n), DWy = ld_nt(D + 7 * np + n);
    const double DWz = ld_nt(D + 8 * np + n), DEz = ld_nt(D + 14 * np + n);
    const double tDUz = ld_nt(dD + 2 * np + n), tDVz = ld_nt(dD + 5 * np + n), tDWx = ld_nt(dD + 6 * np + n), tDWy = ld_nt(dD + 7 * np + n);
    const double tDWz = ld_nt(dD + 8 * np + n), tDEz = ld_nt(dD + 14 * np + n);
    const double tz = -(2.0 / 3.0 * (DUx + DVy + DWz));
    const double dtz = -(2.0 / 3.0 * (tDUx + tDVy + tDWz));
    st_nt(dT + 0 * np + n, drmu * (2.0 * DUx + tz) + rmu * (2.0 * tDUx + dtz));
    st_nt(dT + 1 * np + n, drmu * (2.0 * DVy + tz) + rmu * (2.0 * tDVy + dtz));
    st_nt(dT + 2 * np + n, drmu * (2.0 * DWz + tz) + rmu * (2.0 * tDWz + dtz));
    st_nt(dT + 3 * np + n, drmu * (DUy + DVx) + rmu * (tDUy + tDVx));
    st_nt(dT + 4 * np + n, drmu * (DUz + DWx) + rmu * (tDUz + tDWx));
    st_nt(dT + 5 * np + n, drmu * (DVz + DWy) + rmu * (tDVz + tDWy));
    st_nt(dT + 6 * np + n, -(c.qk * (dmu * DEx + m.mu * tDEx)));
    st_nt(dT + 7 * np + n, -(c.qk * (dmu * DEy + m.mu * tDEy)));
    st_nt(dT + 8 * np + n, -(c.qk * (dmu * DEz + m.mu * tDEz)));
  } else {
    const double dv = 2.0 / 3.0 * (DUx + DVy);
    const double ddv = 2.0 / 3.0 * (tDUx + tDVy);
    st_nt(dT + 0 * np + n, drmu * (2.0 * DUx - dv) + rmu * (2.0 * tDUx - ddv));
    st_nt(dT + 1 * np + n, drmu * (2.0 * DVy - dv) + rmu * (2.0 * tDVy - ddv));
    st_nt(dT + 3 * np + n, drmu * (DUy + DVx) + rmu * (tDUy + tDVx));
    st_nt(dT + 6 * np + n, -(c.qk * (dmu * DEx + m.mu * tDEx)));
    st_nt(dT + 7 * np + n, -(c.qk * (dmu * DEy + m.mu * tDEy)));
  }
}

__global__ __launch_bounds__(256) void visc_adj_kernel(ViscLinArgs a) {
  const long long n = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (n >= a.np) return;
  const long long np = a.np;
  const long long mo = (long long)blockIdx.y * a.member_stride;
  const long long wo = (long long)blockIdx.y * a.work_stride;
  const long long lo = (long long)blockIdx.y * a.lin_stride;
  const double* D = a.D + wo;
  const double* Tb = a.dT + lo;
  double* Db = a.Dbar + lo;
  const FlowConst c = a.c;
  const Mu m = sutherland(c, ld_nt(a.E + mo + n));
  const double rmu = c.invRe * m.mu;
  const double DUx = ld_nt(D + 0 * np + n), DUy = ld_nt(D + 1 * np + n), DVx = ld_nt(D + 3 * np + n), DVy = ld_nt(D + 4 * np + n);
  const double DEx = ld_nt(D + 12 * np + n), DEy = ld_nt(D + 13 * np + n);
  const double b0 = ld_nt(Tb + 0 * np + n), b1 = ld_nt(Tb + 1 * np + n), b3 = ld_nt(Tb + 3 * np + n);
  const double b6 = ld_nt(Tb + 6 * np + n), b7 = ld_nt(Tb + 7 * np + n);
  double rmub, mub;
  if (a.is3d) {
    const double DUz = ld_nt(D + 2 * np + n), DVz = ld_nt(D + 5 * np + n), DWx = ld_nt(D + 6 * np + n), DWy = ld_nt(D + 7 * np + n);
    const double DWz = ld_nt(D + 8 * np + n), DEz = ld_nt(D + 14 * np + n);
    const double b2 = ld_nt(Tb + 2 * np + n), b4 = ld_nt(Tb + 4 * np + n), b5 = ld_nt(Tb + 5 * np + n), b8 = ld_nt(Tb + 8 * np + n);
    const double tz = -(2.0 / 3.0 * (DUx + DVy + DWz));
    rmub = b0 * (2.0 * DUx + tz) + b1 * (2.0 * DVy + tz) + b2 * (2.0 * DWz + tz) + b3 * (DUy + DVx) +
           b4 * (DUz + DWx) + b5 * (DVz + DWy);
    const double tzb = rmu * (b0 + b1 + b2);
    const double divb = -(2.0 / 3.0) * tzb;
    st_nt(Db + 0 * np + n, ld_nt(Db + 0 * np + n) + (2.0 * rmu * b0 + divb));
    st_nt(Db + 4 * np + n, ld_nt(Db + 4 * np + n) + (2.0 * rmu * b1 + divb));
    st_nt(Db + 8 * np + n, ld_nt(Db + 8 * np + n) + (2.0 * rmu * b2 + divb));
    st_nt(Db + 1 * np + n, ld_nt(Db + 1 * np + n) + (rmu * b3));
    st_nt(Db + 3 * np + n, ld_nt(Db + 3 * np + n) + (rmu * b3));
    st_nt(Db + 2 * np + n, ld_nt(Db + 2 * np + n) + (rmu * b4));
    st_nt(Db + 6 * np + n, ld_nt(Db + 6 * np + n) + (rmu * b4));
    st_nt(Db + 5 * np + n, ld_nt(Db + 5 * np + n) + (rmu * b5));
    st_nt(Db + 7 * np + n, ld_nt(Db + 7 * np + n) + (rmu * b5));
    st_nt(Db + 12 * np + n, ld_nt(Db + 12 * np + n) + (-(c.qk * m.mu * b6)));
    st_nt(Db + 13 * np + n, ld_nt(Db + 13 * np + n) + (-(c.qk * m.mu * b7)));
    st_nt(Db + 14 * np + n, ld_nt(Db + 14 * np + n) + (-(c.qk * m.mu * b8)));
    mub = c.invRe * rmub - c.qk * (DEx * b6 + DEy * b7 + DEz * b8);
  } else {
    const double dv = 2.0 / 3.0 * (DUx + DVy);
    rmub = b0 * (2.0 * DUx - dv) + b1 * (2.0 * DVy - dv) + b3 * (DUy + DVx);
    const double divb = -(2.0 / 3.0) * (rmu * (b0 + b1));
    st_nt(Db + 0 * np + n, ld_nt(Db + 0 * np + n) + (2.0 * rmu * b0 + divb));
    st_nt(Db + 4 * np + n, ld_nt(Db + 4 * np + n) + (2.0 * rmu * b1 + divb));
    st_nt(Db + 1 * np + n, ld_nt(Db + 1 * np + n) + (rmu * b3));
    st_nt(Db + 3 * np + n, ld_nt(Db + 3 * np + n) + (rmu * b3));
    st_nt(Db + 12 * np + n, ld_nt(Db + 12 * np + n) + (-(c.qk * m.mu * b6)));
    st_nt(Db + 13 * np + n, ld_nt(Db + 13 * np + n) + (-(c.qk * m.mu * b7)));
    mub = c.invRe * rmub - c.qk * (DEx * b6 + DEy * b7);
  }
  st_nt(a.dEbar + lo + n, ld_nt(a.dEbar + lo + n) + (m.dmu * mub));
}

void launch_visc_tl(const ViscLinArgs& a, int nmembers, hipStream_t st) {
  hipLaunchKernelGGL(visc_tl_kernel, dim3(grid_for(a.np, 256), nmembers), dim3(256), 0, st, a);
}
void launch_visc_adj(const ViscLinArgs& a, int nmembers, hipStream_t st) {
  hipLaunchKernelGGL(visc_adj_kernel, dim3(grid_for(a.np, 256), nmembers), dim3(256), 0, st, a);
}

// -------------------------------------------------------------- RHS assembly
// Base fields of one point, with the combinations the RHS uses.
struct Base {
  double U, V, W, R, E;
  double DUx, DUy, DUz, DVx, DVy, DVz, DWx, DWy, DWz, DRx, DRy, DRz, DEx, DEy, DEz;
  double txx, tyy, tzz, txy, txz, tyz;
  double vtx, vty, vtz, P, DPx, DPy, DPz, div, Phi;
};

__device__ __forceinline__ Base load_base(const AssembleLinArgs& a, long long n) {
  const long long np = a.np;
  const long long mo = (long long)blockIdx.y * a.member_stride;
  const long long wo = (long long)blockIdx.y * a.work_stride;
  const double* S = a.S + mo;
  const double* D = a.D + wo;
  const double* T = a.T + wo;
  const double* G = a.G + wo;
  const double g1 = a.c.gm1;
  Base b;
  b.U = ld_nt(S + 0 * np + n); b.V = ld_nt(S + 1 * np + n); b.R = ld_nt(S + 3 * np + n); b.E = ld_nt(S + 4 * np + n);
  b.DUx = ld_nt(D + 0 * np + n); b.DUy = ld_nt(D + 1 * np + n); b.DVx = ld_nt(D + 3 * np + n); b.DVy = ld_nt(D + 4 * np + n);
  b.DRx = ld_nt(D + 9 * np + n); b.DRy = ld_nt(D + 10 * np + n); b.DEx = ld_nt(D + 12 * np + n); b.DEy = ld_nt(D + 13 * np + n);
  b.txx = ld_nt(T + 0 * np + n); b.tyy = ld_nt(T + 1 * np + n); b.txy = ld_nt(T + 3 * np + n);
  b.P = g1 * b.R * b.E;
  b.DPx = g1 * (b.R * b.DEx + b.DRx * b.E);
  b.DPy = g1 * (b.R * b.DEy + b.DRy * b.E);
  if (a.is3d) {
    b.W = ld_nt(S + 2 * np + n);
    b.DUz = ld_nt(D + 2 * np + n); b.DVz = ld_nt(D + 5 * np + n); b.DWx = ld_nt(D + 6 * np + n); b.DWy = ld_nt(D + 7 * np + n);
    b.DWz = ld_nt(D + 8 * np + n); b.DRz = ld_nt(D + 11 * np + n); b.DEz = ld_nt(D + 14 * np + n);
    b.tzz = ld_nt(T + 2 * np + n); b.txz = ld_nt(T + 4 * np + n); b.tyz = ld_nt(T + 5 * np + n);
    b.DPz = g1 * (b.R * b.DEz + b.DRz * b.E);
    b.vtx = ld_nt(G + 0 * np + n) + ld_nt(G + 4 * np + n) + ld_nt(G + 8 * np + n);
    b.vty = ld_nt(G + 5 * np + n) + ld_nt(G + 1 * np + n) + ld_nt(G + 9 * np + n);
    b.vtz = ld_nt(G + 10 * np + n) + ld_nt(G + 2 * np + n) + ld_nt(G + 6 * np + n);
    b.div = b.DUx + b.DVy + b.DWz;
    b.Phi = -(b.P * b.div) + b.txx * b.DUx + b.tyy * b.DVy + b.tzz * b.DWz + b.txy * (b.DVx + b.DUy) +
            b.txz * (b.DWx + b.DUz) + b.tyz * (b.DWy + b.DVz) - ld_nt(G + 3 * np + n) - ld_nt(G + 7 * np + n) - ld_nt(G + 11 * np + n);
  } else {
    b.W = 0; b.DUz = b.DVz = b.DWx = b.DWy = b.DWz = b.DRz = b.DEz = 0;
    b.tzz = b.txz = b.tyz = 0; b.DPz = 0; b.vtz = 0;
    b.vtx = ld_nt(G + 0 * np + n) + ld_nt(G + 4 * np + n);
    b.vty = ld_nt(G + 5 * np + n) + ld_nt(G + 1 * np + n);
    b.div = b.DUx + b.DVy;
    b.Phi = -(b.P * b.div) + b.txx * b.DUx + b.tyy * b.DVy + b.txy * (b.DVx + b.DUy) - ld_nt(G + 3 * np + n) -
            ld_nt(G + 7 * np + n);
  }
  return b;
}

__global__ __launch_bounds__(256) void assemble_tl_kernel(AssembleLinArgs a) {
  const long long n = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (n >= a.np) return;
  const long long np = a.np;
  const long long lo = (long long)blockIdx.y * a.lin_stride;
  const Base b = load_base(a, n);
  const double g1 = a.c.gm1;
  const double* dS = a.dS + lo;
  const double* dD = a.dD + lo;
  const double* dT = a.dT + lo;
  const double* dG = a.dG + lo;
  const double tU = ld_nt(dS + 0 * np + n), tV = ld_nt(dS + 1 * np + n), tR = ld_nt(dS + 3 * np + n), tE = ld_nt(dS + 4 * np + n);
  const double tDUx = ld_nt(dD + 0 * np + n), tDUy = ld_nt(dD + 1 * np + n), tDVx = ld_nt(dD + 3 * np + n), tDVy = ld_nt(dD + 4 * np + n);
  const double tDRx = ld_nt(dD + 9 * np + n), tDRy = ld_nt(dD + 10 * np + n), tDEx = ld_nt(dD + 12 * np + n), tDEy = ld_nt(dD + 13 * np + n);
  const double ttxx = ld_nt(dT + 0 * np + n), ttyy = ld_nt(dT + 1 * np + n), ttxy = ld_nt(dT + 3 * np + n);
  const double tP = g1 * (tR * b.E + b.R * tE);
  const double tDPx = g1 * (tR * b.DEx + b.R * tDEx + tDRx * b.E + b.DRx * tE);
  const double tDPy = g1 * (tR * b.DEy + b.R * tDEy + tDRy * b.E + b.DRy * tE);
  const double invR = 1.0 / b.R;
  const double invR2 = invR * invR;
  double d[5];
  if (a.is3d) {
    const double tW = ld_nt(dS + 2 * np + n);
    const double tDUz = ld_nt(dD + 2 * np + n), tDVz = ld_nt(dD + 5 * np + n), tDWx = ld_nt(dD + 6 * np + n), tDWy = ld_nt(dD + 7 * np + n);
    const double tDWz = ld_nt(dD + 8 * np + n), tDRz = ld_nt(dD + 11 * np + n), tDEz = ld_nt(dD + 14 * np + n);
    const double ttzz = ld_nt(dT + 2 * np + n), ttxz = ld_nt(dT + 4 * np + n), ttyz = ld_nt(dT + 5 * np + n);
    const double tDPz = g1 * (tR * b.DEz + b.R * tDEz + tDRz * b.E + b.DRz * tE);
    const double tvtx = ld_nt(dG + 0 * np + n) + ld_nt(dG + 4 * np + n) + ld_nt(dG + 8 * np + n);
    const double tvty = ld_nt(dG + 5 * np + n) + ld_nt(dG + 1 * np + n) + ld_nt(dG + 9 * np + n);
    const double tvtz = ld_nt(dG + 10 * np + n) + ld_nt(dG + 2 * np + n) + ld_nt(dG + 6 * np + n);
    const double tdiv = tDUx + tDVy + tDWz;
    d[3] = -(tR * b.div + b.R * tdiv) - (tDRx * b.U + b.DRx * tU) - (tDRy * b.V + b.DRy * tV) -
           (tDRz * b.W + b.DRz * tW);
    d[0] = -(tDUx * b.U + b.DUx * tU + tDUy * b.V + b.DUy * tV + tDUz * b.W + b.DUz * tW) +
           (tvtx - tDPx) * invR - (b.vtx - b.DPx) * tR * invR2;
    d[1] = -(tDVx * b.U + b.DVx * tU + tDVy * b.V + b.DVy * tV + tDVz * b.W + b.DVz * tW) +
           (tvty - tDPy) * invR - (b.vty - b.DPy) * tR * invR2;
    d[2] = -(tDWx * b.U + b.DWx * tU + tDWy * b.V + b.DWy * tV + tDWz * b.W + b.DWz * tW) +
           (tvtz - tDPz) * invR - (b.vtz - b.DPz) * tR * invR2;
    const double tPhi = -(tP * b.div + b.P * tdiv) + ttxx * b.DUx + b.txx * tDUx + ttyy * b.DVy + b.tyy * tDVy +
                        ttzz * b.DWz + b.tzz * tDWz + ttxy * (b.DVx + b.DUy) + b.txy * (tDVx + tDUy) +
                        ttxz * (b.DWx + b.DUz) + b.txz * (tDWx + tDUz) + ttyz * (b.DWy + b.DVz) +
                        b.tyz * (tDWy + tDVz) - ld_nt(dG + 3 * np + n) - ld_nt(dG + 7 * np + n) - ld_nt(dG + 11 * np + n);
    d[4] = -(tDEx * b.U + b.DEx * tU + tDEy * b.V + b.DEy * tV + tDEz * b.W + b.DEz * tW) + tPhi * invR -
           b.Phi * tR * invR2;
  } else {
    const double tvtx = ld_nt(dG + 0 * np + n) + ld_nt(dG + 4 * np + n);
    const double tvty = ld_nt(dG + 5 * np + n) + ld_nt(dG + 1 * np + n);
    const double tdiv = tDUx + tDVy;
    d[3] = -(tR * b.div + b.R * tdiv) - (tDRx * b.U + b.DRx * tU) - (tDRy * b.V + b.DRy * tV);
    d[0] = -(tDUx * b.U + b.DUx * tU + tDUy * b.V + b.DUy * tV) + (tvtx - tDPx) * invR -
           (b.vtx - b.DPx) * tR * invR2;
    d[1] = -(tDVx * b.U + b.DVx * tU + tDVy * b.V + b.DVy * tV) + (tvty - tDPy) * invR -
           (b.vty - b.DPy) * tR * invR2;
    d[2] = 0.0;
    const double tPhi = -(tP * b.div + b.P * tdiv) + ttxx * b.DUx + b.txx * tDUx + ttyy * b.DVy + b.tyy * tDVy +
                        ttxy * (b.DVx + b.DUy) + b.txy * (tDVx + tDUy) - ld_nt(dG + 3 * np + n) - ld_nt(dG + 7 * np + n);
    d[4] = -(tDEx * b.U + b.DEx * tU + tDEy * b.V + b.DEy * tV) + tPhi * invR - b.Phi * tR * invR2;
  }
  if (a.sums && a.is3d) {
    // turbulenceGenerator.F90:57-61 linearised: du = fk u - g Us / N with fk = eps N / (4 ET)
    // and ET = sum(u u + v v + w w), so d(du) = fk du' + u dfk - g dUs / N with
    // dfk = -eps N (2 sum(u u' + v v' + w w')) / (4 ET^2); dE = -g (Es / N - e0) gives -g dEs / N
    const double* sm = a.sums + 8 * blockIdx.y;
    const double* ls = a.lsums + 8 * blockIdx.y;
    const double ET = sm[0], Nd = sm[5];
    const double gam = 10 * a.forcing_eps;
    const double fk = 0.25 * a.forcing_eps * Nd / ET;
    const double dfk = -(0.25 * a.forcing_eps * Nd * (2.0 * ls[0]) / (ET * ET));
    d[0] = d[0] + (fk * tU + dfk * b.U) - gam * ls[1] / Nd;
    d[1] = d[1] + (fk * tV + dfk * b.V) - gam * ls[2] / Nd;
    d[2] = d[2] + (fk * ld_nt(dS + 2 * np + n) + dfk * b.W) - gam * ls[3] / Nd;
    d[4] = d[4] - gam * ls[4] / Nd;
  }
  const double* Q = a.dQ + lo;
  double* O = a.dOut + lo;
  double* Dl = a.dDelta + lo;
  const double dt = a.dt;
#pragma unroll
  for (int f = 0; f < 5; ++f) {
    if (!a.is3d && f == 2) continue;
    const long long o = f * np + n;
    switch (a.stage) {
      case 0: st_nt(O + o, d[f]); break;
      case 1: st_nt(Dl + o, dt / 6.0 * d[f]); st_nt(O + o, ld_nt(Q + o) + dt / 2.0 * d[f]); break;
      case 2: st_nt(Dl + o, ld_nt(Dl + o) + dt / 3.0 * d[f]); st_nt(O + o, ld_nt(Q + o) + dt / 2.0 * d[f]); break;
      case 3: st_nt(Dl + o, ld_nt(Dl + o) + dt / 3.0 * d[f]); st_nt(O + o, ld_nt(Q + o) + dt * d[f]); break;
      default: st_nt(O + o, ld_nt(Q + o) + ld_nt(Dl + o) + dt / 6.0 * d[f]); break;
    }
  }
}

__global__ __launch_bounds__(256) void assemble_adj_kernel(AssembleLinArgs a) {
  const long long n = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (n >= a.np) return;
  const long long np = a.np;
  const long long lo = (long long)blockIdx.y * a.lin_stride;
  const Base b = load_base(a, n);
  const double g1 = a.c.gm1;
  const double* db = a.dbar + lo;
  const double s = a.dbar_scale;
  const double aU = s * ld_nt(db + 0 * np + n), aV = s * ld_nt(db + 1 * np + n), aR = s * ld_nt(db + 3 * np + n),
               aE = s * ld_nt(db + 4 * np + n);
  const double aW = a.is3d ? s * ld_nt(db + 2 * np + n) : 0.0;
  const double invR = 1.0 / b.R;
  const double invR2 = invR * invR;
  // momentum
  double Ub = -(b.DUx * aU) - b.DVx * aV - b.DWx * aW;
  double Vb = -(b.DUy * aU) - b.DVy * aV - b.DWy * aW;
  double Wb = -(b.DUz * aU) - b.DVz * aV - b.DWz * aW;
  double DUxb = -(b.U * aU), DUyb = -(b.V * aU), DUzb = -(b.W * aU);
  double DVxb = -(b.U * aV), DVyb = -(b.V * aV), DVzb = -(b.W * aV);
  double DWxb = -(b.U * aW), DWyb = -(b.V * aW), DWzb = -(b.W * aW);
  const double vtxb = aU * invR, vtyb = aV * invR, vtzb = aW * invR;
  const double DPxb = -(aU * invR), DPyb = -(aV * invR), DPzb = -(aW * invR);
  double Rb = -((b.vtx - b.DPx) * invR2 * aU) - (b.vty - b.DPy) * invR2 * aV - (b.vtz - b.DPz) * invR2 * aW;
  // continuity
  Rb += -(b.div * aR);
  double divb = -(b.R * aR);
  double DRxb = -(b.U * aR), DRyb = -(b.V * aR), DRzb = -(b.W * aR);
  Ub += -(b.DRx * aR);
  Vb += -(b.DRy * aR);
  Wb += -(b.DRz * aR);
  // energy
  Ub += -(b.DEx * aE);
  Vb += -(b.DEy * aE);
  Wb += -(b.DEz * aE);
  double DExb = -(b.U * aE), DEyb = -(b.V * aE), DEzb = -(b.W * aE);
  const double Phib = aE * invR;
  Rb += -(b.Phi * invR2 * aE);
  const double Pb = -(b.div * Phib);
  divb += -(b.P * Phib);
  const double txxb = b.DUx * Phib, tyyb = b.DVy * Phib, tzzb = b.DWz * Phib;
  const double txyb = (b.DVx + b.DUy) * Phib, txzb = (b.DWx + b.DUz) * Phib, tyzb = (b.DWy + b.DVz) * Phib;
  DUxb += b.txx * Phib;
  DVyb += b.tyy * Phib;
  DWzb += b.tzz * Phib;
  DVxb += b.txy * Phib;
  DUyb += b.txy * Phib;
  DWxb += b.txz * Phib;
  DUzb += b.txz * Phib;
  DWyb += b.tyz * Phib;
  DVzb += b.tyz * Phib;
  // pressure gradient DP = g1 (R DE + DR E) and P = g1 R E
  Rb += g1 * (b.DEx * DPxb + b.DEy * DPyb + b.DEz * DPzb) + g1 * b.E * Pb;
  DExb += g1 * b.R * DPxb;
  DEyb += g1 * b.R * DPyb;
  DEzb += g1 * b.R * DPzb;
  DRxb += g1 * b.E * DPxb;
  DRyb += g1 * b.E * DPyb;
  DRzb += g1 * b.E * DPzb;
  double Eb = g1 * (b.DRx * DPxb + b.DRy * DPyb + b.DRz * DPzb) + g1 * b.R * Pb;
  // divergence
  DUxb += divb;
  DVyb += divb;
  DWzb += divb;

  if (a.sums && a.is3d) {
    // transpose of the linearised forcing (see assemble_tl_kernel): the point-wise fk terms,
    // ET's adjoint (2 u per point times sum(a_u u + a_v v + a_w w) dfk/dET) and the means'
    // adjoints (-g sum(a) / N on every point)
    const double* sm = a.sums + 8 * blockIdx.y;
    const double* ls = a.lsums + 8 * blockIdx.y;
    const double ET = sm[0], Nd = sm[5];
    const double gam = 10 * a.forcing_eps;
    const double fk = 0.25 * a.forcing_eps * Nd / ET;
    const double ETb = -(0.25 * a.forcing_eps * Nd / (ET * ET)) * ls[0];
    Ub += fk * aU + 2.0 * b.U * ETb - gam * ls[1] / Nd;
    Vb += fk * aV + 2.0 * b.V * ETb - gam * ls[2] / Nd;
    Wb += fk * aW + 2.0 * b.W * ETb - gam * ls[3] / Nd;
    Eb += -(gam * ls[4] / Nd);
  }
  double* Sb = a.Sbar + lo;   // the stage-state adjoint starts here (later kernels accumulate)
  st_nt(Sb + 0 * np + n, Ub);
  st_nt(Sb + 1 * np + n, Vb);
  if (a.is3d) st_nt(Sb + 2 * np + n, Wb);
  st_nt(Sb + 3 * np + n, Rb);
  st_nt(Sb + 4 * np + n, Eb);
  double* Db = a.Dbar + lo;
  st_nt(Db + 0 * np + n, DUxb); st_nt(Db + 1 * np + n, DUyb); st_nt(Db + 2 * np + n, DUzb);
  st_nt(Db + 3 * np + n, DVxb); st_nt(Db + 4 * np + n, DVyb); st_nt(Db + 5 * np + n, DVzb);
  st_nt(Db + 6 * np + n, DWxb); st_nt(Db + 7 * np + n, DWyb); st_nt(Db + 8 * np + n, DWzb);
  st_nt(Db + 9 * np + n, DRxb); st_nt(Db + 10 * np + n, DRyb); st_nt(Db + 11 * np + n, DRzb);
  st_nt(Db + 12 * np + n, DExb); st_nt(Db + 13 * np + n, DEyb); st_nt(Db + 14 * np + n, DEzb);
  // the stresses (the heat-flux adjoints get their one contribution from the transposed
  // flux solves, which store rather than add)
  double* Tb = a.Tbar + lo;
  st_nt(Tb + 0 * np + n, txxb); st_nt(Tb + 1 * np + n, tyyb); st_nt(Tb + 2 * np + n, tzzb);
  st_nt(Tb + 3 * np + n, txyb); st_nt(Tb + 4 * np + n, txzb); st_nt(Tb + 5 * np + n, tyzb);
  // flux derivatives x: txx,txy,txz,qx | y: txy,tyy,tyz,qy | z: txz,tyz,tzz,qz: the same
  // four values for every direction, stored once per direction or once for all of them
  double* Gb = a.Gbar + lo;
  for (int d = 0; d < a.gbar_copies; ++d) {
    double* g = Gb + (long long)(4 * d) * np + n;
    st_nt(g, vtxb);
    st_nt(g + np, vtyb);
    if (a.is3d) st_nt(g + 2 * np, vtzb);
    st_nt(g + 3 * np, -Phib);
  }
}

// Linear sums of the forcing: block partials, then the second stage of the forcing sums.
__global__ __launch_bounds__(256) void lin_sums_partial_kernel(LinSumArgs a) {
  __shared__ double sh[5][256];
  const long long np = a.np;
  const double* S = a.S + (long long)blockIdx.y * a.base_stride;
  const double* X = a.X + (long long)blockIdx.y * a.lin_stride;
  double acc[5] = {0, 0, 0, 0, 0};
  for (long long n = blockIdx.x * (long long)blockDim.x + threadIdx.x; n < np; n += (long long)gridDim.x * blockDim.x) {
    const double xu = ld_nt(X + n), xv = ld_nt(X + np + n), xw = ld_nt(X + 2 * np + n);
    acc[0] += xu * ld_nt(S + n) + xv * ld_nt(S + np + n) + xw * ld_nt(S + 2 * np + n);
    acc[1] += xu;
    acc[2] += xv;
    acc[3] += xw;
    acc[4] += ld_nt(X + 4 * np + n);
  }
  for (int k = 0; k < 5; ++k) sh[k][threadIdx.x] = acc[k];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int k = 0; k < 5; ++k) sh[k][threadIdx.x] += sh[k][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x < 5) a.partial[((long long)blockIdx.y * a.nblocks + blockIdx.x) * 5 + threadIdx.x] = sh[threadIdx.x][0];
}

void launch_lin_sums(const LinSumArgs& a, int nmembers, hipStream_t st) {
  hipLaunchKernelGGL(lin_sums_partial_kernel, dim3(a.nblocks, nmembers), dim3(256), 0, st, a);
  SumArgs f{};
  f.np = a.np;
  f.partial = a.partial;
  f.sums = a.sums;
  f.nblocks = a.nblocks;
  launch_forcing_final(f, nmembers, st);
}

void launch_assemble_tl(const AssembleLinArgs& a, int nmembers, hipStream_t st) {
  hipLaunchKernelGGL(assemble_tl_kernel, dim3(grid_for(a.np, 256), nmembers), dim3(256), 0, st, a);
}
void launch_assemble_adj(const AssembleLinArgs& a, int nmembers, hipStream_t st) {
  hipLaunchKernelGGL(assemble_adj_kernel, dim3(grid_for(a.np, 256), nmembers), dim3(256), 0, st, a);
}

}  // namespace gatt
