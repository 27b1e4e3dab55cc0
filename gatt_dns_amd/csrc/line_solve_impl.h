// Batched compact line solves: banded right-hand side + pre-factored Thomas
// sweeps, for non-periodic and periodic lines along x, y or z.
//
// Restates solveSystem{X,Y,Z} and solveSystem{X,Y,Z}per
// (source/Fortran/derivs.F90:556-948) for gfx950:
//
//   * A group of G lanes of one wavefront owns one line; lane g holds the m
//     consecutive points [g*m, g*m+m) and their (2nR-1)-point halo in VGPRs.
//     The band product is computed from registers.
//   * The Thomas recurrences are first-order affine recurrences.  Each lane
//     composes the affine map of its segment (the composite's offset follows the
//     reference recurrence exactly), the group combines the maps with a
//     log2(G)-step wavefront scan, and each lane then re-runs the reference
//     recurrence from its exact start value.  Forward: d_i = (r_i - A d_{i-1}) B;
//     backward: x_i = d_i - C x_{i+1}.
//   * Periodic lines first reduce x_1 = D . r over the group, then apply the
//     reference's row-2 / row-N corrections and solve rows 2..N
//     (derivs.F90:785-802, including its index shift of B).
//   * x lines are contiguous: the G lanes of a line are adjacent lanes, so a
//     wavefront reads one 64*m-point stretch.  y and z lines are strided: lanes
//     with consecutive line index i are adjacent, so each load instruction reads
//     64/G consecutive doubles of G rows.
//   * Filter mode writes in place with the FilterCharTime blend fused
//     (derivs.F90:344-349).
#pragma once
#include "gatt_internal.h"

namespace gatt {

__device__ __forceinline__ double shfl_d(double v, int src) { return __shfl(v, src, 64); }
__device__ __forceinline__ double shfl_xor_d(double v, int mask) { return __shfl_xor(v, mask, 64); }

template <int DIR, int NR, int M>
__global__ __launch_bounds__(256) void line_solve_kernel(SolveArgs a) {
  constexpr int H = NR - 1;
  constexpr int W = M + 2 * H;
  const int lane = threadIdx.x & 63;
  const int wave = (int)((blockIdx.x * (unsigned)blockDim.x + threadIdx.x) >> 6);
  const int G = a.G;
  const int LPW = 64 / G;
  constexpr bool xdir = DIR == 0;
  // lane layout: x -> lanes of a line adjacent; y/z -> lines adjacent
  const int g = xdir ? (lane & (G - 1)) : (lane / LPW);
  const int l = xdir ? (lane / G) : (lane & (LPW - 1));
  const int step = xdir ? 1 : LPW;   // lane distance between neighbouring segments
  const int line = wave * LPW + l;
  const bool active = line < a.nlines;
  const int f = (int)blockIdx.y % a.nf;
  const int mem = (int)blockIdx.y / a.nf;
  const double* in = a.in[f] + (long long)mem * a.member_stride;
  double* out = a.out[f] + (long long)mem * a.member_stride;
  const int N = a.op.N;
  const int m = a.m;
  const int per = a.op.periodic;
  const int e0 = per;
  const int s = g * m;
  int cnt = N - s;
  cnt = cnt < 0 ? 0 : (cnt > m ? m : cnt);
  if (!active) cnt = 0;
  const int la = active ? line : 0;
  const int ia = la % a.n1, ib = la / a.n1;
  long long base;
  if (DIR == 0) base = (long long)a.nx * (ia + (long long)a.ny * ib);
  else if (DIR == 1) base = ia + (long long)a.nx * a.ny * ib;
  else base = ia + (long long)a.nx * ib;
  const long long stride = DIR == 0 ? 1 : a.stride;
  const int t = a.types[la];
  const double* Rt = a.op.R + (size_t)t * N * (2 * NR - 1);
  const double* At = a.op.A + (size_t)t * (N - 1);
  const double* Bt = a.op.B + (size_t)t * (N - per);
  const double* Ct = a.op.C + (size_t)t * (N - 1);

  // ---- load the segment and its halo (wrapped like the reference's modulo)
  double fw[W];
#pragma unroll
  for (int u = 0; u < W; ++u) {
    double v = 0.0;
    if (u < cnt + 2 * H) {
      int idx = s - H + u;
      idx = idx < 0 ? idx + N : (idx >= N ? idx - N : idx);
      v = in[base + (long long)idx * stride];
    }
    fw[u] = v;
  }

  // ---- band right-hand side, in the reference's summation order
  double r[M];
#pragma unroll
  for (int tt = 0; tt < M; ++tt) {
    double acc = 0.0;
    if (tt < cnt) {
      const double* Re = Rt + (s + tt);
      acc = Re[(size_t)H * N] * fw[tt + H];
#pragma unroll
      for (int q = 1; q <= H; ++q)
        acc = acc + Re[(size_t)(H + q) * N] * fw[tt + H + q] + Re[(size_t)(H - q) * N] * fw[tt + H - q];
    }
    r[tt] = acc;
  }

  // ---- periodic: x1 = D . r over the whole line, then rows 2 and N corrections
  double x1 = 0.0;
  if (per) {
    const double* Dt = a.op.D + (size_t)t * N;
    double part = 0.0;
#pragma unroll
    for (int tt = 0; tt < M; ++tt)
      if (tt < cnt) part = part + r[tt] * Dt[s + tt];
    for (int k = 1; k < G; k <<= 1) part += shfl_xor_d(part, k * step);
    x1 = part;
#pragma unroll
    for (int tt = 0; tt < M; ++tt) {
      if (tt < cnt) {
        const int e = s + tt;
        if (e == 1) r[tt] = r[tt] - At[0] * x1;
        if (e == N - 1) r[tt] = r[tt] - Ct[0] * x1;
      }
    }
  }

  // ---- forward sweep: composite map of the segment
  double P = 0.0, Q = 1.0;
#pragma unroll
  for (int tt = 0; tt < M; ++tt) {
    const int e = s + tt;
    if (tt < cnt && e >= e0) {
      const double fa = (e > e0) ? At[e - 1] : 0.0;
      const double fb = per ? Bt[e - 1] : Bt[e];
      P = (r[tt] - fa * P) * fb;
      Q = -(fa * fb) * Q;
    }
  }
  // inclusive scan over the segments of the line
  for (int k = 1; k < G; k <<= 1) {
    const double Pp = shfl_d(P, lane - k * step);
    const double Qp = shfl_d(Q, lane - k * step);
    if (g >= k) { P = P + Q * Pp; Q = Q * Qp; }
  }
  double dprev = shfl_d(P, lane - step);
  if (g == 0) dprev = 0.0;
#pragma unroll
  for (int tt = 0; tt < M; ++tt) {
    const int e = s + tt;
    if (tt < cnt && e >= e0) {
      const double fa = (e > e0) ? At[e - 1] : 0.0;
      const double fb = per ? Bt[e - 1] : Bt[e];
      const double d = (r[tt] - fa * dprev) * fb;
      r[tt] = d;
      dprev = d;
    }
  }

  // ---- backward sweep
  P = 0.0;
  Q = 1.0;
#pragma unroll
  for (int tt = M - 1; tt >= 0; --tt) {
    const int e = s + tt;
    if (tt < cnt && e >= e0) {
      if (e == N - 1) { P = r[tt]; Q = 0.0; }
      else { const double ce = Ct[e]; P = r[tt] - ce * P; Q = -ce * Q; }
    }
  }
  for (int k = 1; k < G; k <<= 1) {
    const double Pn = shfl_d(P, lane + k * step);
    const double Qn = shfl_d(Q, lane + k * step);
    if (g + k < G) { P = P + Q * Pn; Q = Q * Qn; }
  }
  double xnext = shfl_d(P, lane + step);
  if (g == G - 1) xnext = 0.0;
#pragma unroll
  for (int tt = M - 1; tt >= 0; --tt) {
    const int e = s + tt;
    if (tt < cnt && e >= e0) {
      const double x = (e == N - 1) ? r[tt] : r[tt] - Ct[e] * xnext;
      r[tt] = x;
      xnext = x;
    }
  }

  // ---- store (filter mode: in place with the FilterCharTime blend)
#pragma unroll
  for (int tt = 0; tt < M; ++tt) {
    if (tt < cnt) {
      const int e = s + tt;
      double v = (per && e == 0) ? x1 : r[tt];
      if (a.mode == 1 && a.theta > 0.0) {
        const double F0 = fw[tt + H];
        v = F0 + a.theta * (v - F0);
      }
      out[base + (long long)e * stride] = v;
    }
  }
}

template <int DIR, int NR, int M>
static void launch_nr_m(const SolveArgs& a, const LineShape& s, int nmembers, hipStream_t st) {
  const int lpw = 64 / s.G;
  const long long waves = (a.nlines + lpw - 1) / lpw;
  const int threads = 64 * s.waves_per_block;
  const long long blocks = (waves + s.waves_per_block - 1) / s.waves_per_block;
  dim3 grid((unsigned)blocks, (unsigned)(a.nf * nmembers));
  hipLaunchKernelGGL((line_solve_kernel<DIR, NR, M>), grid, dim3(threads), 0, st, a);
}

template <int DIR, int NR>
static bool launch_nr(const SolveArgs& a, const LineShape& s, int nmembers, hipStream_t st) {
  switch (s.M) {
    case 4: launch_nr_m<DIR, NR, 4>(a, s, nmembers, st); return true;
    case 8: launch_nr_m<DIR, NR, 8>(a, s, nmembers, st); return true;
    case 16: launch_nr_m<DIR, NR, 16>(a, s, nmembers, st); return true;
    case 32: launch_nr_m<DIR, NR, 32>(a, s, nmembers, st); return true;
  }
  return false;
}

template <int DIR>
void launch_line_solve_dir(const SolveArgs& a, const LineShape& s, int nmembers, hipStream_t st) {
  switch (a.op.nR) {
    case 2: launch_nr<DIR, 2>(a, s, nmembers, st); break;
    case 3: launch_nr<DIR, 3>(a, s, nmembers, st); break;
    case 4: launch_nr<DIR, 4>(a, s, nmembers, st); break;
    case 5: launch_nr<DIR, 5>(a, s, nmembers, st); break;
    case 6: launch_nr<DIR, 6>(a, s, nmembers, st); break;
    case 7: launch_nr<DIR, 7>(a, s, nmembers, st); break;
    case 8: launch_nr<DIR, 8>(a, s, nmembers, st); break;
    case 9: launch_nr<DIR, 9>(a, s, nmembers, st); break;
    case 10: launch_nr<DIR, 10>(a, s, nmembers, st); break;
    case 11: launch_nr<DIR, 11>(a, s, nmembers, st); break;
    default: break;
  }
}

}  // namespace gatt
