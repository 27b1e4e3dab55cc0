// Batched compact line solves: banded right-hand side + pre-factored Thomas
// sweeps, for non-periodic and periodic lines along x, y or z.
//
// Restates solveSystem{X,Y,Z} and solveSystem{X,Y,Z}per
// (source/Fortran/derivs.F90:556-948) for gfx950:
//
//   * A workgroup owns a tile of TL whole lines of one field.  The tile is
//     moved between HBM and LDS with unit-stride accesses: along x a tile is one
//     contiguous stretch of TL*N doubles; along y and z every line point of the
//     tile is a row of TL consecutive doubles.  Halo points of the band product
//     are read from LDS, so every global byte is loaded once and stored once.
//   * Inside the tile G lanes of one wavefront share a line; lane g owns the M
//     points [g*M, g*M+M) (M a power of two).  LDS holds point e of tile line l
//     at (e % M)*ROW + l*G + e/M, so the lanes of a wavefront read consecutive
//     LDS words at every step.
//   * The band product and both Thomas sweeps run in VGPRs.  The Thomas
//     recurrences are first-order affine recurrences: each lane composes the
//     affine map of its segment (the offset of the composite follows the
//     reference recurrence exactly), the G lanes combine their maps with a
//     log2(G)-step wavefront scan, and each lane re-runs the reference recurrence
//     from its exact start value.  Forward: d_i = (r_i - A d_{i-1}) B; backward:
//     x_i = d_i - C x_{i+1}.
//   * Periodic lines first reduce x_1 = D . r over the lanes of the line, then
//     apply the reference's row-2 / row-N corrections and solve rows 2..N
//     (derivs.F90:785-802, including its index shift of B).
//   * Filter mode writes in place with the FilterCharTime blend fused
//     (derivs.F90:344-349).
#pragma once
#include "gatt_internal.h"

namespace gatt {

__device__ __forceinline__ double shfl_d(double v, int src) { return __shfl(v, src, 64); }
__device__ __forceinline__ double shfl_xor_d(double v, int mask) { return __shfl_xor(v, mask, 64); }

// Global offset of the first point of line `line` along direction DIR.
template <int DIR>
__device__ __forceinline__ long long line_base(const SolveArgs& a, int line) {
  if (DIR == 0) return (long long)a.nx * line;                       // line = j + ny k
  if (DIR == 2) return line;                                         // line = i + nx j
  const int k = line / a.nx;                                         // line = i + nx k
  return (long long)(line - k * a.nx) + (long long)a.nx * a.ny * k;
}

template <int DIR, int NR, int M>
__device__ __forceinline__ void line_solve_lg_body(const SolveArgs& a, BlockPos bp) {
  extern __shared__ double lds[];
  constexpr int H = NR - 1;
  constexpr int W = M + 2 * H;
  const int G = a.G;
  const int TL = a.TL;               // lines per tile = (64/G) * waves per block
  const int ROW = a.lds_row;         // TL*G + 1
  const int N = a.op.N;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int g = lane & (G - 1);
  const int l = (tid >> 6) * (64 / G) + lane / G;
  const int line0 = bp.x * TL;
  const int ntile = min(TL, a.nlines - line0);
  const int f = (int)bp.y % a.nf;
  const int mem = (int)bp.y / a.nf;
  const double* in = a.in[f] + (long long)mem * a.in_stride;
  double* out = a.out[f] + (long long)mem * a.out_stride;
  const long long stride = DIR == 0 ? 1 : a.stride;

  // ---- HBM -> LDS, unit stride
  const int total = ntile * N;
  if (DIR == 0) {
    const double* src = in + line_base<0>(a, line0);   // the tile is contiguous
    batched_move<8>(total, [&](int idx) { return src[idx]; }, [&](int idx, double v) {
      const int lp = (int)(((float)idx + 0.5f) * a.inv_n);
      const int ep = idx - lp * N;
      lds[(ep & (M - 1)) * ROW + lp * G + ((unsigned)ep / M)] = v;
    });
  } else {
    // TL is a power of two
    batched_move<8>(
        TL * N,
        [&](int idx) {
          const int ep = (unsigned)idx / (unsigned)TL;
          const int lp = idx - ep * TL;
          return lp < ntile ? in[line_base<DIR>(a, line0 + lp) + ep * stride] : 0.0;
        },
        [&](int idx, double v) {
          const int ep = (unsigned)idx / (unsigned)TL;
          const int lp = idx - ep * TL;
          if (lp < ntile) lds[(ep & (M - 1)) * ROW + lp * G + ((unsigned)ep / M)] = v;
        });
  }
  __syncthreads();

  const bool active = l < ntile;
  const int per = a.op.periodic;
  const int e0 = per;
  const int s = g * M;
  int cnt = N - s;
  cnt = cnt < 0 ? 0 : (cnt > M ? M : cnt);
  if (!active) cnt = 0;
  const int t = a.types[active ? line0 + l : 0];
  const double* Rt = a.op.R + (size_t)t * N * (2 * NR - 1);
  const double* At = a.op.A + (size_t)t * (N - 1);
  const double* Bt = a.op.B + (size_t)t * (N - per);
  const double* Ct = a.op.C + (size_t)t * (N - 1);
  const int lbase = l * G;

  // ---- segment and halo from LDS (wrapped like the reference's modulo)
  double fw[W];
#pragma unroll
  for (int u = 0; u < W; ++u) {
    double v = 0.0;
    if (cnt > 0 && u < cnt + 2 * H) {
      int e = s - H + u;
      e = e < 0 ? e + N : (e >= N ? e - N : e);
      v = lds[(e & (M - 1)) * ROW + lbase + (unsigned)e / M];
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
    for (int k = 1; k < G; k <<= 1) part += shfl_xor_d(part, k);
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

  // ---- forward sweep: composite map of the segment, scan, re-run
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
  for (int k = 1; k < G; k <<= 1) {
    const double Pp = shfl_d(P, lane - k);
    const double Qp = shfl_d(Q, lane - k);
    if (g >= k) { P = P + Q * Pp; Q = Q * Qp; }
  }
  double dprev = shfl_d(P, lane - 1);
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
    const double Pn = shfl_d(P, lane + k);
    const double Qn = shfl_d(Q, lane + k);
    if (g + k < G) { P = P + Q * Pn; Q = Q * Qn; }
  }
  double xnext = shfl_d(P, lane + 1);
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

  // ---- results -> LDS (filter mode: FilterCharTime blend), then LDS -> HBM
  __syncthreads();
#pragma unroll
  for (int tt = 0; tt < M; ++tt) {
    if (tt < cnt) {
      const int e = s + tt;
      double v = (per && e == 0) ? x1 : r[tt];
      if (a.mode == 1 && a.theta > 0.0) {
        const double F0 = fw[tt + H];
        v = F0 + a.theta * (v - F0);
      }
      lds[tt * ROW + lbase + g] = v;
    }
  }
  __syncthreads();
  if (DIR == 0) {
    double* dst = out + line_base<0>(a, line0);
    batched_move<8>(
        total,
        [&](int idx) {
          const int lp = (int)(((float)idx + 0.5f) * a.inv_n);
          const int ep = idx - lp * N;
          return lds[(ep & (M - 1)) * ROW + lp * G + ((unsigned)ep / M)];
        },
        [&](int idx, double v) { dst[idx] = v; });
  } else {
    batched_move<8>(
        TL * N,
        [&](int idx) {
          const int ep = (unsigned)idx / (unsigned)TL;
          const int lp = idx - ep * TL;
          return lp < ntile ? lds[(ep & (M - 1)) * ROW + lp * G + ((unsigned)ep / M)] : 0.0;
        },
        [&](int idx, double v) {
          const int ep = (unsigned)idx / (unsigned)TL;
          const int lp = idx - ep * TL;
          if (lp < ntile) out[line_base<DIR>(a, line0 + lp) + ep * stride] = v;
        });
  }
}

template <int DIR, int NR, int M>
__global__ __launch_bounds__(256) void line_solve_kernel(SolveArgs a) {
  line_solve_lg_body<DIR, NR, M>(a, grid_pos());
}
template <int DIR, int NR, int M>
__global__ __launch_bounds__(256) void line_solve_ragged(const SolveArgs* __restrict__ args, RaggedMap map) {
  BlockPos bp;
  const int m = ragged_pos(map, bp);
  line_solve_lg_body<DIR, NR, M>(args[m], bp);
}

template <int DIR, int NR, int M>
static void launch_nr_m(const SolveArgs& a0, const LineShape& s, int nmembers, hipStream_t st) {
  SolveArgs a = a0;
  a.TL = (64 / s.G) * s.waves_per_block;
  a.lds_row = a.TL * s.G + 1;
  a.inv_n = 1.0f / (float)a.op.N;
  const long long tiles = (a.nlines + a.TL - 1) / a.TL;
  const size_t lds_bytes = (size_t)M * a.lds_row * sizeof(double);
  dim3 grid((unsigned)tiles, (unsigned)(a.nf * nmembers));
  if (record_launch((const void*)line_solve_ragged<DIR, NR, M>, a, grid, dim3(64 * s.waves_per_block), lds_bytes))
    return;
  hipLaunchKernelGGL((line_solve_kernel<DIR, NR, M>), grid, dim3(64 * s.waves_per_block), lds_bytes, st, a);
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
