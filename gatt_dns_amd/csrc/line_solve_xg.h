// Lane-group line solves along x: one wavefront per line.
//
// Restates solveSystemX and solveSystemXper (source/Fortran/derivs.F90:556-615,
// 737-806) with the fused coefficient records of line_solve_wsp.h.  An x line is
// contiguous, so lane l of the wavefront owns the M consecutive points
// [l*M, l*M+M) and every load and store of the wavefront covers one contiguous
// stretch of the line:
//
//   * window: the lane's points and H halo points either side, read straight
//     from HBM (the halo comes from the neighbouring lanes' cache lines);
//   * coefficients: a per-type table stored coefficient-major (table k holds
//     coefficient k of every point), so the lanes read consecutive words;
//   * Thomas sweeps: each lane composes the affine map of its points, the 64
//     maps are combined by a log2(64)-step wavefront scan (shuffles), and each
//     lane re-runs its points from the exact start value;
//   * periodic lines: x1 = D . r is a wavefront reduction.
//
// Any operator type per line (one line per wavefront), any line length up to
// 64*M points.  Results differ from the reference's sequential recurrence by
// rounding only.
#pragma once
#include "line_solve_wsp.h"

namespace gatt {

__device__ __forceinline__ double xg_shfl_up(double v, int d) { return __shfl_up(v, d, 64); }
__device__ __forceinline__ double xg_shfl_down(double v, int d) { return __shfl_down(v, d, 64); }
__device__ __forceinline__ double xg_shfl_xor(double v, int m) { return __shfl_xor(v, m, 64); }

template <int NR, int M, bool BLEND>
__device__ __forceinline__ void line_solve_xg_body(const SolveArgs& a, BlockPos bp) {
  extern __shared__ double shx[];
  constexpr int H = NR - 1;
  constexpr int NW = 2 * H + 1;
  constexpr int W = M + 2 * H;
  const int lane = threadIdx.x & 63;
  const int line = (int)bp.x * 4 + (int)(threadIdx.x >> 6);
  if (line >= a.nlines) return;   // whole wavefronts
  const int f = (int)bp.y % a.nf;
  const int mem = (int)bp.y / a.nf;
  const int N = a.op.N;
  const int per = a.op.periodic;
  const double* src = a.in[f] + (long long)mem * a.in_stride + (long long)a.nx * line;
  double* dst = a.out[f] + (long long)mem * a.out_stride + (long long)a.nx * line;
  const int s = lane * M;
  const int tu = __builtin_amdgcn_readfirstlane(a.types[line]);
  // coefficient k of this lane's point s + tt, lane-interleaved (64 consecutive doubles per
  // (k, tt) across the wavefront); points past the line end hold zeros
  const double* ct = a.fcoef_p + (size_t)tu * (NW + 3) * 64 * M + lane;
  auto coef = [&](int k, int tt) { return ct[((size_t)k * M + tt) * 64]; };

  // the line goes through this wavefront's LDS buffer: unit-stride loads and stores
  // in HBM, each lane's window of M + 2H points read from LDS
  double* buf = shx + (size_t)(threadIdx.x >> 6) * N;
  {
    // all M loads of the lane in flight before the LDS stores (N <= 64 M)
    double v[M];
#pragma unroll
    for (int k = 0; k < M; ++k) {
      const int e = lane + 64 * k;
      if (e < N) v[k] = ld_nt(src + e);
    }
#pragma unroll
    for (int k = 0; k < M; ++k) {
      const int e = lane + 64 * k;
      if (e < N) buf[e] = v[k];
    }
  }
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  double fw[W];
#pragma unroll
  for (int u = 0; u < W; ++u) {
    int e = s - H + u;
    e = e < 0 ? e + N : (e >= N ? e - N : e);
    e = e >= N ? e - N : e;
    e = e < 0 ? 0 : (e >= N ? N - 1 : e);
    fw[u] = buf[e];
  }
  // band product, scaled by the forward factor
  double r[M];
#pragma unroll
  for (int tt = 0; tt < M; ++tt) {
    const int e = s + tt;
    double acc = coef(H, tt) * fw[tt + H];
#pragma unroll
    for (int q = 1; q <= H; ++q) acc = fma(coef(H - q, tt), fw[tt + H - q], fma(coef(H + q, tt), fw[tt + H + q], acc));
    r[tt] = acc;
  }
  double x1 = 0.0;
  if (per) {
    double part = 0.0;
#pragma unroll
    for (int tt = 0; tt < M; ++tt) part = fma(r[tt], coef(NW + 2, tt), part);
#pragma unroll
    for (int m = 1; m < 64; m <<= 1) part += xg_shfl_xor(part, m);
    x1 = part;
    // row 2 and row N corrections (derivs.F90:792-793)
    const double A1 = a.fcoef_ends[(size_t)tu * 2], C1 = a.fcoef_ends[(size_t)tu * 2 + 1];
#pragma unroll
    for (int tt = 0; tt < M; ++tt) {
      if (s + tt == 1) r[tt] = fma(-A1, x1, r[tt]);
      if (s + tt == N - 1) r[tt] = fma(-C1, x1, r[tt]);
    }
  }
  // forward sweep: d_e = r_e - fa'_e d_{e-1}
  double P = 0.0, Q = 1.0;
#pragma unroll
  for (int tt = 0; tt < M; ++tt) {
    const double fa = coef(NW, tt);
    P = fma(-fa, P, r[tt]);
    Q = -fa * Q;
  }
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const double Pp = xg_shfl_up(P, d), Qp = xg_shfl_up(Q, d);
    if (lane >= d) {
      P = fma(Q, Pp, P);
      Q = Q * Qp;
    }
  }
  double dprev = xg_shfl_up(P, 1);
  if (lane == 0) dprev = 0.0;
#pragma unroll
  for (int tt = 0; tt < M; ++tt) {
    const double d = fma(-coef(NW, tt), dprev, r[tt]);
    r[tt] = d;
    dprev = d;
  }
  // backward sweep: x_e = d_e - ce_e x_{e+1}
  P = 0.0;
  Q = 1.0;
#pragma unroll
  for (int tt = M - 1; tt >= 0; --tt) {
    const double ce = coef(NW + 1, tt);
    P = fma(-ce, P, r[tt]);
    Q = -ce * Q;
  }
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const double Pn = xg_shfl_down(P, d), Qn = xg_shfl_down(Q, d);
    if (lane + d < 64) {
      P = fma(Q, Pn, P);
      Q = Q * Qn;
    }
  }
  double xnext = xg_shfl_down(P, 1);
  if (lane == 63) xnext = 0.0;
#pragma unroll
  for (int tt = M - 1; tt >= 0; --tt) {
    const double x = fma(-coef(NW + 1, tt), xnext, r[tt]);
    r[tt] = x;
    xnext = x;
  }
  if (per && lane == 0) r[0] = x1;
  // every lane has read its window (the shuffles above order the wavefront)
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
  for (int tt = 0; tt < M; ++tt) {
    const int e = s + tt;
    if (e < N) {
      double v = r[tt];
      if (BLEND) v = fma(a.theta, v - fw[tt + H], fw[tt + H]);
      buf[e] = v;
    }
  }
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  for (int e = lane; e < N; e += 64) st_nt(dst + e, buf[e]);
}

template <int NR, int M, bool BLEND>
__global__ __launch_bounds__(256) void line_solve_xg_kernel(SolveArgs a) {
  line_solve_xg_body<NR, M, BLEND>(a, grid_pos());
}
// The ragged form takes each member's points per lane from its argument block, so that the
// members of a ragged group share the launch with lines of different lengths, each with all
// the lanes its lines can use.
template <int NR, bool BLEND>
__global__ __launch_bounds__(256) void line_solve_xg_ragged(const SolveArgs* __restrict__ args, RaggedMap map) {
  BlockPos bp;
  const int m = ragged_pos(map, bp);
  const SolveArgs& a = args[m];
  if (a.xg_m == 4) line_solve_xg_body<NR, 4, BLEND>(a, bp);
  else if (a.xg_m == 8) line_solve_xg_body<NR, 8, BLEND>(a, bp);
  else line_solve_xg_body<NR, 16, BLEND>(a, bp);
}

template <int NR, int M>
static void launch_xg_m(const SolveArgs& a0, int nmembers, hipStream_t st) {
  static_assert(M == 4 || M == 8 || M == 16, "points per lane of the ragged form");
  SolveArgs a = a0;
  a.xg_m = M;
  dim3 grid((unsigned)((a.nlines + 3) / 4), (unsigned)(a.nf * nmembers));
  const size_t lds = sizeof(double) * 4 * (size_t)a.op.N;
  const bool blend = a.mode == 1 && a.theta > 0.0;
  if (record_launch(blend ? (const void*)line_solve_xg_ragged<NR, true> : (const void*)line_solve_xg_ragged<NR, false>,
                    a, grid, dim3(256), lds))
    return;
  if (blend)
    hipLaunchKernelGGL((line_solve_xg_kernel<NR, M, true>), grid, dim3(256), lds, st, a);
  else
    hipLaunchKernelGGL((line_solve_xg_kernel<NR, M, false>), grid, dim3(256), lds, st, a);
}

template <int NR>
static void launch_xg_nr(const SolveArgs& a, const LineShape& s, int nmembers, hipStream_t st) {
  switch (s.M) {
    case 4: launch_xg_m<NR, 4>(a, nmembers, st); break;
    case 8: launch_xg_m<NR, 8>(a, nmembers, st); break;
    case 16: launch_xg_m<NR, 16>(a, nmembers, st); break;
  }
}

inline void launch_line_solve_xg_impl(const SolveArgs& a, const LineShape& s, int nmembers, hipStream_t st) {
  switch (a.op.nR) {
    case 2: launch_xg_nr<2>(a, s, nmembers, st); break;
    case 3: launch_xg_nr<3>(a, s, nmembers, st); break;
    case 4: launch_xg_nr<4>(a, s, nmembers, st); break;
    case 5: launch_xg_nr<5>(a, s, nmembers, st); break;
    case 6: launch_xg_nr<6>(a, s, nmembers, st); break;
    case 7: launch_xg_nr<7>(a, s, nmembers, st); break;
    case 8: launch_xg_nr<8>(a, s, nmembers, st); break;
    case 9: launch_xg_nr<9>(a, s, nmembers, st); break;
    case 10: launch_xg_nr<10>(a, s, nmembers, st); break;
    case 11: launch_xg_nr<11>(a, s, nmembers, st); break;
    default: break;
  }
}

}  // namespace gatt

namespace gatt {

// Transposed x solve with its transposed band, one wavefront per line: the
// adjoint of the forward x solve above (and of solveSystemX[per],
// source/Fortran/derivs.F90:556-615, 737-806).  With the fused records the
// forward operator is r' = B'F (rows scaled by the forward factor),
// x1 = D'.r', r'_1 -= A1' x1, r'_{N-1} -= C1' x1, L'd = r', Ux = d, out_0 = x1.
// Its transpose applied to xbar:
//   U^T z = xbar:   z_e   = xbar_e - ce_{e-1} z_{e-1}      (first solved row: no predecessor)
//   L'^T rho = z:   rho_e = z_e - fa'_{e+1} rho_{e+1}
//   periodic:       t = xbar_0 - A1' rho_1 - C1' rho_{N-1};  r'bar_0 = D'_0 t, r'bar_e = rho_e + D'_e t
//   B'^T:           Fbar_j = sum_l B'((j-l) mod N, l) r'bar_{j-l}
// and out = beta*src + (accumulate ? out : 0) + alpha*Fbar.  The line's r'bar
// passes through an LDS buffer of N words per wavefront for the band.
template <int NR, int M>
__global__ __launch_bounds__(256) void line_solve_xgt_kernel(BandTArgs a) {
  extern __shared__ double sh[];
  constexpr int H = NR - 1;
  constexpr int NW = 2 * H + 1;
  const int lane = threadIdx.x & 63;
  const int wv = threadIdx.x >> 6;
  const int line = (int)blockIdx.x * 4 + wv;
  if (line >= a.nlines) return;   // whole wavefronts; only wavefront-level synchronisation below
  const int f = (int)blockIdx.y % a.nf;
  const int mem = (int)blockIdx.y / a.nf;
  const int N = a.N;
  const int per = a.periodic;
  const double* in = a.in[f] + (long long)mem * a.in_stride + (long long)a.nx * line;
  double* out = a.out[f] + (long long)mem * a.out_stride + (long long)a.nx * line;
  const double* src = a.src[f] ? a.src[f] + (long long)mem * a.out_stride + (long long)a.nx * line : nullptr;
  const int tu = __builtin_amdgcn_readfirstlane(a.types[line]);
  // lane-interleaved records of this lane's points s + tt: rows 0..NW-1 the transposed
  // band column, NW the shifted ce(e-1), NW+1 the shifted fa'(e+1), NW+2 D(e)
  const double* ct = a.fcoef_pt + (size_t)tu * (NW + 3) * 64 * M + lane;
  auto cp = [&](int k, int tt) { return ct[((size_t)k * M + tt) * 64]; };
  const int s = lane * M;
  // the line in through this wavefront's LDS buffer (unit-stride loads)
  double* buf = sh + (size_t)wv * N;
  double* obuf = sh + (size_t)(4 + wv) * N;
  {
    // all M loads of the lane in flight before the LDS stores (N <= 64 M)
    double v[M];
#pragma unroll
    for (int k = 0; k < M; ++k) {
      const int e = lane + 64 * k;
      if (e < N) v[k] = ld_nt(in + e);
    }
#pragma unroll
    for (int k = 0; k < M; ++k) {
      const int e = lane + 64 * k;
      if (e < N) buf[e] = v[k];
    }
  }
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  double r[M];
#pragma unroll
  for (int tt = 0; tt < M; ++tt) r[tt] = s + tt < N ? buf[s + tt] : 0.0;

  // U^T z = xbar
  double P = 0.0, Q = 1.0;
#pragma unroll
  for (int tt = 0; tt < M; ++tt) {
    const double cm = cp(NW, tt);
    P = fma(-cm, P, r[tt]);
    Q = -cm * Q;
  }
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const double Pp = xg_shfl_up(P, d), Qp = xg_shfl_up(Q, d);
    if (lane >= d) {
      P = fma(Q, Pp, P);
      Q = Q * Qp;
    }
  }
  double zprev = xg_shfl_up(P, 1);
  if (lane == 0) zprev = 0.0;
#pragma unroll
  for (int tt = 0; tt < M; ++tt) {
    const double cm = cp(NW, tt);
    const double z = fma(-cm, zprev, r[tt]);
    r[tt] = z;
    zprev = z;
  }
  // L'^T rho = z
  P = 0.0;
  Q = 1.0;
#pragma unroll
  for (int tt = M - 1; tt >= 0; --tt) {
    const double fp = cp(NW + 1, tt);
    P = fma(-fp, P, r[tt]);
    Q = -fp * Q;
  }
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const double Pn = xg_shfl_down(P, d), Qn = xg_shfl_down(Q, d);
    if (lane + d < 64) {
      P = fma(Q, Pn, P);
      Q = Q * Qn;
    }
  }
  double rnext = xg_shfl_down(P, 1);
  if (lane == 63) rnext = 0.0;
#pragma unroll
  for (int tt = M - 1; tt >= 0; --tt) {
    const double rho = fma(-cp(NW + 1, tt), rnext, r[tt]);
    r[tt] = rho;
    rnext = rho;
  }
  // transposed periodic correction (row 0 carried xbar_0 through both sweeps unchanged)
  if (per) {
    const double A1 = a.fcoef_ends[(size_t)tu * 2], C1 = a.fcoef_ends[(size_t)tu * 2 + 1];
    double part = 0.0;
#pragma unroll
    for (int tt = 0; tt < M; ++tt) {
      const int e = s + tt;
      if (e == 0) part = part + r[tt];
      if (e == 1) part = part - A1 * r[tt];
      if (e == N - 1) part = part - C1 * r[tt];
    }
#pragma unroll
    for (int m = 1; m < 64; m <<= 1) part += xg_shfl_xor(part, m);
#pragma unroll
    for (int tt = 0; tt < M; ++tt) {
      const int e = s + tt;
      const double De = cp(NW + 2, tt);
      r[tt] = e == 0 ? De * part : fma(De, part, r[tt]);
    }
  }
  // transposed band through the wavefront's line buffer (the input was read into
  // registers before the sweeps, so the buffer can be overwritten)
#pragma unroll
  for (int tt = 0; tt < M; ++tt)
    if (s + tt < N) buf[s + tt] = r[tt];
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
  for (int tt = 0; tt < M; ++tt) {
    const int j = s + tt;
    if (j < N) {
      double acc = 0.0;
#pragma unroll
      for (int l = -H; l <= H; ++l) {
        int se = j - l;
        se = se < 0 ? se + N : (se >= N ? se - N : se);
        acc = fma(cp(H + l, tt), buf[se], acc);
      }
      obuf[j] = a.alpha * acc;
    }
  }
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  // unit-stride combination with src and out (N <= 64 M): every load of the lane in flight
  // before the first store
  double v[M];
#pragma unroll
  for (int k = 0; k < M; ++k) {
    const int j = lane + 64 * k;
    if (j < N) {
      v[k] = obuf[j];
      if (src) v[k] = a.beta * ld_nt(src + j) + v[k];
      if ((a.accumulate >> f) & 1) v[k] = ld_nt(out + j) + v[k];
    }
  }
#pragma unroll
  for (int k = 0; k < M; ++k) {
    const int j = lane + 64 * k;
    if (j < N) st_nt(out + j, v[k]);
  }
}

template <int NR, int M>
static void launch_xgt_m(const BandTArgs& a, int nmembers, hipStream_t st) {
  dim3 grid((unsigned)((a.nlines + 3) / 4), (unsigned)(a.nf * nmembers));
  hipLaunchKernelGGL((line_solve_xgt_kernel<NR, M>), grid, dim3(256), sizeof(double) * 8 * (size_t)a.N, st, a);
}

template <int NR>
static bool launch_xgt_nr(const BandTArgs& a, int nmembers, hipStream_t st) {
  const int M = a.N <= 256 ? 4 : (a.N <= 512 ? 8 : 0);
  if (M == 4) launch_xgt_m<NR, 4>(a, nmembers, st);
  else if (M == 8) launch_xgt_m<NR, 8>(a, nmembers, st);
  else return false;
  return true;
}

inline bool launch_line_solve_xgt_impl(const BandTArgs& a, int nmembers, hipStream_t st) {
  if (a.nR - 1 > a.N) return false;
  switch (a.nR) {
    case 2: return launch_xgt_nr<2>(a, nmembers, st);
    case 3: return launch_xgt_nr<3>(a, nmembers, st);
    case 4: return launch_xgt_nr<4>(a, nmembers, st);
    case 5: return launch_xgt_nr<5>(a, nmembers, st);
    case 6: return launch_xgt_nr<6>(a, nmembers, st);
    case 7: return launch_xgt_nr<7>(a, nmembers, st);
    case 8: return launch_xgt_nr<8>(a, nmembers, st);
    case 9: return launch_xgt_nr<9>(a, nmembers, st);
    case 10: return launch_xgt_nr<10>(a, nmembers, st);
    case 11: return launch_xgt_nr<11>(a, nmembers, st);
    default: return false;
  }
}

}  // namespace gatt
