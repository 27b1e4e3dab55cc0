// Transposed compact line solves: the adjoint of solveSystem{X,Y,Z}[per]
// (source/Fortran/derivs.F90:556-948) without its RHS band, which is applied
// afterwards by band_t_kernel (adjoint.hip).
//
// The forward solve on rows e0..N-1 (e0 = 1 for periodic lines, else 0) is
//   d_e = (r_e - fa_e d_{e-1}) fb_e,    x_e = d_e - ce_e x_{e+1}      (L d = r, U x = d)
// with the packed per-point coefficients fa, fb, ce (line_solve_ws.h).  Its
// transpose applied to xbar is
//   z_e   = xbar_e - ce_{e-1} z_{e-1}                                  (U^T z = xbar)
//   rho_e = (z_e - fa_{e+1} rho_{e+1}) fb_e                            (L^T rho = z)
// For periodic lines the forward step also formed x_0 = sum_e D_e r_e and
// corrected r_1 -= A(1) x_0, r_{N-1} -= C(1) x_0; transposed, with
// t = xbar_0 - A(1) rho_1 - C(1) rho_{N-1}:  rbar_0 = D_0 t and
// rbar_e = rho_e + D_e t.
//
// Layout as the forward wave-segment kernel: lane = line, wavefront w owns
// points [w*M, w*M+M) of 64 lines, segments are chained by composing affine
// maps through LDS; x lines go through an LDS tile.  Each lane reads its own
// line's coefficients, so tiles may mix line types.  The result may overwrite
// the input (every lane reads its points before any store).
#include "gatt_internal.h"
#include "line_solve_ws.h"

namespace gatt {

// UNI: the block's 64 lines share one type, so the coefficient addresses are
// wave-uniform and the records go through the scalar cache.
template <int DIR, int M, bool XLDS, bool UNI>
__global__ __launch_bounds__(1024) void line_solve_t_kernel(SolveArgs a) {
  extern __shared__ double sh[];
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int G = blockDim.x >> 6;
  const int N = a.op.N;
  const int CS = a.cs;
  const int NW = 2 * a.op.nR - 1;
  const int line0 = blockIdx.x * 64;
  const int line = line0 + lane;
  const int nact = min(64, a.nlines - line0);
  const bool active = lane < nact;
  const int f = (int)blockIdx.y % a.nf;
  const int mem = (int)blockIdx.y / a.nf;
  const double* in = a.in[f] + (long long)mem * a.in_stride;
  double* out = a.out[f] + (long long)mem * a.out_stride;
  const long long stride = a.stride;
  const int s = w * M;
  int cnt = N - s;
  cnt = cnt < 0 ? 0 : (cnt > M ? M : cnt);
  double* shP = sh;                 // forward maps   [G][64][2]
  double* shB = sh + 2 * G * 64;    // backward maps  [G][64][2]
  double* shD = sh + 4 * G * 64;    // periodic partials [G][64]
  double* tile = sh;

  const int lsafe = active ? line : line0;
  long long base = 0;
  if (DIR == 1) {
    const int k = lsafe / a.nx;
    base = (long long)(lsafe - k * a.nx) + (long long)a.nx * a.ny * k;
  } else if (DIR == 2) {
    base = lsafe;
  } else if (!XLDS) {
    base = (long long)a.nx * lsafe;
  }
  if (XLDS) {
    const double* src = in + (long long)a.nx * line0;
    const int total = nact * N;
    batched_move<8>(total, [&](int idx) { return src[idx]; }, [&](int idx, double v) {
      const int lp = (int)(((float)idx + 0.5f) * a.inv_n);
      const int ep = idx - lp * N;
      tile[ep * 65 + lp] = v;
    });
    __syncthreads();
  }
  double r[M];
#pragma unroll
  for (int tt = 0; tt < M; ++tt) {
    const int e = min(s + tt, N - 1);
    r[tt] = XLDS ? tile[e * 65 + lane] : in[base + (long long)e * stride];
  }
  if (XLDS) __syncthreads();   // the tile words are reused by the maps

  const int type = UNI ? __builtin_amdgcn_readfirstlane(a.types[line0]) : a.types[lsafe];
  const double* cf = a.coef + (size_t)type * N * CS;   // record of point e at cf + e*CS
  const int per = a.op.periodic;
  const int e0 = per;

  // U^T z = xbar
  double P = 0.0, Q = 1.0;
#pragma unroll
  for (int tt = 0; tt < M; ++tt) {
    const int e = s + tt;
    if (tt < cnt && e >= e0) {
      const double c = e > e0 ? cf[(e - 1) * CS + NW + 2] : 0.0;
      P = r[tt] - c * P;
      Q = -c * Q;
    }
  }
  shP[(w * 64 + lane) * 2] = P;
  shP[(w * 64 + lane) * 2 + 1] = Q;
  __syncthreads();
  double zprev = 0.0;
#pragma unroll
  for (int k = 0; k < kMaxWaveSegments; ++k)
    if (k < w) zprev = shP[(k * 64 + lane) * 2] + shP[(k * 64 + lane) * 2 + 1] * zprev;
#pragma unroll
  for (int tt = 0; tt < M; ++tt) {
    const int e = s + tt;
    if (tt < cnt && e >= e0) {
      const double c = e > e0 ? cf[(e - 1) * CS + NW + 2] : 0.0;
      const double z = r[tt] - c * zprev;
      r[tt] = z;
      zprev = z;
    }
  }

  // L^T rho = z
  P = 0.0;
  Q = 1.0;
#pragma unroll
  for (int tt = M - 1; tt >= 0; --tt) {
    const int e = s + tt;
    if (tt < cnt && e >= e0) {
      const double fa1 = e < N - 1 ? cf[(e + 1) * CS + NW] : 0.0;
      const double fb = cf[e * CS + NW + 1];
      P = (r[tt] - fa1 * P) * fb;
      Q = -(fa1 * fb) * Q;
    }
  }
  shB[(w * 64 + lane) * 2] = P;
  shB[(w * 64 + lane) * 2 + 1] = Q;
  __syncthreads();
  double rnext = 0.0;
#pragma unroll
  for (int k = kMaxWaveSegments - 1; k >= 0; --k)
    if (k > w && k < G) rnext = shB[(k * 64 + lane) * 2] + shB[(k * 64 + lane) * 2 + 1] * rnext;
#pragma unroll
  for (int tt = M - 1; tt >= 0; --tt) {
    const int e = s + tt;
    if (tt < cnt && e >= e0) {
      const double fa1 = e < N - 1 ? cf[(e + 1) * CS + NW] : 0.0;
      const double rho = (r[tt] - fa1 * rnext) * cf[e * CS + NW + 1];
      r[tt] = rho;
      rnext = rho;
    }
  }

  // transposed periodic correction
  if (per) {
    const double A1 = a.coef_ends[(size_t)type * 2], C1 = a.coef_ends[(size_t)type * 2 + 1];
    double part = 0.0;
#pragma unroll
    for (int tt = 0; tt < M; ++tt) {
      const int e = s + tt;
      if (tt < cnt) {
        if (e == 0) part = part + r[tt];
        if (e == 1) part = part - A1 * r[tt];
        if (e == N - 1) part = part - C1 * r[tt];
      }
    }
    shD[w * 64 + lane] = part;
    __syncthreads();
    double t = 0.0;
#pragma unroll
    for (int k = 0; k < kMaxWaveSegments; ++k)
      if (k < G) t += shD[k * 64 + lane];
#pragma unroll
    for (int tt = 0; tt < M; ++tt) {
      const int e = s + tt;
      if (tt < cnt) {
        const double De = cf[e * CS + NW + 3];
        r[tt] = e == 0 ? De * t : r[tt] + De * t;
      }
    }
  }

  if (XLDS) __syncthreads();   // all maps read before the tile is rewritten
#pragma unroll
  for (int tt = 0; tt < M; ++tt) {
    if (tt < cnt) {
      const int e = s + tt;
      if (XLDS) tile[e * 65 + lane] = r[tt];
      else if (active) out[base + (long long)e * stride] = r[tt];
    }
  }
  if (XLDS) {
    __syncthreads();
    double* dst = out + (long long)a.nx * line0;
    const int total = nact * N;
    batched_move<8>(
        total,
        [&](int idx) {
          const int lp = (int)(((float)idx + 0.5f) * a.inv_n);
          const int ep = idx - lp * N;
          return tile[ep * 65 + lp];
        },
        [&](int idx, double v) { dst[idx] = v; });
  }
}

// y and z: the transposed solve followed by the transposed band in one launch.
// The solve is the one above (lane = line, wavefront = segment of M points); its
// result stays in registers, and each segment's first and last H = nR-1 values go
// to LDS, from where the neighbouring segments (circularly, as the reference's
// modulo indexing) read the halo of the band.  The band sums run over l = -H..H in
// the order of band_t_seg_kernel (adjoint.hip).  Every segment holds at least H
// points (checked by the launcher).
template <int DIR, int M, bool UNI, int NR>
__global__ __launch_bounds__(1024) void line_solve_band_t_kernel(SolveArgs a, BandTArgs b) {
  constexpr int H = NR - 1;
  extern __shared__ double sh[];
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int G = blockDim.x >> 6;
  const int N = a.op.N;
  const int CS = a.cs;
  const int NW = 2 * a.op.nR - 1;
  const int line0 = blockIdx.x * 64;
  const int nact = min(64, a.nlines - line0);
  const bool active = lane < nact;
  const int f = (int)blockIdx.y % a.nf;
  const int mem = (int)blockIdx.y / a.nf;
  const double* in = a.in[f] + (long long)mem * a.in_stride;
  const long long stride = a.stride;
  const int s = w * M;
  int cnt = N - s;
  cnt = cnt < 0 ? 0 : (cnt > M ? M : cnt);
  double* shP = sh;                 // forward maps   [G][64][2]
  double* shB = sh + 2 * G * 64;    // backward maps  [G][64][2]
  double* shD = sh + 4 * G * 64;    // periodic partials [G][64]
  double* shH = sh + 5 * G * 64;    // segment ends [G][2H][64]: first H, then last H values

  const int lsafe = active ? line0 + lane : line0;
  long long base;
  if (DIR == 1) {
    const int k = lsafe / a.nx;
    base = (long long)(lsafe - k * a.nx) + (long long)a.nx * a.ny * k;
  } else {
    base = lsafe;
  }
  double r[M];
#pragma unroll
  for (int tt = 0; tt < M; ++tt) {
    const int e = min(s + tt, N - 1);
    r[tt] = ld_nt(in + base + (long long)e * stride);
  }

  const int type = UNI ? __builtin_amdgcn_readfirstlane(a.types[line0]) : a.types[lsafe];
  const double* cf = a.coef + (size_t)type * N * CS;
  const int per = a.op.periodic;
  const int e0 = per;

  // U^T z = xbar
  double P = 0.0, Q = 1.0;
#pragma unroll
  for (int tt = 0; tt < M; ++tt) {
    const int e = s + tt;
    if (tt < cnt && e >= e0) {
      const double c = e > e0 ? cf[(e - 1) * CS + NW + 2] : 0.0;
      P = r[tt] - c * P;
      Q = -c * Q;
    }
  }
  shP[(w * 64 + lane) * 2] = P;
  shP[(w * 64 + lane) * 2 + 1] = Q;
  __syncthreads();
  double zprev = 0.0;
#pragma unroll
  for (int k = 0; k < kMaxWaveSegments; ++k)
    if (k < w) zprev = shP[(k * 64 + lane) * 2] + shP[(k * 64 + lane) * 2 + 1] * zprev;
#pragma unroll
  for (int tt = 0; tt < M; ++tt) {
    const int e = s + tt;
    if (tt < cnt && e >= e0) {
      const double c = e > e0 ? cf[(e - 1) * CS + NW + 2] : 0.0;
      const double z = r[tt] - c * zprev;
      r[tt] = z;
      zprev = z;
    }
  }
  // L^T rho = z
  P = 0.0;
  Q = 1.0;
#pragma unroll
  for (int tt = M - 1; tt >= 0; --tt) {
    const int e = s + tt;
    if (tt < cnt && e >= e0) {
      const double fa1 = e < N - 1 ? cf[(e + 1) * CS + NW] : 0.0;
      const double fb = cf[e * CS + NW + 1];
      P = (r[tt] - fa1 * P) * fb;
      Q = -(fa1 * fb) * Q;
    }
  }
  shB[(w * 64 + lane) * 2] = P;
  shB[(w * 64 + lane) * 2 + 1] = Q;
  __syncthreads();
  double rnext = 0.0;
#pragma unroll
  for (int k = kMaxWaveSegments - 1; k >= 0; --k)
    if (k > w && k < G) rnext = shB[(k * 64 + lane) * 2] + shB[(k * 64 + lane) * 2 + 1] * rnext;
#pragma unroll
  for (int tt = M - 1; tt >= 0; --tt) {
    const int e = s + tt;
    if (tt < cnt && e >= e0) {
      const double fa1 = e < N - 1 ? cf[(e + 1) * CS + NW] : 0.0;
      const double rho = (r[tt] - fa1 * rnext) * cf[e * CS + NW + 1];
      r[tt] = rho;
      rnext = rho;
    }
  }
  if (per) {
    const double A1 = a.coef_ends[(size_t)type * 2], C1 = a.coef_ends[(size_t)type * 2 + 1];
    double part = 0.0;
#pragma unroll
    for (int tt = 0; tt < M; ++tt) {
      const int e = s + tt;
      if (tt < cnt) {
        if (e == 0) part = part + r[tt];
        if (e == 1) part = part - A1 * r[tt];
        if (e == N - 1) part = part - C1 * r[tt];
      }
    }
    shD[w * 64 + lane] = part;
    __syncthreads();
    double t = 0.0;
#pragma unroll
    for (int k = 0; k < kMaxWaveSegments; ++k)
      if (k < G) t += shD[k * 64 + lane];
#pragma unroll
    for (int tt = 0; tt < M; ++tt) {
      const int e = s + tt;
      if (tt < cnt) {
        const double De = cf[e * CS + NW + 3];
        r[tt] = e == 0 ? De * t : r[tt] + De * t;
      }
    }
  }

  // segment ends for the neighbours' halos (cnt >= H on every segment)
#pragma unroll
  for (int h = 0; h < H; ++h) {
    double last = r[0];
#pragma unroll
    for (int tt = 0; tt < M; ++tt)
      if (tt == cnt - H + h) last = r[tt];
    shH[(w * 2 * H + h) * 64 + lane] = r[h];
    shH[(w * 2 * H + H + h) * 64 + lane] = last;
  }
  __syncthreads();
  if (!active || cnt == 0) return;
  const int wl = w == 0 ? G - 1 : w - 1;
  const int wr = w == G - 1 ? 0 : w + 1;
  const double* Rt = b.R + (size_t)type * N * (2 * NR - 1);
  double* out = b.out[f] + (long long)mem * b.out_stride;
  const double* src = b.src[f] ? b.src[f] + (long long)mem * b.out_stride : nullptr;
#pragma unroll
  for (int tt = 0; tt < M; ++tt) {
    if (tt < cnt) {
      const int e = s + tt;
      double acc = 0.0;
#pragma unroll
      for (int l = -H; l <= H; ++l) {
        const int u = tt - l;           // position of point e - l relative to the segment start
        double v;
        if (u < 0) v = shH[(wl * 2 * H + H + (H + u)) * 64 + lane];
        else if (u >= cnt) v = shH[(wr * 2 * H + (u - cnt)) * 64 + lane];
        else {
          v = r[0];
#pragma unroll
          for (int q = 0; q < M; ++q)
            if (q == u) v = r[q];
        }
        int se = e - l;
        se = se < 0 ? se + N : (se >= N ? se - N : se);
        acc = acc + Rt[(size_t)(H + l) * N + se] * v;
      }
      const long long p = base + (long long)e * stride;
      double v = b.alpha * acc;
      if (src) v = b.beta * ld_nt(src + p) + v;
      if ((b.accumulate >> f) & 1) v = ld_nt(out + p) + v;
      st_nt(out + p, v);
    }
  }
}

size_t line_solve_band_t_lds_bytes(int N, int M, int nR) {
  const int G = (N + M - 1) / M;
  return sizeof(double) * (size_t)(5 + 2 * (nR - 1)) * G * 64;
}

template <int DIR, int M, int NR>
static void launch_band_t_m(const SolveArgs& a0, const BandTArgs& b, int nmembers, hipStream_t st) {
  SolveArgs a = a0;
  const int G = (a.op.N + M - 1) / M;
  const long long tiles = (a.nlines + 63) / 64;
  const size_t lds = line_solve_band_t_lds_bytes(a.op.N, M, NR);
  dim3 grid((unsigned)tiles, (unsigned)(a.nf * nmembers));
  auto k = a.uniform ? line_solve_band_t_kernel<DIR, M, true, NR> : line_solve_band_t_kernel<DIR, M, false, NR>;
  if (lds > 65536) (void)hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  hipLaunchKernelGGL(k, grid, dim3(64 * G), lds, st, a, b);
}

template <int DIR, int M>
static bool launch_band_t_nr(const SolveArgs& a, const BandTArgs& b, int nmembers, hipStream_t st) {
  switch (a.op.nR) {
    case 2: launch_band_t_m<DIR, M, 2>(a, b, nmembers, st); return true;
    case 3: launch_band_t_m<DIR, M, 3>(a, b, nmembers, st); return true;
    case 4: launch_band_t_m<DIR, M, 4>(a, b, nmembers, st); return true;
    case 5: launch_band_t_m<DIR, M, 5>(a, b, nmembers, st); return true;
    case 6: launch_band_t_m<DIR, M, 6>(a, b, nmembers, st); return true;
  }
  return false;
}

bool line_solve_band_t_fits(int dir, int N, int nR, int nf, int nmembers, long long max_lds) {
  if (dir != 1 && dir != 2) return false;
  const int H = nR - 1;
  const int M = line_solve_t_points_per_lane(N);
  if (M == 0 || H < 1 || H > 5 || H > M) return false;
  const int G = (N + M - 1) / M;
  if (N - (G - 1) * M < H) return false;                  // the last segment must hold a halo
  if ((long long)line_solve_band_t_lds_bytes(N, M, nR) > max_lds) return false;
  return (long long)nf * nmembers <= 65535;
}

bool launch_line_solve_band_t(const SolveArgs& a, const BandTArgs& b, int nmembers, long long max_lds,
                              hipStream_t st) {
  if (!line_solve_band_t_fits(a.dir, a.op.N, a.op.nR, a.nf, nmembers, max_lds)) return false;
  const int M = line_solve_t_points_per_lane(a.op.N);
  if (a.dir == 1) return M == 16 ? launch_band_t_nr<1, 16>(a, b, nmembers, st) : launch_band_t_nr<1, 32>(a, b, nmembers, st);
  return M == 16 ? launch_band_t_nr<2, 16>(a, b, nmembers, st) : launch_band_t_nr<2, 32>(a, b, nmembers, st);
}

// x lines are staged through an LDS tile of 65*N words when it fits in 160 KiB
// of LDS; longer x lines are read in place (strided across lanes).
constexpr int kMaxTileN = 300;

size_t line_solve_t_lds_bytes(int dir, int N, int M) {
  const int G = (N + M - 1) / M;
  const size_t maps = (size_t)5 * G * 64;
  const size_t t = (dir == 0 && N <= kMaxTileN) ? (size_t)65 * N : 0;
  return sizeof(double) * (maps > t ? maps : t);
}

template <int DIR, int M>
static void launch_t_m(const SolveArgs& a0, int nmembers, hipStream_t st) {
  SolveArgs a = a0;
  a.inv_n = 1.0f / (float)a.op.N;
  const int G = (a.op.N + M - 1) / M;
  const long long tiles = (a.nlines + 63) / 64;
  const size_t lds = line_solve_t_lds_bytes(DIR, a.op.N, M);
  dim3 grid((unsigned)tiles, (unsigned)(a.nf * nmembers));
  const bool tile = DIR == 0 && a.op.N <= kMaxTileN;
  auto k = a.uniform ? (tile ? line_solve_t_kernel<DIR, M, DIR == 0, true> : line_solve_t_kernel<DIR, M, false, true>)
                     : (tile ? line_solve_t_kernel<DIR, M, DIR == 0, false> : line_solve_t_kernel<DIR, M, false, false>);
  if (lds > 65536) (void)hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  hipLaunchKernelGGL(k, grid, dim3(64 * G), lds, st, a);
}

// M = 16 for lines up to 256 points, 32 up to 512 (at most 16 wavefronts).
int line_solve_t_points_per_lane(int N) { return N <= 256 ? 16 : (N <= 512 ? 32 : 0); }

void launch_line_solve_t(const SolveArgs& a, int nmembers, hipStream_t st) {
  const int M = line_solve_t_points_per_lane(a.op.N);
  if (a.dir == 0) {
    if (M == 16) launch_t_m<0, 16>(a, nmembers, st);
    else launch_t_m<0, 32>(a, nmembers, st);
  } else if (a.dir == 1) {
    if (M == 16) launch_t_m<1, 16>(a, nmembers, st);
    else launch_t_m<1, 32>(a, nmembers, st);
  } else {
    if (M == 16) launch_t_m<2, 16>(a, nmembers, st);
    else launch_t_m<2, 32>(a, nmembers, st);
  }
}

}  // namespace gatt
