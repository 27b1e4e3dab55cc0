// Fused wave-segment line solves.
//
// Same algorithm as line_solve_ws.h (band product, periodic first-row dot
// product and corrections, segmented Thomas sweeps chained by an LDS scan of
// affine maps), restating solveSystem{X,Y,Z}[per]
// (source/Fortran/derivs.F90:556-948), with a leaner instruction stream:
//
//   * A workgroup owns 64 lines; lane l of every wavefront works on line
//     line0 + l and wavefront w on the points [w*M, w*M+M) of all of them.
//   * Per-point coefficient records are pre-multiplied on the host so that the
//     band product and every sweep update are single fused multiply-adds.
//   * No point carries a branch: coefficients past the line end are zero, the
//     forward factor of the first solved row and the backward factor of the
//     last row are zero, so every point runs the same arithmetic.
//   * Scan maps keep P and Q in separate LDS rows of 64 lanes (conflict-free).
//   * Along y and z every point of a segment is one 512-byte row; along x each
//     lane reads and writes its own contiguous stretch of its line (pairs of
//     doubles where aligned), so no LDS tile is needed.
//   * Results differ from the reference's sequential recurrence by rounding
//     only (the fused operations and the scan's composite maps).
#pragma once
#include "gatt_internal.h"

namespace gatt {

// Fused per-point coefficient record of this kernel (built by the host from the
// reference's A, B, C, D, R):  the band row scaled by the forward factor,
// c'_q = R_q B, then fa' = A B, ce = C, D' = D / B (D at the periodic row 0,
// whose band row is kept unscaled), padded to an even count.  With it the
// forward sweep d_e = (r_e - A d_{e-1}) B becomes d_e = r'_e - fa' d_{e-1},
// and every update is one fused multiply-add.  The two periodic end factors
// are A(1) B and C(1) B of their rows.
constexpr int wsp_coef_stride(int nR) { return (2 * nR - 1 + 3 + 1) & ~1; }
constexpr int kMaxSegments = 16;   // wavefronts (segments) per line

// Compiler-only fence between the phases of a sweep: the coefficients are read
// from LDS again in each phase instead of being held in registers across it.
__device__ __forceinline__ void wsp_phase() { asm volatile("" ::: "memory"); }

// Sweeps of one segment: r holds the scaled band product on entry and the
// solution on exit; returns the periodic first-row value.  Every point of the
// segment runs through the same arithmetic: the packed forward factor is zero
// at the first solved row (and at the periodic row 0 before it) and the
// backward factor is zero at the last row, so no point needs a branch.
// The caller is segment sg (of G) of line l (of the L lines of the workgroup);
// the maps of segment k of line l sit at [k*L + l].  KMAX bounds G.
// ROLLED composes the maps of earlier segments in a loop over their count instead of a
// fully unrolled one (the same operations in the same order; fewer registers in flight).
template <int NR, int M, int KMAX = kMaxSegments, bool ROLLED = false>
__device__ __forceinline__ double wsp_sweeps(double (&r)[M], const double* cf, const double* cends, double* shD,
                                             double* shP, double* shB, int cnt, int sg, int G, int l, int per,
                                             int L = 64, int GA = 0) {
  // GA: segments the map arrays hold (at least G)
  if (GA < G) GA = G;
  double* shQ = shP + GA * L;
  double* shC = shB + GA * L;
  constexpr int H = NR - 1;
  constexpr int NW = 2 * H + 1;
  constexpr int CS = wsp_coef_stride(NR);
  const int me = sg * L + l;
  double x1 = 0.0;
  if (per) {
    double part = 0.0;
#pragma unroll
    for (int tt = 0; tt < M; ++tt) part = fma(r[tt], cf[tt * CS + NW + 2], part);
    shD[me] = part;
    __syncthreads();
    wsp_phase();
    if constexpr (ROLLED) {
#pragma unroll 4
      for (int k = 0; k < G; ++k) x1 += shD[k * L + l];
    } else {
#pragma unroll
      for (int k = 0; k < KMAX; ++k)
        if (k < G) x1 += shD[k * L + l];
    }
    // row 2 and row N corrections (derivs.F90:792-793)
    if (sg == 0) r[1] = fma(-cends[0], x1, r[1]);
    if (sg == G - 1) {
#pragma unroll
      for (int tt = 0; tt < M; ++tt)
        if (tt == cnt - 1) r[tt] = fma(-cends[1], x1, r[tt]);
    }
  }

  // forward sweep: composite affine map of the segment, scan, re-run
  double P = 0.0, Q = 1.0;
#pragma unroll
  for (int tt = 0; tt < M; ++tt) {
    const double fa = cf[tt * CS + NW];
    P = fma(-fa, P, r[tt]);
    Q = -fa * Q;
  }
  shP[me] = P;
  shQ[me] = Q;
  __syncthreads();
  wsp_phase();
  double dprev = 0.0;
  if constexpr (ROLLED) {
#pragma unroll 4
    for (int k = 0; k < sg; ++k) dprev = fma(shQ[k * L + l], dprev, shP[k * L + l]);
  } else {
#pragma unroll
    for (int k = 0; k < KMAX; ++k)
      if (k < sg) dprev = fma(shQ[k * L + l], dprev, shP[k * L + l]);
  }
#pragma unroll
  for (int tt = 0; tt < M; ++tt) {
    const double d = fma(-cf[tt * CS + NW], dprev, r[tt]);
    r[tt] = d;
    dprev = d;
  }

  // backward sweep
  wsp_phase();
  P = 0.0;
  Q = 1.0;
#pragma unroll
  for (int tt = M - 1; tt >= 0; --tt) {
    const double ce = cf[tt * CS + NW + 1];
    P = fma(-ce, P, r[tt]);
    Q = -ce * Q;
  }
  shB[me] = P;
  shC[me] = Q;
  __syncthreads();
  wsp_phase();
  double xnext = 0.0;
  if constexpr (ROLLED) {
#pragma unroll 4
    for (int k = G - 1; k > sg; --k) xnext = fma(shC[k * L + l], xnext, shB[k * L + l]);
  } else {
#pragma unroll
    for (int k = KMAX - 1; k >= 0; --k)
      if (k > sg && k < G) xnext = fma(shC[k * L + l], xnext, shB[k * L + l]);
  }
#pragma unroll
  for (int tt = M - 1; tt >= 0; --tt) {
    const double x = fma(-cf[tt * CS + NW + 1], xnext, r[tt]);
    r[tt] = x;
    xnext = x;
  }
  return x1;
}

template <int NR, int M, bool BLEND>
__device__ __forceinline__ void wsp_band(double (&r)[M], double (&f0)[BLEND ? M : 1],
                                         const double (&fw)[M + 2 * (NR - 1)], const double* cf) {
  constexpr int H = NR - 1;
  constexpr int CS = wsp_coef_stride(NR);
#pragma unroll
  for (int tt = 0; tt < M; ++tt) {
    const double* c = cf + tt * CS;
    double acc = c[H] * fw[tt + H];
#pragma unroll
    for (int q = 1; q <= H; ++q) acc = fma(c[H - q], fw[tt + H - q], fma(c[H + q], fw[tt + H + q], acc));
    r[tt] = acc;
  
    if (BLEND) f0[tt] = fw[tt + H];
  }
}

// Flux-mode input of member mem at point n (SolveArgs::has_flux): viscous flux q
// computed from the stage state's E and the first derivatives.
__device__ __forceinline__ double wsp_flux_at(const SolveArgs& a, int q, int mem, long long n) {
  const long long np = (long long)a.nx * a.ny * a.nz;
  const double* E = a.flux_E + (long long)mem * a.in_stride;
  const double* D = a.flux_D + (long long)mem * a.flux_d_stride;
  return visc_flux(q, a.fc, E[n], D, n, np, a.is3d);
}

template <int DIR>
__device__ __forceinline__ long long wsp_base(const SolveArgs& a, int line) {
  if (DIR == 1) {
    const int k = line / a.nx;
    return (long long)(line - k * a.nx) + (long long)a.nx * a.ny * k;
  }
  return line;   // DIR 2 (x uses the contiguous stretch)
}

// x: lane l holds the window of line line0 + l, points st .. st+W-1 (wrapped),
// read straight from its contiguous line.  Where the window does not wrap and
// the line is 16-byte aligned, it is read as pairs of doubles.
template <int W>
__device__ __forceinline__ void wsp_load_x(const double* src, int N, int st, bool vec, double (&fw)[W]) {
  constexpr int NP = (W + 2) / 2;
  const int off = st & 1;
  const int a0 = st - off;
  if (vec && a0 >= 0 && a0 + 2 * NP <= N) {
    double t[2 * NP];
    const double2* p = reinterpret_cast<const double2*>(src + a0);
#pragma unroll
    for (int u = 0; u < NP; ++u) {
      const double2 v = p[u];
      t[2 * u] = v.x;
      t[2 * u + 1] = v.y;
    }
#pragma unroll
    for (int u = 0; u < W; ++u) fw[u] = off ? t[u + 1] : t[u];
  } else {
#pragma unroll
    for (int u = 0; u < W; ++u) {
      int e = st + u;
      e = e < 0 ? e + N : (e >= N ? e - N : e);
      e = e >= N ? N - 1 : e;
      fw[u] = src[e];
    }
  }
}

// Two workgroups per compute unit where the registers allow it (64 lines x 256
// points: 16 wavefronts of 16 points, or 8 of 32).  Flux-mode loads evaluate the
// viscous flux at every window point and get 128 registers (one workgroup).
template <int NR, int M, bool BLEND, bool FLUX = false>
constexpr int wsp_min_waves() { return M >= 32 ? 4 : (FLUX ? 4 : ((NR <= 4 && !BLEND) ? 8 : 1)); }

template <int DIR, int NR, int M, bool BLEND, bool FLUX = false>
__device__ __forceinline__ void line_solve_wsp_body(const SolveArgs& a, BlockPos bp) {
  extern __shared__ double sh[];
  constexpr int H = NR - 1;
  constexpr int CS = wsp_coef_stride(NR);
  constexpr int W = M + 2 * H;
  constexpr int SLOT = M * CS + 2;             // coefficient words of one segment + the two end factors
  constexpr int CP = (SLOT + 63) / 64;          // coefficient words per lane
  constexpr bool XL = DIR == 0;
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int G = blockDim.x >> 6;
  const int N = a.op.N;
  const int per = a.op.periodic;
  const int s = w * M;
  int cnt = N - s;
  cnt = cnt < 0 ? 0 : (cnt > M ? M : cnt);
  double* shD = sh;                              // periodic dot partials [G][64]
  double* shP = sh + G * 64;                     // forward maps  [2][G][64]
  double* shB = sh + 3 * G * 64;                 // backward maps [2][G][64]
  double* cl = sh + 5 * G * 64 + w * SLOT;       // this wavefront's coefficients
  const long long stride = a.stride;
  const int line0 = bp.x * 64;
  const int nact = min(64, a.nlines - line0);
  const int f = (int)bp.y % a.nf;
  const int mem = (int)bp.y / a.nf;
  const double* in = a.in[f] + (long long)mem * a.in_stride;
  double* out = a.out[f] + (long long)mem * a.out_stride;
  const int line = lane < nact ? line0 + lane : line0;

  // the segment's coefficients, in flight with the field loads; points past the
  // line end get zeros, which make their band rows, forward factors and
  // backward factors vanish
  double cpre[CP];
  {
    const int tu = __builtin_amdgcn_readfirstlane(a.types[line0]);
    const double* src = a.fcoef + ((size_t)tu * N + s) * CS;
    const double* ends = a.fcoef_ends + (size_t)tu * 2;
#pragma unroll
    for (int k = 0; k < CP; ++k) {
      const int jj = lane + 64 * k;
      double v = 0.0;
      if (jj < cnt * CS) v = src[jj];
      else if (jj >= M * CS && jj < SLOT) v = ends[jj - M * CS];
      cpre[k] = v;
    }
  }

  double fw[W];
  if (FLUX) {
    const int q = a.flux_q[f];
    const long long base = XL ? (long long)a.nx * line : wsp_base<DIR>(a, line);
    const long long st = XL ? 1 : stride;
#pragma unroll
    for (int u = 0; u < W; ++u) {
      int e = s - H + u;
      e = e < 0 ? e + N : (e >= N ? e - N : e);
      e = e >= N ? N - 1 : e;
      fw[u] = wsp_flux_at(a, q, mem, base + (long long)e * st);
    }
  } else if (XL) {
    wsp_load_x<W>(in + (long long)a.nx * line, N, s - H, a.xvec != 0, fw);
  } else {
    const long long base = wsp_base<DIR>(a, line);
#pragma unroll
    for (int u = 0; u < W; ++u) {
      int e = s - H + u;
      e = e < 0 ? e + N : (e >= N ? e - N : e);
      e = e >= N ? N - 1 : e;
      fw[u] = in[base + (long long)e * stride];   // neighbouring segments re-read the halo rows
    }
  }
  // only this wavefront reads its slot, and its LDS accesses run in order
#pragma unroll
  for (int k = 0; k < CP; ++k) {
    const int jj = lane + 64 * k;
    if (jj < SLOT) cl[jj] = cpre[k];
  }
  __builtin_amdgcn_wave_barrier();
  const double* cf = cl;
  const double* cends = cl + M * CS;
  double r[M];
  double f0[BLEND ? M : 1];
  wsp_band<NR, M, BLEND>(r, f0, fw, cf);
  const double x1 = wsp_sweeps<NR, M>(r, cf, cends, shD, shP, shB, cnt, w, G, lane, per);
  if (per && w == 0) r[0] = x1;
  if (BLEND) {
#pragma unroll
    for (int tt = 0; tt < M; ++tt) r[tt] = fma(a.theta, r[tt] - f0[tt], f0[tt]);
  }
  const bool full = cnt == M && nact == 64;
  if (XL) {
    double* dst = out + (long long)a.nx * line + s;
    if (full && a.xvec) {
      double2* p = reinterpret_cast<double2*>(dst);
#pragma unroll
      for (int u = 0; u < M / 2; ++u) p[u] = make_double2(r[2 * u], r[2 * u + 1]);
    } else if (lane < nact) {
#pragma unroll
      for (int tt = 0; tt < M; ++tt)
        if (tt < cnt) dst[tt] = r[tt];
    }
  } else {
    const long long base = wsp_base<DIR>(a, line) + (long long)s * stride;
    if (full) {
#pragma unroll
      for (int tt = 0; tt < M; ++tt) st_nt(out + base + (long long)tt * stride, r[tt]);
    } else if (lane < nact) {
#pragma unroll
      for (int tt = 0; tt < M; ++tt)
        if (tt < cnt) st_nt(out + base + (long long)tt * stride, r[tt]);
    }
  }
}

template <int DIR, int NR, int M, bool BLEND, bool FLUX = false>
__global__ __launch_bounds__(M >= 32 ? 512 : 1024, (wsp_min_waves<NR, M, BLEND, FLUX>())) void line_solve_wsp_kernel(SolveArgs a) {
  line_solve_wsp_body<DIR, NR, M, BLEND, FLUX>(a, grid_pos());
}
template <int DIR, int NR, int M, bool BLEND>
__global__ __launch_bounds__(M >= 32 ? 512 : 1024, (wsp_min_waves<NR, M, BLEND, false>())) void line_solve_wsp_ragged(
    const SolveArgs* __restrict__ args, RaggedMap map) {
  BlockPos bp;
  const int m = ragged_pos(map, bp);
  line_solve_wsp_body<DIR, NR, M, BLEND, false>(args[m], bp);
}

// Variant 3, x lines through an LDS tile: a workgroup owns LW (16 or 32) lines,
// one contiguous stretch of LW*N doubles moved between HBM and the tile (N rows
// of LW words) with unit-stride accesses.  Lane l of wavefront w works on line
// l % LW and segment (64/LW)*w + l/LW; the scan maps and the coefficient slots
// (one per segment) reuse the tile's words between the window reads and the
// result writes.  At N = 256 a 32-line tile takes 64 KiB (two workgroups per
// compute unit), a 16-line tile 32 KiB (four).
template <int NR, int M, bool BLEND, int LW, bool FLUX = false>
__device__ __forceinline__ void line_solve_wspx_body(const SolveArgs& a, BlockPos bp) {
  extern __shared__ double sh[];
  constexpr int H = NR - 1;
  constexpr int CS = wsp_coef_stride(NR);
  constexpr int W = M + 2 * H;
  constexpr int SLOT = M * CS + 2;
  constexpr int SPW = 64 / LW;                 // segments per wavefront
  constexpr int CP = (SLOT + LW - 1) / LW;      // coefficient words per lane (LW lanes per segment)
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int l = lane & (LW - 1);
  const int sg = SPW * w + lane / LW;
  const int N = a.op.N;
  const int G = (N + M - 1) / M;               // segments per line
  const int GA = (G + SPW - 1) / SPW * SPW;     // segments of the wavefronts (the last may lie past the line)
  const int per = a.op.periodic;
  const int s = sg * M;
  int cnt = N - s;
  cnt = cnt < 0 ? 0 : (cnt > M ? M : cnt);
  double* tile = sh;
  // tile word of point e of line l: rows of LW words, the line index XOR-swizzled by the row
  // so that the pair writes and reads (one line, consecutive points) spread over the banks
  auto tix = [](int e, int ll) { return e * LW + (ll ^ ((e >> 1) & (LW - 1))); };
  double* shD = sh;                            // periodic dot partials [GA][LW]
  double* shP = sh + GA * LW;                  // forward maps  [2][GA][LW]
  double* shB = sh + 3 * GA * LW;              // backward maps [2][GA][LW]
  double* cl = sh + 5 * GA * LW + sg * SLOT;   // this segment's coefficients
  const int line0 = bp.x * LW;
  const int nact = min(LW, a.nlines - line0);
  const int f = (int)bp.y % a.nf;
  const int mem = (int)bp.y / a.nf;
  const double* in = a.in[f] + (long long)mem * a.in_stride + (long long)a.nx * line0;
  double* out = a.out[f] + (long long)mem * a.out_stride + (long long)a.nx * line0;

  double cpre[CP];
  {
    const int tu = __builtin_amdgcn_readfirstlane(a.types[line0]);
    const double* src = a.fcoef + ((size_t)tu * N + (sg < G ? s : 0)) * CS;
    const double* ends = a.fcoef_ends + (size_t)tu * 2;
#pragma unroll
    for (int k = 0; k < CP; ++k) {
      const int jj = l + LW * k;
      double v = 0.0;
      if (jj < cnt * CS) v = src[jj];
      else if (jj >= M * CS && jj < SLOT) v = ends[jj - M * CS];
      cpre[k] = v;
    }
  }
  const int total = nact * N;
  // pairs of doubles where the stretch is 16-byte aligned and a pair never straddles two lines
  const bool pairs = a.xvec && (N & 1) == 0;
  if (FLUX) {
    // the tile holds the viscous flux computed from E and the first derivatives
    const int q = a.flux_q[f];
    const long long p0 = (long long)a.nx * line0;
    for (int idx = threadIdx.x; idx < total; idx += blockDim.x) {
      const int lp = (int)(((float)idx + 0.5f) * a.inv_n);
      const int ep = idx - lp * N;
      tile[tix(ep, lp)] = wsp_flux_at(a, q, mem, p0 + idx);
    }
  } else if (pairs) {
    // batches of XB pairs per thread: every load of a batch is issued before the first
    // LDS store waits for its data
    constexpr int XB = 8;
    const int npair = total / 2;
    const int bs = blockDim.x;
    for (int b0 = threadIdx.x; b0 < npair; b0 += XB * bs) {
      double2 v[XB];
#pragma unroll
      for (int k = 0; k < XB; ++k) {
        const int i2 = b0 + k * bs;
        if (i2 < npair) {
          const nt_double2 t = ld_nt2(in + 2 * i2);
          v[k] = make_double2(t.x, t.y);
        }
      }
#pragma unroll
      for (int k = 0; k < XB; ++k) {
        const int i2 = b0 + k * bs;
        if (i2 < npair) {
          const int idx = 2 * i2;
          const int lp = (int)(((float)idx + 0.5f) * a.inv_n);
          const int ep = idx - lp * N;
          tile[tix(ep, lp)] = v[k].x;
          tile[tix(ep + 1, lp)] = v[k].y;
        }
      }
    }
  } else {
    batched_move<8>(total, [&](int idx) { return in[idx]; }, [&](int idx, double v) {
      const int lp = (int)(((float)idx + 0.5f) * a.inv_n);
      const int ep = idx - lp * N;
      tile[tix(ep, lp)] = v;
    });
  }
  __syncthreads();
  double fw[W];
#pragma unroll
  for (int u = 0; u < W; ++u) {
    int e = s - H + u;
    e = e < 0 ? e + N : (e >= N ? e - N : e);
    e = e >= N ? e - N : e;
    e = e < 0 ? 0 : (e >= N ? N - 1 : e);
    fw[u] = tile[tix(e, l)];
  }
  __syncthreads();   // the tile's words now hold the maps and the coefficient slots
  // a segment's slot is written and read by the same wavefront, whose LDS accesses run in order
#pragma unroll
  for (int k = 0; k < CP; ++k) {
    const int jj = l + LW * k;
    if (jj < SLOT) cl[jj] = cpre[k];
  }
  __builtin_amdgcn_wave_barrier();
  const double* cf = cl;
  const double* cends = cl + M * CS;
  double r[M];
  double f0[BLEND ? M : 1];
  wsp_band<NR, M, BLEND>(r, f0, fw, cf);
  const double x1 = wsp_sweeps<NR, M, 2 * kMaxSegments>(r, cf, cends, shD, shP, shB, cnt, sg, G, l, per, LW, GA);
  if (per && sg == 0) r[0] = x1;
  if (BLEND) {
#pragma unroll
    for (int tt = 0; tt < M; ++tt) r[tt] = fma(a.theta, r[tt] - f0[tt], f0[tt]);
  }
  __syncthreads();   // every map and slot read before the tile is rewritten
#pragma unroll
  for (int tt = 0; tt < M; ++tt)
    if (tt < cnt) tile[tix(s + tt, l)] = r[tt];
  __syncthreads();
  if (pairs) {
    constexpr int XB = 8;
    const int npair = total / 2;
    const int bs = blockDim.x;
    for (int b0 = threadIdx.x; b0 < npair; b0 += XB * bs) {
      double2 v[XB];
#pragma unroll
      for (int k = 0; k < XB; ++k) {
        const int i2 = b0 + k * bs;
        if (i2 < npair) {
          const int idx = 2 * i2;
          const int lp = (int)(((float)idx + 0.5f) * a.inv_n);
          const int ep = idx - lp * N;
          v[k] = make_double2(tile[tix(ep, lp)], tile[tix(ep + 1, lp)]);
        }
      }
#pragma unroll
      for (int k = 0; k < XB; ++k) {
        const int i2 = b0 + k * bs;
        if (i2 < npair) st_nt2(out + 2 * i2, v[k].x, v[k].y);
      }
    }
  } else {
    batched_move<8>(
        total,
        [&](int idx) {
          const int lp = (int)(((float)idx + 0.5f) * a.inv_n);
          const int ep = idx - lp * N;
          return tile[tix(ep, lp)];
        },
        [&](int idx, double v) { out[idx] = v; });
  }
}

template <int NR, int M, bool BLEND, int LW, bool FLUX = false>
__global__ __launch_bounds__(512) void line_solve_wspx_kernel(SolveArgs a) {
  line_solve_wspx_body<NR, M, BLEND, LW, FLUX>(a, grid_pos());
}
template <int NR, int M, bool BLEND, int LW>
__global__ __launch_bounds__(512) void line_solve_wspx_ragged(const SolveArgs* __restrict__ args, RaggedMap map) {
  BlockPos bp;
  const int m = ragged_pos(map, bp);
  line_solve_wspx_body<NR, M, BLEND, LW, false>(args[m], bp);
}

// LDS words: scan maps (5*G*64), then one coefficient slot per wavefront.
inline size_t wsp_lds_bytes(int G, int M, int nR) {
  return sizeof(double) * ((size_t)5 * G * 64 + (size_t)G * (M * wsp_coef_stride(nR) + 2));
}
// Variant 3: the tile (lw*N words), which holds the maps and slots in turn.
inline size_t wspx_lds_bytes(int N, int M, int nR, int lw) {
  const int spw = 64 / lw;
  const int G = ((N + M - 1) / M + spw - 1) / spw * spw;
  const size_t work = (size_t)5 * G * lw + (size_t)G * (M * wsp_coef_stride(nR) + 2);
  const size_t tile = (size_t)lw * N;
  return sizeof(double) * (work > tile ? work : tile);
}

template <int DIR, int NR, int M, bool BLEND, bool FLUX = false>
static void launch_wsp_k(const SolveArgs& a, int G, hipStream_t st) {
  static_assert(M % 2 == 0, "pairs of points per lane");
  auto k = line_solve_wsp_kernel<DIR, NR, M, BLEND, FLUX>;
  const size_t lds = wsp_lds_bytes(G, M, NR);
  dim3 grid((unsigned)((a.nlines + 63) / 64), (unsigned)(a.nf * a.nmem));
  if constexpr (FLUX) {
    if (record_unsupported()) return;
  } else {
    if (record_launch((const void*)line_solve_wsp_ragged<DIR, NR, M, BLEND>, a, grid, dim3(64 * G), lds)) return;
  }
  if (lds > 65536) (void)hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  hipLaunchKernelGGL(k, grid, dim3(64 * G), lds, st, a);
}

template <int NR, int M, bool BLEND, int LW, bool FLUX = false>
static void launch_wspx_k(const SolveArgs& a, hipStream_t st) {
  auto k = line_solve_wspx_kernel<NR, M, BLEND, LW, FLUX>;
  const size_t lds = wspx_lds_bytes(a.op.N, M, NR, LW);
  constexpr int SPW = 64 / LW;
  const int G = (a.op.N + M - 1) / M;
  dim3 grid((unsigned)((a.nlines + LW - 1) / LW), (unsigned)(a.nf * a.nmem));
  if constexpr (FLUX) {
    if (record_unsupported()) return;
  } else {
    if (record_launch((const void*)line_solve_wspx_ragged<NR, M, BLEND, LW>, a, grid, dim3(64 * ((G + SPW - 1) / SPW)),
                      lds))
      return;
  }
  if (lds > 65536) (void)hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  hipLaunchKernelGGL(k, grid, dim3(64 * ((G + SPW - 1) / SPW)), lds, st, a);
}

template <int DIR, int NR, int M>
static void launch_wsp_m(const SolveArgs& a0, int nmembers, hipStream_t st) {
  SolveArgs a = a0;
  a.inv_n = 1.0f / (float)a.op.N;
  a.nmem = nmembers;
  // x lines are read and written as pairs of doubles when every line starts 16-byte aligned
  bool vec = DIR == 0 && a.nx % 2 == 0 && a.in_stride % 2 == 0 && a.out_stride % 2 == 0;
  for (int f = 0; f < a.nf && vec; ++f)
    vec = ((uintptr_t)a.in[f] % 16 == 0) && ((uintptr_t)a.out[f] % 16 == 0);
  a.xvec = vec ? 1 : 0;
  // segments (wavefronts) per line: enough for the line, or the shape's count when that is larger
  // (the members of a ragged group share one workgroup size; segments past the line end are empty)
  const int G = std::max((a.op.N + M - 1) / M, a.xtile ? 0 : a.G);
  const bool blend = a.mode == 1 && a.theta > 0.0;
  if constexpr (NR <= kFluxMaxNR) if (a.has_flux) {
    // derivatives of the viscous fluxes (mode 0; derivative bands only, the host checks nR)
    if constexpr (DIR == 0) {
      if (a.xtile == 16) {
        launch_wspx_k<NR, M, false, 16, true>(a, st);
        return;
      }
      if (a.xtile) {
        launch_wspx_k<NR, M, false, 32, true>(a, st);
        return;
      }
    }
    launch_wsp_k<DIR, NR, M, false, true>(a, G, st);
    return;
  }
  if constexpr (DIR == 0) {
    if (a.xtile == 16) {
      if (blend) launch_wspx_k<NR, M, true, 16>(a, st);
      else launch_wspx_k<NR, M, false, 16>(a, st);
      return;
    }
    if (a.xtile) {
      if (blend) launch_wspx_k<NR, M, true, 32>(a, st);
      else launch_wspx_k<NR, M, false, 32>(a, st);
      return;
    }
  }
  if (blend) launch_wsp_k<DIR, NR, M, true>(a, G, st);
  else launch_wsp_k<DIR, NR, M, false>(a, G, st);
}

template <int DIR, int NR>
static void launch_wsp_nr(const SolveArgs& a, const LineShape& s, int nmembers, hipStream_t st) {
  switch (s.M) {
    case 16: launch_wsp_m<DIR, NR, 16>(a, nmembers, st); break;
    case 32: launch_wsp_m<DIR, NR, 32>(a, nmembers, st); break;
  }
}

template <int DIR>
void launch_line_solve_wsp_dir(const SolveArgs& a, const LineShape& s, int nmembers, hipStream_t st) {
  switch (a.op.nR) {
    case 2: launch_wsp_nr<DIR, 2>(a, s, nmembers, st); break;
    case 3: launch_wsp_nr<DIR, 3>(a, s, nmembers, st); break;
    case 4: launch_wsp_nr<DIR, 4>(a, s, nmembers, st); break;
    case 5: launch_wsp_nr<DIR, 5>(a, s, nmembers, st); break;
    case 6: launch_wsp_nr<DIR, 6>(a, s, nmembers, st); break;
    case 7: launch_wsp_nr<DIR, 7>(a, s, nmembers, st); break;
    case 8: launch_wsp_nr<DIR, 8>(a, s, nmembers, st); break;
    case 9: launch_wsp_nr<DIR, 9>(a, s, nmembers, st); break;
    case 10: launch_wsp_nr<DIR, 10>(a, s, nmembers, st); break;
    case 11: launch_wsp_nr<DIR, 11>(a, s, nmembers, st); break;
    default: break;
  }
}

}  // namespace gatt
