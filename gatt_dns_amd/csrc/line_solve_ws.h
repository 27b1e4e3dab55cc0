// Wave-segment variant of the compact line solves (see line_solve_impl.h for
// the algorithm and the reference routines it restates).
//
// Layout: a workgroup owns 64 lines; lane l of every wavefront works on line
// line0 + l, and wavefront w owns the M points [w*M, w*M+M) of all 64 lines.
// Every lane of a wavefront is at the same point at the same time, so the
// coefficients of that point are the same for all 64 lines; each wavefront
// copies its segment's coefficients to LDS once and the sweeps read broadcasts.
//   * y and z lines: lanes are consecutive i, so each point of the segment is
//     one 512-byte row load and store.
//   * x lines: the 64 lines are one contiguous stretch of 64*N doubles, moved
//     through LDS (point-major, 65 words per row) with unit-stride accesses.
//   * The segments of a line are chained through LDS: each wavefront publishes
//     the affine map of its segment, and every lane composes the maps of the
//     segments before (forward) or after (backward) its own, in order.
#pragma once
#include "gatt_internal.h"

namespace gatt {

// Packed coefficients per point: band (2nR-1), fa, fb, ce, D, padded to an even
// count so that a point's record is 16-byte aligned.
constexpr int ws_coef_stride(int nR) { return (2 * nR - 1 + 4 + 1) & ~1; }
constexpr int kMaxWaveSegments = 16;

// Band product, periodic correction, both sweeps and the store of one segment.
template <int DIR, int NR, int M, bool XLDS, bool BLEND, typename CP>
__device__ __forceinline__ void ws_solve(const SolveArgs& a, double* sh, double* tile, const double (&fw)[M + 2 * (NR - 1)],
                                         CP cf, CP cends, int N, int s, int cnt, int w, int G, int lane,
                                         bool active, double* out, long long base, long long stride) {
  constexpr int H = NR - 1;
  constexpr int NW = 2 * H + 1;
  constexpr int CS = ws_coef_stride(NR);
  const int per = a.op.periodic;
  const int e0 = per;
  double* shP = sh;                       // forward maps   [G][64][2]
  double* shB = sh + 2 * G * 64;          // backward maps  [G][64][2]
  double* shD = sh + 4 * G * 64;          // periodic dot partials [G][64]

  double r[M];
#pragma unroll
  for (int tt = 0; tt < M; ++tt) {
    double acc = 0.0;
    if (tt < cnt) {
      CP c = cf + tt * CS;
      acc = c[H] * fw[tt + H];
#pragma unroll
      for (int q = 1; q <= H; ++q) acc = acc + c[H + q] * fw[tt + H + q] + c[H - q] * fw[tt + H - q];
    }
    r[tt] = acc;
  }

  double x1 = 0.0;
  if (per) {
    double part = 0.0;
#pragma unroll
    for (int tt = 0; tt < M; ++tt)
      if (tt < cnt) part = part + r[tt] * cf[tt * CS + NW + 3];
    shD[w * 64 + lane] = part;
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kMaxWaveSegments; ++k)
      if (k < G) x1 += shD[k * 64 + lane];
#pragma unroll
    for (int tt = 0; tt < M; ++tt) {
      if (tt < cnt) {
        const int e = s + tt;
        if (e == 1) r[tt] = r[tt] - cends[0] * x1;
        if (e == N - 1) r[tt] = r[tt] - cends[1] * x1;
      }
    }
  }

  // forward sweep
  double P = 0.0, Q = 1.0;
#pragma unroll
  for (int tt = 0; tt < M; ++tt) {
    if (tt < cnt && s + tt >= e0) {
      const double fa = cf[tt * CS + NW], fb = cf[tt * CS + NW + 1];
      P = (r[tt] - fa * P) * fb;
      Q = -(fa * fb) * Q;
    }
  }
  shP[(w * 64 + lane) * 2] = P;
  shP[(w * 64 + lane) * 2 + 1] = Q;
  __syncthreads();
  double dprev = 0.0;
#pragma unroll
  for (int k = 0; k < kMaxWaveSegments; ++k)
    if (k < w) dprev = shP[(k * 64 + lane) * 2] + shP[(k * 64 + lane) * 2 + 1] * dprev;
#pragma unroll
  for (int tt = 0; tt < M; ++tt) {
    if (tt < cnt && s + tt >= e0) {
      const double fa = cf[tt * CS + NW], fb = cf[tt * CS + NW + 1];
      const double d = (r[tt] - fa * dprev) * fb;
      r[tt] = d;
      dprev = d;
    }
  }

  // backward sweep
  P = 0.0;
  Q = 1.0;
#pragma unroll
  for (int tt = M - 1; tt >= 0; --tt) {
    const int e = s + tt;
    if (tt < cnt && e >= e0) {
      if (e == N - 1) { P = r[tt]; Q = 0.0; }
      else { const double ce = cf[tt * CS + NW + 2]; P = r[tt] - ce * P; Q = -ce * Q; }
    }
  }
  shB[(w * 64 + lane) * 2] = P;
  shB[(w * 64 + lane) * 2 + 1] = Q;
  __syncthreads();
  double xnext = 0.0;
#pragma unroll
  for (int k = kMaxWaveSegments - 1; k >= 0; --k)
    if (k > w && k < G) xnext = shB[(k * 64 + lane) * 2] + shB[(k * 64 + lane) * 2 + 1] * xnext;
#pragma unroll
  for (int tt = M - 1; tt >= 0; --tt) {
    const int e = s + tt;
    if (tt < cnt && e >= e0) {
      const double x = (e == N - 1) ? r[tt] : r[tt] - cf[tt * CS + NW + 2] * xnext;
      r[tt] = x;
      xnext = x;
    }
  }

  // store
  if (XLDS) __syncthreads();   // all maps read before the tile is rewritten
#pragma unroll
  for (int tt = 0; tt < M; ++tt) {
    if (tt < cnt) {
      const int e = s + tt;
      double v = (per && e == 0) ? x1 : r[tt];
      if (BLEND) {
        const double F0 = fw[tt + H];
        v = F0 + a.theta * (v - F0);
      }
      if (XLDS) tile[e * 65 + lane] = v;
      else if (active) out[base + (long long)e * stride] = v;
    }
  }
}

template <int DIR, int NR, int M, bool XLDS, bool BLEND>
__device__ __forceinline__ void line_solve_ws_body(const SolveArgs& a, BlockPos bp) {
  extern __shared__ double sh[];
  constexpr int H = NR - 1;
  constexpr int CS = ws_coef_stride(NR);
  constexpr int W = M + 2 * H;
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int G = blockDim.x >> 6;
  const int N = a.op.N;
  const int line0 = bp.x * 64;
  const int line = line0 + lane;
  const int nact = min(64, a.nlines - line0);
  const bool active = lane < nact;
  const int f = (int)bp.y % a.nf;
  const int mem = (int)bp.y / a.nf;
  const double* in = a.in[f] + (long long)mem * a.in_stride;
  double* out = a.out[f] + (long long)mem * a.out_stride;
  const long long stride = a.stride;
  const int s = w * M;
  int cnt = N - s;
  cnt = cnt < 0 ? 0 : (cnt > M ? M : cnt);
  // LDS: scan maps and dot partials; the x tile (N rows of 65 words) shares the
  // same words, separated from them by barriers
  double* tile = sh;

  long long base = 0;
  if (DIR == 1) {
    const int k = (active ? line : line0) / a.nx;
    base = (long long)((active ? line : line0) - k * a.nx) + (long long)a.nx * a.ny * k;
  } else if (DIR == 2) {
    base = active ? line : line0;
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

  // every load is in bounds (wrapped, then clamped for empty segments), so the
  // window is fetched without branches
  double fw[W];
#pragma unroll
  for (int u = 0; u < W; ++u) {
    int e = s - H + u;
    e = e < 0 ? e + N : (e >= N ? e - N : e);
    e = e >= N ? N - 1 : e;
    fw[u] = XLDS ? tile[e * 65 + lane] : in[base + (long long)e * stride];
  }
  if (XLDS) __syncthreads();   // the tile words are reused by the scan maps

  // operator type of the 64 lines: the host only selects this kernel for
  // directions whose 64-line tiles each have a single type
  const int tu = __builtin_amdgcn_readfirstlane(a.types[line0]);
  // the segment's coefficients (M points) and the two periodic factors go to
  // this wavefront's LDS slot in one batch; the sweeps then read broadcasts
  double* cl = sh + a.lds_coef_offset + (size_t)w * (M * CS + 2);
  const double* src = a.coef + ((size_t)tu * N + s) * CS;
  for (int q = lane; q < cnt * CS; q += 64) cl[q] = src[q];
  if (lane < 2) cl[M * CS + lane] = a.coef_ends[(size_t)tu * 2 + lane];
  __builtin_amdgcn_s_waitcnt(0);
  __builtin_amdgcn_wave_barrier();
  ws_solve<DIR, NR, M, XLDS, BLEND>(a, sh, tile, fw, (const double*)cl, (const double*)(cl + M * CS), N, s, cnt, w,
                                    G, lane, active, out, base, stride);
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

template <int DIR, int NR, int M, bool XLDS, bool BLEND>
__global__ __launch_bounds__(M >= 32 ? 512 : 1024) void line_solve_ws_kernel(SolveArgs a) {
  line_solve_ws_body<DIR, NR, M, XLDS, BLEND>(a, grid_pos());
}
template <int DIR, int NR, int M, bool XLDS, bool BLEND>
__global__ __launch_bounds__(M >= 32 ? 512 : 1024) void line_solve_ws_ragged(const SolveArgs* __restrict__ args,
                                                                              RaggedMap map) {
  BlockPos bp;
  const int m = ragged_pos(map, bp);
  line_solve_ws_body<DIR, NR, M, XLDS, BLEND>(args[m], bp);
}

// LDS words: scan maps (5*G*64), then per wavefront M*CS + 2 coefficient words.
// Along x the tile (65*N words) is dead once the windows are in registers, so
// maps and coefficients reuse its words.
inline size_t ws_lds_words_front(int dir, int G, int N) {
  (void)dir;
  (void)N;
  return (size_t)5 * G * 64;
}
inline size_t ws_lds_bytes(int dir, int G, int N, int M, int nR) {
  const size_t work = ws_lds_words_front(dir, G, N) + (size_t)G * (M * ws_coef_stride(nR) + 2);
  const size_t tile = dir == 0 ? (size_t)N * 65 : 0;
  return sizeof(double) * (work > tile ? work : tile);
}

template <int DIR, int NR, int M>
static void launch_ws_m(const SolveArgs& a0, const LineShape& s, int nmembers, hipStream_t st) {
  SolveArgs a = a0;
  a.inv_n = 1.0f / (float)a.op.N;
  const int G = (a.op.N + M - 1) / M;
  const long long tiles = (a.nlines + 63) / 64;
  a.lds_coef_offset = (int)ws_lds_words_front(DIR, G, a.op.N);
  const size_t lds = ws_lds_bytes(DIR, G, a.op.N, M, NR);
  dim3 grid((unsigned)tiles, (unsigned)(a.nf * nmembers));
  const bool blend = a.mode == 1 && a.theta > 0.0;
  constexpr bool X = DIR == 0;
  auto kr = blend ? line_solve_ws_ragged<DIR, NR, M, X, true> : line_solve_ws_ragged<DIR, NR, M, X, false>;
  if (record_launch((const void*)kr, a, grid, dim3(64 * G), lds)) return;
  auto k = blend ? line_solve_ws_kernel<DIR, NR, M, X, true> : line_solve_ws_kernel<DIR, NR, M, X, false>;
  if (lds > 65536) (void)hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  hipLaunchKernelGGL(k, grid, dim3(64 * G), lds, st, a);
}

template <int DIR, int NR>
static void launch_ws_nr(const SolveArgs& a, const LineShape& s, int nmembers, hipStream_t st) {
  switch (s.M) {
    case 16: launch_ws_m<DIR, NR, 16>(a, s, nmembers, st); break;
    case 32: launch_ws_m<DIR, NR, 32>(a, s, nmembers, st); break;
  }
}

template <int DIR>
void launch_line_solve_ws_dir(const SolveArgs& a, const LineShape& s, int nmembers, hipStream_t st) {
  switch (a.op.nR) {
    case 2: launch_ws_nr<DIR, 2>(a, s, nmembers, st); break;
    case 3: launch_ws_nr<DIR, 3>(a, s, nmembers, st); break;
    case 4: launch_ws_nr<DIR, 4>(a, s, nmembers, st); break;
    case 5: launch_ws_nr<DIR, 5>(a, s, nmembers, st); break;
    case 6: launch_ws_nr<DIR, 6>(a, s, nmembers, st); break;
    case 7: launch_ws_nr<DIR, 7>(a, s, nmembers, st); break;
    case 8: launch_ws_nr<DIR, 8>(a, s, nmembers, st); break;
    case 9: launch_ws_nr<DIR, 9>(a, s, nmembers, st); break;
    case 10: launch_ws_nr<DIR, 10>(a, s, nmembers, st); break;
    case 11: launch_ws_nr<DIR, 11>(a, s, nmembers, st); break;
    default: break;
  }
}

}  // namespace gatt
