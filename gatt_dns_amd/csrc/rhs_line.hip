// The last direction of the second derivatives fused with the right-hand side and
// the stage update (nseq.F90:241-318 and timeIntegration.F90:382-496 along one
// direction of lines).
//
// A workgroup owns 16 consecutive y (or z) lines; lane l of wavefront w works on
// line l % 16 and on the 8-point segment 4w + l / 16 of it, so a line of up to 256
// points is one workgroup of up to 512 threads, whose wavefronts may hold 256
// registers: the four solved fields stay in registers through the assembly.  For
// each viscous flux of the
// direction (y: txy, tyy, tyz, qy; z: txz, tyz, tzz, qz; 2-D: txy, tyy, qy) the
// lane computes the flux at its window points from E and the first derivatives
// (visc_flux, the expressions of visc_kernel), applies the derivative operator
// (the band and segmented Thomas sweeps of line_solve_wsp.h) and keeps the 8
// results in registers.  Then every point of the segment is assembled with the
// flux derivatives of the other directions read from the workspace
// (rhs_point.h) and the stage update is written.  Neither the stresses nor the
// derivatives of this direction's fluxes ever reach memory.
#include "line_solve_wsp.h"
#include "rhs_point.h"

namespace gatt {

namespace {

constexpr int kRhsLines = 16;    // lines per workgroup
constexpr int kRhsM = 8;         // points per lane
constexpr int kRhsMaxSeg = 32;   // segments per line (lines up to 256 points)

inline int rhs_segments(int N) {
  const int G = (N + kRhsM - 1) / kRhsM;
  constexpr int spw = 64 / kRhsLines;
  return (G + spw - 1) / spw * spw;    // whole wavefronts of segments
}

template <int NR>
inline size_t rhs_lds_doubles(int GA) {
  return (size_t)5 * GA * kRhsLines + (size_t)GA * (kRhsM * wsp_coef_stride(NR) + 2) + 16 * 5;
}

template <int DIR, int NR>
__global__ __launch_bounds__(512) void rhs_line_kernel(RhsLineArgs r) {
  extern __shared__ double sh[];
  constexpr int M = kRhsM;
  constexpr int LW = kRhsLines;
  constexpr int SPW = 64 / LW;
  constexpr int H = NR - 1;
  constexpr int CS = wsp_coef_stride(NR);
  constexpr int W = M + 2 * H;
  constexpr int SLOT = M * CS + 2;
  constexpr int CP = (SLOT + LW - 1) / LW;
  const SolveArgs& s = r.s;
  const AssembleArgs& a = r.a;
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int l = lane & (LW - 1);
  const int sg = SPW * w + lane / LW;
  const int N = s.op.N;
  const int G = (N + M - 1) / M;
  const int GA = (G + SPW - 1) / SPW * SPW;
  const int per = s.op.periodic;
  const int st = sg * M;
  int cnt = N - st;
  cnt = cnt < 0 ? 0 : (cnt > M ? M : cnt);
  double* shD = sh;                              // periodic dot partials [GA][LW]
  double* shP = sh + GA * LW;                    // forward maps  [2][GA][LW]
  double* shB = sh + 3 * GA * LW;                // backward maps [2][GA][LW]
  double* cl = sh + 5 * GA * LW + sg * SLOT;     // this segment's coefficients
  double* wsum = sh + 5 * GA * LW + GA * SLOT;   // forcing-sum partials of the wavefronts [16][5]
  const int line0 = blockIdx.x * LW;
  const int nact = min(LW, s.nlines - line0);
  const int mem = blockIdx.y;
  const int line = l < nact ? line0 + l : line0;
  const long long np = a.np;
  const long long mo = (long long)mem * a.member_stride;
  const long long wo = (long long)mem * a.work_stride;
  const double* S = a.S + mo;
  const double* E = S + 4 * np;
  const double* D = a.D + wo;
  const double* Gw = a.G + wo;
  const long long stride = s.stride;
  const long long base = wsp_base<DIR>(s, line);

  // the segment's fused coefficient records (zeros past the line end); a segment's slot is
  // written and read by the lanes of one wavefront, whose LDS accesses run in order
  {
    const int tu = __builtin_amdgcn_readfirstlane(s.types[line0]);
    const double* src = s.fcoef + ((size_t)tu * N + (sg < G ? st : 0)) * CS;
    const double* ends = s.fcoef_ends + (size_t)tu * 2;
#pragma unroll
    for (int k = 0; k < CP; ++k) {
      const int jj = l + LW * k;
      double v = 0.0;
      if (jj < cnt * CS) v = src[jj];
      else if (jj >= M * CS && jj < SLOT) v = ends[jj - M * CS];
      if (jj < SLOT) cl[jj] = v;
    }
  }
  __builtin_amdgcn_wave_barrier();

  // derivatives of this direction's fluxes, kept in registers
  const int nfl = a.is3d ? 4 : 3;
  double res[4][M];
#pragma unroll
  for (int fl = 0; fl < 4; ++fl) {
    if (fl < nfl) {
      int q;
      if (DIR == 1) q = fl == 0 ? 3 : (fl == 1 ? 1 : ((fl == 2 && a.is3d) ? 5 : 7));
      else q = fl == 0 ? 4 : (fl == 1 ? 5 : (fl == 2 ? 2 : 8));
      double fw[W];
#pragma unroll
      for (int u = 0; u < W; ++u) {
        int e = st - H + u;
        e = e < 0 ? e + N : (e >= N ? e - N : e);
        e = e >= N ? N - 1 : e;
        const long long n = base + (long long)e * stride;
        fw[u] = visc_flux(q, a.c, E[n], D, n, np, a.is3d);
      }
      double rr[M];
      double f0[1];
      wsp_band<NR, M, false>(rr, f0, fw, cl);
      const double x1 = wsp_sweeps<NR, M, kRhsMaxSeg, true>(rr, cl, cl + M * CS, shD, shP, shB, cnt, sg, G, l, per, LW, GA);
      if (per && sg == 0) rr[0] = x1;
#pragma unroll
      for (int t = 0; t < M; ++t) res[fl][t] = rr[t];
    }
  }

  // the right-hand side and the stage update at the segment's points, one point per
  // iteration (a rolled loop keeps the register use bounded); the results move down
  // one register each iteration, so point t's values are always in res[.][0]
  double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  const int todo = l < nact ? cnt : 0;
#pragma unroll 1
  for (int t = 0; t < todo; ++t) {
    const long long n = base + (long long)(st + t) * stride;
    double g[12];
#pragma unroll
    for (int k = 0; k < 12; ++k) g[k] = 0.0;
    if (DIR == 1) {
      g[0] = Gw[0 * np + n];
      g[1] = Gw[1 * np + n];
      g[3] = Gw[3 * np + n];
      g[4] = res[0][0];
      g[5] = res[1][0];
      if (a.is3d) {
        g[2] = Gw[2 * np + n];
        g[6] = res[2][0];
        g[7] = res[3][0];
#pragma unroll
        for (int k = 8; k < 12; ++k) g[k] = Gw[k * np + n];
      } else {
        g[7] = res[2][0];
      }
    } else {
#pragma unroll
      for (int k = 0; k < 8; ++k) g[k] = Gw[k * np + n];
#pragma unroll
      for (int k = 0; k < 4; ++k) g[8 + k] = res[k][0];
    }
    double d[5], sv[5], ov[5];
    rhs_point(a, S, D, n, g, mem, d, sv);
    stage_point(a, mo, n, d, sv, ov);
    acc[0] += ov[0] * ov[0] + ov[1] * ov[1] + ov[2] * ov[2];
    acc[1] += ov[0];
    acc[2] += ov[1];
    acc[3] += ov[2];
    acc[4] += ov[4];
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
      for (int j = 0; j + 1 < M; ++j) res[k][j] = res[k][j + 1];
  }
  if (a.sum_partial) {
    // turbulenceGenerator.F90:39-44 on the stage state just written: the workgroup's
    // partial sums (wavefront trees, then the wavefronts in order)
#pragma unroll
    for (int m = 32; m > 0; m >>= 1)
#pragma unroll
      for (int k = 0; k < 5; ++k) acc[k] += __shfl_xor(acc[k], m);
    if (lane == 0)
#pragma unroll
      for (int k = 0; k < 5; ++k) wsum[w * 5 + k] = acc[k];
    __syncthreads();
    if (threadIdx.x < 5) {
      const int nw = blockDim.x >> 6;
      double t = 0.0;
      for (int v = 0; v < nw; ++v) t += wsum[v * 5 + threadIdx.x];
      a.sum_partial[((long long)mem * gridDim.x + blockIdx.x) * 5 + threadIdx.x] = t;
    }
  }
}

template <int DIR, int NR>
void launch_k(const RhsLineArgs& r, int nmembers, hipStream_t st) {
  auto k = rhs_line_kernel<DIR, NR>;
  const int GA = rhs_segments(r.s.op.N);
  const size_t lds = sizeof(double) * rhs_lds_doubles<NR>(GA);
  if (lds > 65536) (void)hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  dim3 grid((unsigned)((r.s.nlines + kRhsLines - 1) / kRhsLines), (unsigned)nmembers);
  hipLaunchKernelGGL(k, grid, dim3(GA * kRhsLines), lds, st, r);
}

template <int DIR>
void launch_dir(const RhsLineArgs& r, int nmembers, hipStream_t st) {
  switch (r.s.op.nR) {
    case 2: launch_k<DIR, 2>(r, nmembers, st); break;
    case 3: launch_k<DIR, 3>(r, nmembers, st); break;
    case 4: launch_k<DIR, 4>(r, nmembers, st); break;
    case 5: launch_k<DIR, 5>(r, nmembers, st); break;
    case 6: launch_k<DIR, 6>(r, nmembers, st); break;
    default: break;
  }
}

}  // namespace

bool rhs_line_fits(int N, int nR) {
  return nR >= 2 && nR <= kFluxMaxNR && N >= 2 && (N + kRhsM - 1) / kRhsM <= kRhsMaxSeg &&
         kRhsM + nR - 1 <= N + 1 && nR - 1 <= N;
}

int rhs_line_blocks(int nlines) { return (nlines + kRhsLines - 1) / kRhsLines; }

void launch_rhs_line(const RhsLineArgs& r, int nmembers, hipStream_t st) {
  if (record_unsupported()) return;
  if (r.s.dir == 2) launch_dir<2>(r, nmembers, st);
  else launch_dir<1>(r, nmembers, st);
}

}  // namespace gatt
