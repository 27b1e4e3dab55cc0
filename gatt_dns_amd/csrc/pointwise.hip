// Point-wise parts of the time step: viscous stresses and heat fluxes, the
// RHS assembly fused with the RK4 stage update, the forcing sums, the SFD
// blend and the per-step reductions.  One thread per grid point; every field
// is an FP64 column-major array (x fastest), so all accesses are unit-stride.
#include "gatt_internal.h"
#include "rhs_point.h"

namespace gatt {

static inline unsigned grid_for(long long n, int threads) {
  long long b = (n + threads - 1) / threads;
  return (unsigned)b;
}

// nseq.F90:218-239 (3-D) and 72-85 (2-D): Sutherland viscosity, stresses, heat flux.
__device__ __forceinline__ void visc_body(const ViscArgs& a, BlockPos bp) {
  const long long n = bp.x * (long long)blockDim.x + threadIdx.x;
  if (n >= a.np) return;
  const long long mo = (long long)bp.y * a.member_stride;
  const long long wo = (long long)bp.y * a.work_stride;
  const long long np = a.np;
  const double* D = a.D + wo;
  double* T = a.T + wo;
  const double E = ld_nt(a.E + mo + n);
  const FlowConst c = a.c;
  const double mu1 = c.muA / (c.muC * E + c.muB);
  const double mu2 = c.muC * E;
  const double mu = mu1 * mu2 * sqrt(mu2);
  const double rmu = c.invRe * mu;
  const double DUx = ld_nt(D + 0 * np + n), DUy = ld_nt(D + 1 * np + n);
  const double DVx = ld_nt(D + 3 * np + n), DVy = ld_nt(D + 4 * np + n);
  const double DEx = ld_nt(D + 12 * np + n), DEy = ld_nt(D + 13 * np + n);
  if (a.is3d) {
    const double DUz = ld_nt(D + 2 * np + n), DVz = ld_nt(D + 5 * np + n);
    const double DWx = ld_nt(D + 6 * np + n), DWy = ld_nt(D + 7 * np + n), DWz = ld_nt(D + 8 * np + n);
    const double DEz = ld_nt(D + 14 * np + n);
    const double tz = -(2.0 / 3.0 * (DUx + DVy + DWz));
    st_nt(T + 0 * np + n, rmu * (2.0 * DUx + tz));
    st_nt(T + 1 * np + n, rmu * (2.0 * DVy + tz));
    st_nt(T + 2 * np + n, rmu * (2.0 * DWz + tz));
    st_nt(T + 3 * np + n, rmu * (DUy + DVx));
    st_nt(T + 4 * np + n, rmu * (DUz + DWx));
    st_nt(T + 5 * np + n, rmu * (DVz + DWy));
    st_nt(T + 6 * np + n, -(c.qk * mu * DEx));
    st_nt(T + 7 * np + n, -(c.qk * mu * DEy));
    st_nt(T + 8 * np + n, -(c.qk * mu * DEz));
  } else {
    const double dv = 2.0 / 3.0 * (DUx + DVy);
    st_nt(T + 0 * np + n, rmu * (2.0 * DUx - dv));
    st_nt(T + 1 * np + n, rmu * (2.0 * DVy - dv));
    st_nt(T + 3 * np + n, rmu * (DUy + DVx));
    st_nt(T + 6 * np + n, -(c.qk * mu * DEx));
    st_nt(T + 7 * np + n, -(c.qk * mu * DEy));
  }
}

__global__ __launch_bounds__(256) void visc_kernel(ViscArgs a) { visc_body(a, grid_pos()); }
__global__ __launch_bounds__(256) void visc_ragged(const ViscArgs* __restrict__ args, RaggedMap map) {
  BlockPos bp;
  const int m = ragged_pos(map, bp);
  visc_body(args[m], bp);
}

void launch_visc(const ViscArgs& a, int nmembers, hipStream_t st) {
  const dim3 grid(grid_for(a.np, 256), nmembers);
  if (record_launch((const void*)visc_ragged, a, grid, dim3(256), 0)) return;
  hipLaunchKernelGGL(visc_kernel, grid, dim3(256), 0, st, a);
}

// nseq.F90:280-318 (3-D), 106-134 (2-D), the forcings and one stage update (rhs_point.h),
// one thread per point; the flux derivatives come from the G arrays.
__device__ __forceinline__ void assemble_body(const AssembleArgs& a, BlockPos bp) {
  const long long n = bp.x * (long long)blockDim.x + threadIdx.x;
  if (n >= a.np) return;
  const long long mo = (long long)bp.y * a.member_stride;
  const long long np = a.np;
  const long long wo = (long long)bp.y * a.work_stride;
  const double* G = a.G + wo;
  double g[12];
#pragma unroll
  for (int k = 0; k < 12; ++k) g[k] = (a.is3d || k == 0 || k == 1 || k == 3 || k == 4 || k == 5 || k == 7)
                                          ? ld_nt(G + k * np + n) : 0.0;
  double d[5], sv[5], ov[5];
  rhs_point(a, a.S + mo, a.D + wo, n, g, (int)bp.y, d, sv);
  stage_point(a, mo, n, d, sv, ov);
  if (a.sum_partial) {
    // turbulenceGenerator.F90:39-44 on the stage state just written: wavefront trees, then
    // the four wavefronts of the block in order (no thread of the block returned early)
    double acc[5] = {ov[0] * ov[0] + ov[1] * ov[1] + ov[2] * ov[2], ov[0], ov[1], ov[2], ov[4]};
#pragma unroll
    for (int m = 32; m > 0; m >>= 1)
#pragma unroll
      for (int k = 0; k < 5; ++k) acc[k] += __shfl_xor(acc[k], m);
    __shared__ double wsum[4][5];
    const int wv = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
#pragma unroll
      for (int k = 0; k < 5; ++k) wsum[wv][k] = acc[k];
    __syncthreads();
    if (threadIdx.x < 5) {
      const int k = threadIdx.x;
      a.sum_partial[((long long)bp.y * bp.gx + bp.x) * 5 + k] =
          ((wsum[0][k] + wsum[1][k]) + wsum[2][k]) + wsum[3][k];
    }
  }
}

__global__ __launch_bounds__(256) void assemble_kernel(AssembleArgs a) { assemble_body(a, grid_pos()); }
__global__ __launch_bounds__(256) void assemble_ragged(const AssembleArgs* __restrict__ args, RaggedMap map) {
  BlockPos bp;
  const int m = ragged_pos(map, bp);
  assemble_body(args[m], bp);
}

void launch_assemble(const AssembleArgs& a, int nmembers, hipStream_t st) {
  const dim3 grid(grid_for(a.np, 256), nmembers);
  if (record_launch((const void*)assemble_ragged, a, grid, dim3(256), 0)) return;
  hipLaunchKernelGGL(assemble_kernel, grid, dim3(256), 0, st, a);
}

// Global sums of turbulenceGenerator.F90:39-44, deterministic two-stage reduction.
__device__ __forceinline__ void forcing_partial_body(const SumArgs& a, BlockPos bp) {
  __shared__ double sh[6][256];
  const long long mo = (long long)bp.y * a.member_stride;
  const long long np = a.np;
  const double* S = a.S + mo;
  double acc[5] = {0, 0, 0, 0, 0};
  for (long long n = bp.x * (long long)blockDim.x + threadIdx.x; n < np; n += (long long)bp.gx * blockDim.x) {
    const double U = S[n], V = S[np + n], W = S[2 * np + n], E = S[4 * np + n];
    acc[0] += U * U + V * V + W * W;
    acc[1] += U;
    acc[2] += V;
    acc[3] += W;
    acc[4] += E;
  }
  for (int k = 0; k < 5; ++k) sh[k][threadIdx.x] = acc[k];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int k = 0; k < 5; ++k) sh[k][threadIdx.x] += sh[k][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x < 5) a.partial[((long long)bp.y * a.nblocks + bp.x) * 5 + threadIdx.x] = sh[threadIdx.x][0];
}
__global__ __launch_bounds__(256) void forcing_partial_kernel(SumArgs a) { forcing_partial_body(a, grid_pos()); }
__global__ __launch_bounds__(256) void forcing_partial_ragged(const SumArgs* __restrict__ args, RaggedMap map) {
  BlockPos bp;
  const int m = ragged_pos(map, bp);
  forcing_partial_body(args[m], bp);
}

// Fixed-order tree over the block partials: thread k sums partials k, k+256, ...
__device__ __forceinline__ void forcing_final_body(const SumArgs& a, BlockPos bp) {
  __shared__ double sh[5][256];
  const int mem = bp.x;
  double acc[5] = {0, 0, 0, 0, 0};
  for (int b = threadIdx.x; b < a.nblocks; b += blockDim.x)
    for (int k = 0; k < 5; ++k) acc[k] += a.partial[((long long)mem * a.nblocks + b) * 5 + k];
  for (int k = 0; k < 5; ++k) sh[k][threadIdx.x] = acc[k];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int k = 0; k < 5; ++k) sh[k][threadIdx.x] += sh[k][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x < 5) a.sums[8 * mem + threadIdx.x] = sh[threadIdx.x][0];
  if (threadIdx.x == 5) a.sums[8 * mem + 5] = (double)a.np;
}

__global__ __launch_bounds__(256) void forcing_final_kernel(SumArgs a) { forcing_final_body(a, grid_pos()); }
__global__ __launch_bounds__(256) void forcing_final_ragged(const SumArgs* __restrict__ args, RaggedMap map) {
  BlockPos bp;
  const int m = ragged_pos(map, bp);
  forcing_final_body(args[m], bp);
}

void launch_forcing_final(const SumArgs& a, int nmembers, hipStream_t st) {
  if (record_launch((const void*)forcing_final_ragged, a, dim3(nmembers), dim3(256), 0)) return;
  hipLaunchKernelGGL(forcing_final_kernel, dim3(nmembers), dim3(256), 0, st, a);
}

void launch_forcing_sums(const SumArgs& a, int nmembers, hipStream_t st) {
  const dim3 grid(a.nblocks, nmembers);
  if (!record_launch((const void*)forcing_partial_ragged, a, grid, dim3(256), 0))
    hipLaunchKernelGGL(forcing_partial_kernel, grid, dim3(256), 0, st, a);
  launch_forcing_final(a, nmembers, st);
}

// the tracked points of the log line (main.F90:478-486): one thread per (point, field)
__global__ __launch_bounds__(256) void gather_points_kernel(const double* __restrict__ state, long long np, int is3d,
                                                            int npts, const long long* __restrict__ idx,
                                                            double* __restrict__ out) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= 5 * npts) return;
  const int p = i / 5, f = i - 5 * p;
  out[i] = (f == 2 && !is3d) ? 0.0 : state[(long long)f * np + idx[p]];
}

void launch_gather_points(const double* state, long long np, int is3d, int npts, const long long* idx, double* out,
                          hipStream_t st) {
  if (npts <= 0) return;
  hipLaunchKernelGGL(gather_points_kernel, dim3((5 * npts + 255) / 256), dim3(256), 0, st, state, np, is3d, npts, idx,
                     out);
}

// main.F90:382-401
__device__ __forceinline__ void sfd_body(const SfdArgs& a, BlockPos bp) {
  const long long n = bp.x * (long long)blockDim.x + threadIdx.x;
  if (n >= a.np) return;
  const long long mo = (long long)bp.y * a.member_stride;
  const double X = a.chi ? a.chi[n] : a.chi_val;
  const double dtX = a.dt * X;
  const double dtD = a.delta > 0 ? a.dt / a.delta : 0.0;
  for (int f = 0; f < 5; ++f) {
    if (!(a.nf_mask & (1 << f))) continue;
    const long long o = mo + f * a.np + n;
    if (a.transpose) {
      // transpose of  M' = M + dtD (q - M),  q' = q - dtX (q - M')  on the adjoints (qb', Mb'):
      // Mb'' = Mb' + dtX qb',  qb = qb' - dtX qb' + dtD Mb'',  Mb = Mb'' - dtD Mb''
      const double qb = a.Q[o];
      if (a.delta > 0 && a.Mean) {
        const double mb = a.Mean[o] + dtX * qb;
        a.Q[o] = (qb - dtX * qb) + dtD * mb;
        a.Mean[o] = mb - dtD * mb;
      } else {
        a.Q[o] = qb - dtX * qb;
      }
      continue;
    }
    // without a mean array the blend acts on a perturbation of a frozen mean
    double M = a.Mean ? a.Mean[o] : 0.0;
    const double q = a.Q[o];
    if (a.delta > 0 && a.Mean) {
      M = M + dtD * (q - M);
      if (!a.mean_ro) a.Mean[o] = M;
    }
    a.Q[o] = q - dtX * (q - M);
  }
}

__global__ __launch_bounds__(256) void sfd_kernel(SfdArgs a) { sfd_body(a, grid_pos()); }
__global__ __launch_bounds__(256) void sfd_ragged(const SfdArgs* __restrict__ args, RaggedMap map) {
  BlockPos bp;
  const int m = ragged_pos(map, bp);
  sfd_body(args[m], bp);
}

void launch_sfd(const SfdArgs& a, int nmembers, hipStream_t st) {
  const dim3 grid(grid_for(a.np, 256), nmembers);
  if (record_launch((const void*)sfd_ragged, a, grid, dim3(256), 0)) return;
  hipLaunchKernelGGL(sfd_kernel, grid, dim3(256), 0, st, a);
}

__global__ void clear_u64_kernel(unsigned long long* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0ull;
}
__global__ void clear_u64_ragged(const ClearArgs* __restrict__ args, RaggedMap map) {
  BlockPos bp;
  const ClearArgs& a = args[ragged_pos(map, bp)];
  const int i = bp.x * blockDim.x + threadIdx.x;
  if (i < a.n) a.p[i] = 0ull;
}

void launch_clear_u64(unsigned long long* p, int n, hipStream_t st) {
  const dim3 grid((n + 255) / 256);
  if (record_launch((const void*)clear_u64_ragged, ClearArgs{p, n}, grid, dim3(256), 0)) return;
  hipLaunchKernelGGL(clear_u64_kernel, grid, dim3(256), 0, st, p, n);
}

__global__ void clear_i32_kernel(int* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

void launch_clear_i32(int* p, int n, hipStream_t st) {
  hipLaunchKernelGGL(clear_i32_kernel, dim3((n + 255) / 256), dim3(256), 0, st, p, n);
}

// max |q_new - q| per field (main.F90:414-418) and max |U|,|V|,|W| of the new
// state (main.F90:197-203 of the next step).  Non-negative doubles order like
// their bit patterns, and a NaN pattern orders above +inf, so NaN propagates.
__device__ __forceinline__ void step_reduce_body(const ReduceArgs& a, BlockPos bp) {
  __shared__ unsigned long long sh[8][256];
  const long long mo = (long long)bp.y * a.member_stride;
  const long long np = a.np;
  unsigned long long acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  // several points per thread: the loads of four iterations go out together
#pragma unroll 4
  for (long long n = bp.x * (long long)blockDim.x + threadIdx.x; n < np; n += (long long)bp.gx * blockDim.x) {
#pragma unroll
    for (int f = 0; f < 5; ++f) {
      if (!a.is3d && f == 2) continue;
      const double qn = ld_nt(a.Qn + mo + f * np + n);
      const double d = fabs(qn - ld_nt(a.Q + mo + f * np + n));
      const unsigned long long b = (unsigned long long)__double_as_longlong(d);
      acc[f] = b > acc[f] ? b : acc[f];
      if (f < 3) {   // max |u|, |v|, |w| from the same loads
        const unsigned long long u = (unsigned long long)__double_as_longlong(fabs(qn));
        acc[5 + f] = u > acc[5 + f] ? u : acc[5 + f];
      }
    }
  }
  for (int k = 0; k < 8; ++k) sh[k][threadIdx.x] = acc[k];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int k = 0; k < 8; ++k) {
        const unsigned long long o = sh[k][threadIdx.x + w];
        if (o > sh[k][threadIdx.x]) sh[k][threadIdx.x] = o;
      }
    __syncthreads();
  }
  if (threadIdx.x < 8) atomicMax(a.out + 8 * bp.y + threadIdx.x, sh[threadIdx.x][0]);
}
__global__ __launch_bounds__(256) void step_reduce_kernel(ReduceArgs a) { step_reduce_body(a, grid_pos()); }
__global__ __launch_bounds__(256) void step_reduce_ragged(const ReduceArgs* __restrict__ args, RaggedMap map) {
  BlockPos bp;
  const int m = ragged_pos(map, bp);
  step_reduce_body(args[m], bp);
}

void launch_step_reduce(const ReduceArgs& a, int nmembers, hipStream_t st) {
  // four points per thread where that still leaves 64 workgroups (fewer contend for the eight
  // atomic maxima); on smaller grids one point per thread, up to 64 workgroups
  long long blocks = std::max((a.np + 1023) / 1024, std::min((a.np + 255) / 256, 64LL));
  if (blocks > 2048) blocks = 2048;
  const dim3 grid((unsigned)blocks, nmembers);
  if (record_launch((const void*)step_reduce_ragged, a, grid, dim3(256), 0)) return;
  hipLaunchKernelGGL(step_reduce_kernel, grid, dim3(256), 0, st, a);
}

__device__ __forceinline__ void copy_body(double* dst, const double* src, long long n, BlockPos bp) {
  for (long long i = bp.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)bp.gx * blockDim.x)
    dst[i] = src[i];
}
__global__ __launch_bounds__(256) void copy_kernel(double* dst, const double* src, long long n) {
  copy_body(dst, src, n, grid_pos());
}
__global__ __launch_bounds__(256) void copy_ragged(const CopyArgs* __restrict__ args, RaggedMap map) {
  BlockPos bp;
  const CopyArgs& a = args[ragged_pos(map, bp)];
  copy_body(a.dst, a.src, a.n, bp);
}

void launch_copy(double* dst, const double* src, long long n, hipStream_t st) {
  long long blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  const dim3 grid((unsigned)blocks);
  if (record_launch((const void*)copy_ragged, CopyArgs{dst, src, n}, grid, dim3(256), 0)) return;
  hipLaunchKernelGGL(copy_kernel, grid, dim3(256), 0, st, dst, src, n);
}

// The step record of member 0's eight maxima (max changes, max |u|) and the sticky divergence
// flag, raised when a member's change is NaN or all its changes are zero (main.F90:434); then
// the maxima are cleared for the next step's reduction and the clock is advanced by dt
// (main.F90:424) when given.
__device__ __forceinline__ void record_body(unsigned long long* red, double* rec, int* flag, int nmembers,
                                            double* clock, double dt) {
  if (threadIdx.x != 0) return;
  int div = 0;
  for (int mbr = 0; mbr < nmembers; ++mbr) {
    double s = 0;
    int nan = 0;
    for (int f = 0; f < 5; ++f) {
      const double v = __longlong_as_double((long long)red[8 * mbr + f]);
      if (v != v) nan = 1;
      s += v;
    }
    if (nan || s == 0.0) div = 1;
  }
  for (int k = 0; k < 8; ++k) rec[k] = __longlong_as_double((long long)red[k]);
  for (int k = 0; k < 8 * nmembers; ++k) red[k] = 0ull;
  if (div) *flag = 1;
  if (clock) *clock = *clock + dt;
}

__global__ void record_kernel(RecordArgs a) { record_body(a.red, a.rec, a.flag, a.nmembers, a.clock, a.dt); }
__global__ void record_ragged(const RecordArgs* __restrict__ args, RaggedMap map) {
  BlockPos bp;
  const RecordArgs& a = args[ragged_pos(map, bp)];
  record_body(a.red, a.rec, a.flag, a.nmembers, a.clock, a.dt);
}

void launch_record(unsigned long long* red, double* rec, int* flag, int nmembers, hipStream_t st, double* clock,
                   double dt) {
  const RecordArgs a{red, rec, flag, nmembers, clock, dt};
  if (record_launch((const void*)record_ragged, a, dim3(1), dim3(64), 0)) return;
  hipLaunchKernelGGL(record_kernel, dim3(1), dim3(64), 0, st, a);
}

}  // namespace gatt
