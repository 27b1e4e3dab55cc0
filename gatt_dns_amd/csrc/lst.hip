// Kernels of the Jacobian-free stability analysis on a batch of members:
// perturbation states built from Krylov vectors, Krylov vectors collected from
// advanced states, and the tall-skinny products of block Gram-Schmidt.
//
// A Krylov vector packs the flow-region points of U,V,(W),R,E one field after
// the other (flow2vec / vec2flow, calcInstability.m:480-515, 517-582): entry
// s*nflow + r holds field slot s at the r-th flow-region point, the region
// points in increasing column-major order.  Slots are U,V,R,E in 2-D and
// U,V,W,R,E in 3-D.
#include "gatt_internal.h"

namespace gatt {

// state(member) = base + scale * vec2flow(V[:, member]); the SFD mean of the
// member, when given, is reset to the base (frozen mean = base flow).  With the
// singleBeta packing (calcInstability.m:557-579) the vector holds one z plane and
// plane k of the disturbance is (scale v) cos(z_k), or (scale v) sin(z_k) for W.
__global__ __launch_bounds__(256) void perturb_kernel(PerturbArgs a) {
  const long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (p >= a.np) return;
  const int b = blockIdx.y;
  double* q = a.state + (long long)b * a.member_stride;
  double* mean = a.mean ? a.mean + (long long)b * a.member_stride : nullptr;
  const long long pp = a.single_beta ? p % a.plane : p;
  const int kz = a.single_beta ? (int)(p / a.plane) : 0;
  const int r = a.rmap ? a.rmap[pp] : (int)pp;
  const double* v = a.V ? a.V + (long long)b * a.ldv : nullptr;
  int s = 0;
#pragma unroll
  for (int f = 0; f < 5; ++f) {
    const long long o = (long long)f * a.np + p;
    const double base = a.base[o];
    double val = base;
    const bool active = a.is3d || f != 2;
    if (active && v && r >= 0) {
      const double d = a.scale * v[(long long)s * a.nflow + r];
      val = a.single_beta ? base + d * a.zw[(f == 2 ? a.nz : 0) + kz] : base + d;
    }
    if (active) ++s;
    q[o] = val;
    if (mean) mean[o] = base;
  }
}

// singleBeta projection of one field at plane point pp (flow2vec, calcInstability.m:493-509):
// sum over the z planes in order of weight * value
__device__ __forceinline__ double sb_project(const double* q, long long pp, long long plane, int nz,
                                             const double* w) {
  double acc = w[0] * q[pp];
  for (int k = 1; k < nz; ++k) acc = acc + w[k] * q[pp + k * plane];
  return acc;
}

// W[:, b] = scale * (flow2vec(member first+b) - flow2vec(member minus+b) or ref or 0)
__global__ __launch_bounds__(256) void collect_kernel(CollectArgs a) {
  const long long r = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (r >= a.nflow) return;
  const int b = blockIdx.y;
  const long long p = a.idx ? (long long)a.idx[r] : r;
  const double* q = a.plus + (long long)b * a.member_stride;
  const double* m = a.minus ? a.minus + (long long)b * a.member_stride : nullptr;
  double* w = a.W + (long long)b * a.ldw;
  int s = 0;
#pragma unroll
  for (int f = 0; f < 5; ++f) {
    if (!a.is3d && f == 2) continue;
    const long long o = (long long)f * a.np + p;
    const long long vo = (long long)s * a.nflow + r;
    double d;
    if (a.single_beta) {
      const double* zw = a.zw + (f == 2 ? 3 : 2) * a.nz;
      d = sb_project(q + (long long)f * a.np, p, a.plane, a.nz, zw);
      if (m) d = d - sb_project(m + (long long)f * a.np, p, a.plane, a.nz, zw);
      else if (a.ref) d = d - a.ref[vo];
    } else {
      d = q[o];
      if (m) d = d - m[o];
      else if (a.ref) d = d - a.ref[vo];
    }
    w[vo] = a.scale * d;
    ++s;
  }
}

// Partial sums of H = X^T Y over row chunks.  A workgroup owns a 32x32 tile of
// H and the rows [chunk*rows_per, ...); 64-row slabs of X and Y are staged in
// LDS, each thread accumulates four entries of the tile.
constexpr int kDotTile = 32;
constexpr int kDotRows = 64;

__global__ __launch_bounds__(256) void block_dot_kernel(BlockDotArgs a) {
  __shared__ double xs[kDotRows][kDotTile + 1];
  __shared__ double ys[kDotRows][kDotTile + 1];
  const int tiles_j = (a.p + kDotTile - 1) / kDotTile;
  const int ti = blockIdx.y / tiles_j, tj = blockIdx.y % tiles_j;
  const int i0 = ti * kDotTile, j0 = tj * kDotTile;
  const long long r_begin = (long long)blockIdx.x * a.rows_per;
  const long long r_end = min(a.n, r_begin + a.rows_per);
  const int t = threadIdx.x;
  const int li = t >> 3, lj = (t & 7) * 4;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (long long r0 = r_begin; r0 < r_end; r0 += kDotRows) {
    // stage X[r0:r0+64, i0:i0+32] and Y[r0:r0+64, j0:j0+32]; zero outside
    for (int e = t; e < kDotRows * kDotTile; e += 256) {
      const int rr = e % kDotRows, cc = e / kDotRows;
      const long long r = r0 + rr;
      const bool rin = r < r_end;
      xs[rr][cc] = (rin && i0 + cc < a.k) ? a.X[(long long)(i0 + cc) * a.ldx + r] : 0.0;
      ys[rr][cc] = (rin && j0 + cc < a.p) ? a.Y[(long long)(j0 + cc) * a.ldy + r] : 0.0;
    }
    __syncthreads();
#pragma unroll 8
    for (int rr = 0; rr < kDotRows; ++rr) {
      const double x = xs[rr][li];
#pragma unroll
      for (int q = 0; q < 4; ++q) acc[q] = fma(x, ys[rr][lj + q], acc[q]);
    }
    __syncthreads();
  }
  const int i = i0 + li;
  if (i < a.k) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int j = j0 + lj + q;
      if (j < a.p) a.partial[((long long)blockIdx.x * a.p + j) * a.k + i] = acc[q];
    }
  }
}

// H(i,j) = sum over chunks in chunk order (deterministic).
__global__ __launch_bounds__(256) void block_dot_reduce_kernel(const double* partial, double* H, int nchunks,
                                                               long long kp) {
  const long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (e >= kp) return;
  double s = 0.0;
  for (int c = 0; c < nchunks; ++c) s += partial[(long long)c * kp + e];
  H[e] = s;
}

// Y[:, j] = alpha * Y[:, j] + beta * sum_i X[:, i] M(i, j) for a chunk of 32
// columns j; M is staged in LDS, each thread owns one row.
constexpr int kCombCols = 32;

__global__ __launch_bounds__(256) void block_combine_kernel(BlockCombineArgs a) {
  extern __shared__ double ms[];  // k x kCombCols
  const int j0 = blockIdx.y * kCombCols;
  const int nc = min(kCombCols, a.p - j0);
  for (int e = threadIdx.x; e < a.k * nc; e += blockDim.x) {
    const int i = e % a.k, jj = e / a.k;
    ms[jj * a.k + i] = a.M[(long long)(j0 + jj) * a.k + i];
  }
  __syncthreads();
  const long long r = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (r >= a.n) return;
  double acc[kCombCols];
#pragma unroll
  for (int jj = 0; jj < kCombCols; ++jj) acc[jj] = 0.0;
  for (int i = 0; i < a.k; ++i) {
    const double x = a.X[(long long)i * a.ldx + r];
#pragma unroll
    for (int jj = 0; jj < kCombCols; ++jj)
      if (jj < nc) acc[jj] = fma(x, ms[jj * a.k + i], acc[jj]);
  }
#pragma unroll
  for (int jj = 0; jj < kCombCols; ++jj) {
    if (jj < nc) {
      double* y = a.Y + (long long)(j0 + jj) * a.ldy + r;
      const double prev = a.alpha != 0.0 ? a.alpha * *y : 0.0;
      *y = prev + a.beta * acc[jj];
    }
  }
}

void launch_perturb(const PerturbArgs& a, int count, hipStream_t st) {
  const unsigned gx = (unsigned)((a.np + 255) / 256);
  hipLaunchKernelGGL(perturb_kernel, dim3(gx, count), dim3(256), 0, st, a);
}

void launch_collect(const CollectArgs& a, int count, hipStream_t st) {
  const unsigned gx = (unsigned)((a.nflow + 255) / 256);
  hipLaunchKernelGGL(collect_kernel, dim3(gx, count), dim3(256), 0, st, a);
}

int block_dot_chunks(long long n) {
  const long long c = (n + 8191) / 8192;
  return (int)std::max(1LL, std::min(c, 512LL));
}

void launch_block_dot(BlockDotArgs a, double* H, hipStream_t st) {
  const int nchunks = block_dot_chunks(a.n);
  a.rows_per = (a.n + nchunks - 1) / nchunks;
  const int tiles = ((a.k + kDotTile - 1) / kDotTile) * ((a.p + kDotTile - 1) / kDotTile);
  hipLaunchKernelGGL(block_dot_kernel, dim3(nchunks, tiles), dim3(256), 0, st, a);
  const long long kp = (long long)a.k * a.p;
  hipLaunchKernelGGL(block_dot_reduce_kernel, dim3((unsigned)((kp + 255) / 256)), dim3(256), 0, st,
                     (const double*)a.partial, H, nchunks, kp);
}

void launch_block_combine(const BlockCombineArgs& a, hipStream_t st) {
  const unsigned gx = (unsigned)((a.n + 255) / 256);
  const unsigned gy = (unsigned)((a.p + kCombCols - 1) / kCombCols);
  const size_t lds = sizeof(double) * (size_t)a.k * kCombCols;
  hipLaunchKernelGGL(block_combine_kernel, dim3(gx, gy), dim3(256), lds, st, a);
}

}  // namespace gatt
