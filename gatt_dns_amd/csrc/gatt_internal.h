// Internal declarations shared by the HIP translation units of libgatt_dns.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/gatt_dns.h"
#include "ragged.h"

namespace gatt {

constexpr int kMaxPtr = 8;   // fields per line-solve launch
constexpr int kFluxMaxNR = 6;   // widest derivative band (2nR-1 diagonals) of the flux-mode solves

// Device copy of one pre-factored operator (see gatt_lineop).
struct LineOpDev {
  int N = 0, ntypes = 0, nR = 0, periodic = 0;
  const double* A = nullptr;  // (N-1) x nt
  const double* B = nullptr;  // (N-periodic) x nt
  const double* C = nullptr;  // (N-1) x nt
  const double* D = nullptr;  // N x nt (periodic)
  const double* R = nullptr;  // N x (2nR-1) x nt
};

// Launch shape of the line solves along one direction: G lanes of a wavefront
// share a line, each owning M consecutive points (M a power of two, G*M >= N);
// a workgroup of waves_per_block wavefronts holds (64/G)*waves_per_block lines.
// variant 1 ("wave-segment", line_solve_ws.h): every lane owns a line, wavefront
// w owns points [w*M, w*M+M) of 64 lines, G = ceil(N/M) wavefronts per workgroup.
// variant 2 ("fused wave-segment", line_solve_wsp.h): the same layout with fused
// coefficient records and branch-free sweeps; variant 3: the same sweeps with
// 32 x lines per workgroup staged through an LDS tile; variant 4 ("x lane group", line_solve_xg.h):
// one wavefront per x line, M consecutive points per lane (G = 64).
struct LineShape {
  int variant = 0;
  int G = 64;
  int M = 4;
  int waves_per_block = 4;
};

// ---------------------------------------------------------------- pointwise constants
struct FlowConst {
  double gm1;        // gamma - 1
  double invRe;      // 1/Re
  double muA;        // 1 + 110/T0
  double muB;        // 110/T0
  double muC;        // gamma*(gamma-1)*Ma*Ma
  double qk;         // gamma/Re/Pr
};

// Viscous stresses and heat fluxes of nseq.F90:218-239 (3-D) and 72-85 (2-D): flux q
// (0 txx, 1 tyy, 2 tzz, 3 txy, 4 txz, 5 tyz, 6 qx, 7 qy, 8 qz) at point n from E and the
// first derivatives D + (3f + d) np.  The expressions are those of visc_kernel, so the
// value is bitwise the one that kernel stores.
__device__ __forceinline__ double visc_flux(int q, const FlowConst& c, double E, const double* D, long long n,
                                            long long np, int is3d) {
  const double mu1 = c.muA / (c.muC * E + c.muB);
  const double mu2 = c.muC * E;
  const double mu = mu1 * mu2 * sqrt(mu2);
  const double rmu = c.invRe * mu;
  if (q >= 6) return -(c.qk * mu * D[(12 + (q - 6)) * np + n]);
  if (q >= 3) {
    // txy: DUy + DVx, txz: DUz + DWx, tyz: DVz + DWy
    const int a = q == 3 ? 1 : (q == 4 ? 2 : 5), b = q == 3 ? 3 : (q == 4 ? 6 : 7);
    return rmu * (D[a * np + n] + D[b * np + n]);
  }
  const double DUx = D[0 * np + n], DVy = D[4 * np + n];
  const double own = q == 0 ? DUx : (q == 1 ? DVy : D[8 * np + n]);
  if (is3d) {
    const double tz = -(2.0 / 3.0 * (DUx + DVy + D[8 * np + n]));
    return rmu * (2.0 * own + tz);
  }
  const double dv = 2.0 / 3.0 * (DUx + DVy);
  return rmu * (2.0 * own - dv);
}

// Streaming accesses of arrays a kernel reads or writes once per point: the non-temporal
// hint lets the caches evict those lines first (the point-wise kernels of the step stream
// 20 to 50 full arrays per launch).
__device__ __forceinline__ double ld_nt(const double* p) { return __builtin_nontemporal_load(p); }
__device__ __forceinline__ void st_nt(double* p, double v) { __builtin_nontemporal_store(v, p); }
typedef double nt_double2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ nt_double2 ld_nt2(const double* p) {
  return __builtin_nontemporal_load(reinterpret_cast<const nt_double2*>(p));
}
__device__ __forceinline__ void st_nt2(double* p, double a, double b) {
  __builtin_nontemporal_store(nt_double2{a, b}, reinterpret_cast<nt_double2*>(p));
}

// Moves n values with a block-strided loop in batches of XB per thread: every load of a
// batch is issued before the first store needs its value, so each thread keeps XB loads
// in flight instead of one (the tile loads of the x-line kernels are latency-bound
// otherwise).  ld(i) returns value i, st(i, v) stores it.
template <int XB, typename Load, typename Store>
__device__ __forceinline__ void batched_move(int n, Load ld, Store st) {
  const int bs = blockDim.x;
  for (int b0 = threadIdx.x; b0 < n; b0 += XB * bs) {
    double v[XB];
#pragma unroll
    for (int k = 0; k < XB; ++k) {
      const int i = b0 + k * bs;
      if (i < n) v[k] = ld(i);
    }
#pragma unroll
    for (int k = 0; k < XB; ++k) {
      const int i = b0 + k * bs;
      if (i < n) st(i, v[k]);
    }
  }
}

struct SolveArgs {
  LineOpDev op;
  const int* types;           // 0-based type per line (a + n1*b)
  int dir;                    // 0 x, 1 y, 2 z
  int nx, ny, nz;
  int n1;                     // extent of the first line index
  int nlines;
  long long stride;           // element stride along the line
  int G;
  int TL;                     // lines per workgroup tile
  int lds_row;                // LDS words per point row of the tile
  float inv_n;                // 1/N
  int uniform;                // every 64-line tile carries one line type (transposed solves)
  int lds_coef_offset;        // wave-segment: LDS word offset of the per-wavefront coefficient slots
  int cs;                     // doubles per packed coefficient record (transposed solve)
  int nf;                     // fields per member
  int nmem;                   // members (fused wave-segment kernel)
  int xvec;                   // x lines 16-byte aligned (fused wave-segment kernel)
  int xtile;                  // x lines staged through an LDS tile: lines per workgroup (16 or 32), 0 = no tile
  int xg_m;                   // points per lane of the one-wavefront x solve (variant 4; its ragged form)
  long long in_stride;        // doubles between members in the input arrays
  long long out_stride;       // doubles between members in the output arrays
  const double* in[kMaxPtr];
  double* out[kMaxPtr];
  const double* coef;         // packed per-point coefficients [type][point][2nR-1+4] (wave-segment)
  const double* coef_ends;    // per type: A(1), C(1) of the periodic corrections
  const double* fcoef;        // fused per-point coefficients [type][point][2nR-1+3] (persistent wave-segment)
  const double* fcoef_ends;   // per type: A(1) B and C(1) B of the periodic corrections
  const double* fcoef_t;      // the fused records coefficient-major: [type][coefficient][point] (x, variant 4)
  const double* fcoef_p;      // the same lane-interleaved for M points per lane: [type][k][tt][lane] (variant 4)
  double theta;               // filter blend (mode 1)
  int mode;                   // 0 write result; 1 in-place filter with FilterCharTime blend
  // viscous-flux input (fused wave-segment kernels, mode 0): when has_flux is set, field f's
  // input at point n is visc_flux(flux_q[f], E[n], D, n) of member m (E + m in_stride,
  // D + m flux_d_stride) instead of in[f][n], so no stress array is stored
  int has_flux;
  int flux_q[kMaxPtr];
  const double* flux_E;
  const double* flux_D;
  long long flux_d_stride;
  FlowConst fc;
  int is3d;
};

template <int DIR>
void launch_line_solve_dir(const SolveArgs& a, const LineShape& s, int nmembers, hipStream_t st);
template <int DIR>
void launch_line_solve_ws_dir(const SolveArgs& a, const LineShape& s, int nmembers, hipStream_t st);
template <int DIR>
void launch_line_solve_wsp_dir(const SolveArgs& a, const LineShape& s, int nmembers, hipStream_t st);
// each direction of each variant is instantiated in its own translation unit
// (line_solve_{x,y,z}_{lg,ws,wsp}.hip)
extern template void launch_line_solve_dir<0>(const SolveArgs&, const LineShape&, int, hipStream_t);
extern template void launch_line_solve_dir<1>(const SolveArgs&, const LineShape&, int, hipStream_t);
extern template void launch_line_solve_dir<2>(const SolveArgs&, const LineShape&, int, hipStream_t);
extern template void launch_line_solve_ws_dir<0>(const SolveArgs&, const LineShape&, int, hipStream_t);
extern template void launch_line_solve_ws_dir<1>(const SolveArgs&, const LineShape&, int, hipStream_t);
extern template void launch_line_solve_ws_dir<2>(const SolveArgs&, const LineShape&, int, hipStream_t);
extern template void launch_line_solve_wsp_dir<0>(const SolveArgs&, const LineShape&, int, hipStream_t);
extern template void launch_line_solve_wsp_dir<1>(const SolveArgs&, const LineShape&, int, hipStream_t);
extern template void launch_line_solve_wsp_dir<2>(const SolveArgs&, const LineShape&, int, hipStream_t);
void launch_line_solve_xg(const SolveArgs& a, const LineShape& s, int nmembers, hipStream_t st);
// Lane-group solves of up to three directions (one SolveArgs each, with its own lanes per
// line a.G; one band width, points per lane and workgroup size, given by s) in one launch
// (line_solve_multi.hip); false, with nothing launched, where these are not compiled for it.
struct SolveMulti {
  SolveArgs a[3];
  unsigned first[4];   // first flat workgroup of each part; first[n] is the total
  unsigned gx[3];      // x extent of each part's own grid
  int n;
};
bool multi_solve_supported(int nR, int M);
bool launch_line_solve_multi(const SolveArgs* parts, int n, const LineShape& s, int nmembers, hipStream_t st);
inline void launch_line_solve(const SolveArgs& a, const LineShape& s, int nmembers, hipStream_t st) {
  if (s.variant == 4) {
    launch_line_solve_xg(a, s, nmembers, st);
    return;
  }
  if (s.variant == 2 || s.variant == 3) {
    SolveArgs b = a;
    b.xtile = s.variant == 3 ? 32 / (s.waves_per_block == 2 ? 2 : 1) : 0;   // lines per workgroup
    if (a.dir == 0) launch_line_solve_wsp_dir<0>(b, s, nmembers, st);
    else if (a.dir == 1) launch_line_solve_wsp_dir<1>(b, s, nmembers, st);
    else launch_line_solve_wsp_dir<2>(b, s, nmembers, st);
    return;
  }
  if (s.variant == 1) {
    if (a.dir == 0) launch_line_solve_ws_dir<0>(a, s, nmembers, st);
    else if (a.dir == 1) launch_line_solve_ws_dir<1>(a, s, nmembers, st);
    else launch_line_solve_ws_dir<2>(a, s, nmembers, st);
    return;
  }
  if (a.dir == 0) launch_line_solve_dir<0>(a, s, nmembers, st);
  else if (a.dir == 1) launch_line_solve_dir<1>(a, s, nmembers, st);
  else launch_line_solve_dir<2>(a, s, nmembers, st);
}
// A shape is usable when its register tile is compiled (M in 4..32), the
// segments cover the line, the wrapped halo stays within one period
// (M + nR - 1 <= N + 1) and the LDS tile fits in 64 KiB.
inline bool line_shape_ok(const LineShape& s, int N, int nR, int dir = 1, long long max_lds = 65536,
                          bool uniform_tiles = false) {
  if (!(s.M == 4 || s.M == 8 || s.M == 16 || s.M == 32) || nR < 2 || nR > 11) return false;
  if (s.variant == 4) {
    // one wavefront per x line, M points per lane
    // (16 points per lane only for bands up to 13 diagonals, which keep the window in registers)
    return dir == 0 && (s.M == 4 || s.M == 8 || (s.M == 16 && nR <= 7)) && 64LL * s.M >= N && nR - 1 <= N &&
           nR <= 11;
  }
  if (s.variant == 2 || s.variant == 3) {
    // fused wave-segment: as variant 1, every segment read straight from HBM
    // (variant 2) or, along x only, through an LDS tile (variant 3); LDS holds
    // the scan maps and coefficient slots (in the tile's words for variant 3)
    if (!(s.M == 16 || s.M == 32) || !uniform_tiles || (s.variant == 3 && dir != 0)) return false;
    // variant 2 may carry more segments than the line needs (empty ones past its end)
    const int G = s.variant == 2 && s.G > (N + s.M - 1) / s.M ? s.G : (N + s.M - 1) / s.M;
    const int gmax = s.variant == 3 ? 16 : (s.M >= 32 ? 8 : 16);   // variant 3: two segments per wavefront
    if (G < 1 || G > gmax || s.M + nR - 1 > N + 1 || nR - 1 > N) return false;
    const long long cs = (2LL * nR - 1 + 3 + 1) & ~1LL;
    if (s.variant == 3) {   // 32 (waves_per_block 1) or 16 (2) lines per workgroup through a tile of lw*N words
      const long long lw = s.waves_per_block == 2 ? 16 : 32, spw = 64 / lw;
      const long long GA = (G + spw - 1) / spw * spw;
      if (GA / spw > 8 || GA > 32) return false;   // at most 512 threads
      const long long work3 = 5LL * GA * lw + GA * (s.M * cs + 2), tile3 = lw * N;
      return 8 * (work3 > tile3 ? work3 : tile3) <= max_lds;
    }
    const long long work = 5LL * G * 64 + G * (s.M * cs + 2);
    return 8 * work <= max_lds;
  }
  if (s.variant == 1) {
    // wave-segment: compiled for 16 and 32 points per lane, and only for
    // directions whose 64-line tiles each carry one operator type
    if (!(s.M == 16 || s.M == 32) || !uniform_tiles) return false;
    const int G = (N + s.M - 1) / s.M;
    if (G < 1 || G > (s.M >= 32 ? 8 : 16) || s.M + nR - 1 > N + 1 || nR - 1 > N) return false;
    const long long cs = (2LL * nR - 1 + 4 + 1) & ~1LL;
    const long long work = 5LL * G * 64 + G * (s.M * cs + 2), tile = dir == 0 ? 65LL * N : 0;
    return 8 * (work > tile ? work : tile) <= max_lds;
  }
  if (s.G < 2 || s.G > 64 || (s.G & (s.G - 1)) || s.waves_per_block < 1 || s.waves_per_block > 4) return false;
  if (s.G * s.M < N || s.M + nR - 1 > N + 1 || nR - 1 > N) return false;
  const long long rows = (long long)(64 / s.G) * s.waves_per_block * s.G + 1;
  return (long long)s.M * rows * 8 <= 65536;
}
inline int pow2_at_least(int v) {
  int p = 1;
  while (p < v) p <<= 1;
  return p;
}

// ---------------------------------------------------------------- pointwise
struct ViscArgs {
  FlowConst c;
  long long np, member_stride;  // member stride of E
  long long work_stride;        // member stride of D and T
  const double* E;
  const double* D;   // 15 derivative arrays, D + (3f+d)*np
  double* T;         // 9 arrays: txx tyy tzz txy txz tyz qx qy qz
  int is3d;
};
void launch_visc(const ViscArgs& a, int nmembers, hipStream_t st);

struct AssembleArgs {
  FlowConst c;
  long long np, member_stride;  // member stride of S, Q, Out, Delta
  long long work_stride;        // member stride of D, T, G
  const double* S;      // stage state (RHS evaluated here), 5 arrays
  const double* Q;      // base state q^n, 5 arrays
  double* Out;          // stage output (may alias S), 5 arrays
  double* Delta;        // RK accumulator, 5 arrays
  const double* D;      // 15 first derivatives
  const double* T;      // 9 stresses / heat fluxes
  const double* G;      // 12 second derivatives (x: txx,txy,txz,qx | y: txy,tyy,tyz,qy | z: txz,tyz,tzz,qz)
  const double* sums;   // forcing sums per member (6 doubles) or nullptr
  double forcing_eps, forcing_e0;
  int stage;            // 0 raw RHS into Out, 1..4 RK4, 5 Euler, 6..8 SSPRK3
  double dt;
  int is3d;
  // pressureGradient forcing: d(pg_u) += pg_dpdx / pg_r on the 0-based box pg_box
  int pg;
  int pg_r, pg_u;
  int pg_box[6];
  double pg_dpdx;
  int nx, ny;
  // when set (np a multiple of 256): the block partials of the turbulence-forcing sums
  // of the values written to Out, 5 per block, member-major (the next stage's sums)
  double* sum_partial;
};
void launch_assemble(const AssembleArgs& a, int nmembers, hipStream_t st);

// The last direction's flux derivatives fused with the assembly and the stage update
// (rhs_line.hip): s describes the derivative operator and the lines of direction s.dir
// (1 or 2, fused records, one line type per 32-line tile); a is what assemble_kernel takes,
// its G holding the other directions' flux derivatives.  With a.sum_partial set, the
// forcing sums are written per workgroup: rhs_line_blocks(s.nlines) partials per member.
struct RhsLineArgs {
  SolveArgs s;
  AssembleArgs a;
};
bool rhs_line_fits(int N, int nR);
int rhs_line_blocks(int nlines);
void launch_rhs_line(const RhsLineArgs& r, int nmembers, hipStream_t st);

struct SumArgs {
  long long np, member_stride;
  const double* S;
  double* partial;      // nblocks x 6 per member
  double* sums;         // 6 per member: ET, U, V, W, E, N
  int nblocks;
};
void launch_forcing_sums(const SumArgs& a, int nmembers, hipStream_t st);
// the second stage of the forcing sums alone, over `nblocks` partials per member
void launch_forcing_final(const SumArgs& a, int nmembers, hipStream_t st);

struct SfdArgs {
  long long np, member_stride;
  double* Q;            // state, 5 arrays
  double* Mean;         // 5 arrays, or nullptr: Q is a perturbation about a frozen mean
  const double* chi;    // field (type 2) or nullptr
  double chi_val;
  double dt, delta;
  int nf_mask;          // bit f set => field f is damped
  int mean_ro;          // 1: the mean is read and updated in registers only (Mean is not written)
  int transpose;        // 1: the transpose of the blend on adjoints (Q: state adjoint, Mean: mean adjoint)
};
void launch_sfd(const SfdArgs& a, int nmembers, hipStream_t st);

struct ReduceArgs {
  long long np, member_stride;
  const double* Qn;     // new state
  const double* Q;      // old state
  unsigned long long* out;  // per member 8 slots (bit patterns): mc[5], umax[3]
  int is3d;
};
void launch_clear_u64(unsigned long long* p, int n, hipStream_t st);
void launch_clear_i32(int* p, int n, hipStream_t st);
// out[5*i + f] = field f of `state` at point idx[i] (W = 0 in 2-D); idx checked by the caller
void launch_gather_points(const double* state, long long np, int is3d, int npts, const long long* idx, double* out,
                          hipStream_t st);
void launch_step_reduce(const ReduceArgs& a, int nmembers, hipStream_t st);
void launch_copy(double* dst, const double* src, long long n, hipStream_t st);
// The step record: member 0's eight maxima (max changes, max |u|) as doubles into rec, the
// divergence flag raised over every member's, red cleared for the next step's reduction,
// and the clock (when given) advanced by dt (main.F90:424), in one single-thread launch.
void launch_record(unsigned long long* red, double* rec, int* flag, int nmembers, hipStream_t st,
                   double* clock = nullptr, double dt = 0.0);

// ------------------------------------------------ tangent and adjoint step
// Transposed Thomas sweeps of one direction (line_solve_t.hip); in may equal out.
void launch_line_solve_t(const SolveArgs& a, int nmembers, hipStream_t st);
int line_solve_t_points_per_lane(int N);
struct BandTArgs;
// y and z: the transposed solve and the transposed band in one launch; false when the
// line does not qualify (the caller then runs launch_line_solve_t and launch_band_t)
bool launch_line_solve_band_t(const SolveArgs& a, const BandTArgs& b, int nmembers, long long max_lds,
                              hipStream_t st);
bool line_solve_band_t_fits(int dir, int N, int nR, int nf, int nmembers, long long max_lds);

// Transposed RHS band along one direction:
//   out = beta*src + (accumulate bit f ? out : 0) + alpha * R^T in   (field f of the launch)
struct BandTArgs {
  int dir, nx, ny, nz, N, nR, cs;
  long long np;
  const double* coef;     // packed records [type][point][cs], band first
  const double* R;        // band rows as the reference stores them: N x (2nR-1) x ntypes
  const double* fcoef_t;  // fused records coefficient-major (x transposed solve)
  const double* fcoef_pt; // the transposed solve's records lane-interleaved (M = 4 for N <= 256, else 8)
  const double* fcoef_ends;
  int periodic, nlines;
  int uniform;            // every 64-line tile carries one line type (wave-uniform coefficients)
  const int* types;       // per line
  int nf;
  long long in_stride, out_stride;
  const double* in[kMaxPtr];
  double* out[kMaxPtr];
  const double* src[kMaxPtr];   // or nullptr
  double alpha, beta;
  int accumulate;         // bit f: field f adds to what out holds
};
void launch_band_t(const BandTArgs& a, int nmembers, hipStream_t st);
// Transposed solve and band along x in one launch (line_solve_xg.h); false if the
// line does not fit (the caller then runs launch_line_solve_t and launch_band_t).
bool launch_line_solve_xgt(const BandTArgs& a, int nmembers, hipStream_t st);

// Tangent of visc: dT from base E, D and tangents dE, dD.  Adjoint: accumulates
// Ebar and Dbar from Tbar (Dbar, Ebar accumulate; Tbar is read).
struct ViscLinArgs {
  FlowConst c;
  long long np, member_stride, work_stride, lin_stride;
  const double* E;      // base E (state stride)
  const double* D;      // base derivatives (work stride)
  const double* dE;     // tangent / adjoint E (lin stride)
  double* dEbar;        // adjoint: Ebar (lin stride)
  const double* dD;     // tangent dD (lin stride)
  double* dT;           // tangent output / adjoint input Tbar (lin stride)
  double* Dbar;         // adjoint output (lin stride)
  int is3d;
};
void launch_visc_tl(const ViscLinArgs& a, int nmembers, hipStream_t st);
void launch_visc_adj(const ViscLinArgs& a, int nmembers, hipStream_t st);

// Tangent of the RHS assembly (dd from dS, dD, dT, dG at base S, D, T, G) fused
// with the linear RK stage update of the tangent state; adjoint: from dbar,
// sets Sbar and writes Dbar, Tbar (stresses; heat fluxes zero) and Gbar.
struct AssembleLinArgs {
  FlowConst c;
  long long np, member_stride, work_stride, lin_stride;
  const double* S; const double* D; const double* T; const double* G;   // base
  // tangent
  const double* dS; const double* dD; const double* dT; const double* dG;
  const double* dQ; double* dOut; double* dDelta;
  int stage; double dt;
  // adjoint
  const double* dbar;   // 5 arrays (lin stride)
  double dbar_scale;    // the adjoint point-wise RHS reads dbar_scale * dbar (set by the caller)
  double* Sbar;         // 5 arrays, set (lin stride)
  double* Dbar; double* Tbar; double* Gbar;   // written (lin stride): Tbar 0-5, Gbar 4 per copy
  int gbar_copies;      // flux-derivative adjoints stored once for all directions (1) or per direction
  int is3d;
  // turbulenceGenerator (3-D): the base state's sums (ET, U, V, W, E, N; 8 per member of the
  // launch) and the linear sums of lin_sums (8 per member), or null without the forcing
  const double* sums;
  const double* lsums;
  double forcing_eps;
};
void launch_assemble_tl(const AssembleLinArgs& a, int nmembers, hipStream_t st);
// Global sums of a linearised field X against the base state S for the turbulence forcing:
// sums[8 m + 0..4] = sum(X_u S_u + X_v S_v + X_w S_w), sum X_u, sum X_v, sum X_w, sum X_e, and
// sums[8 m + 5] = np; deterministic two-stage reduction through `partial` (nblocks x 5 per member).
struct LinSumArgs {
  long long np, base_stride, lin_stride;
  const double* S;
  const double* X;
  double* partial;
  double* sums;
  int nblocks;
};
void launch_lin_sums(const LinSumArgs& a, int nmembers, hipStream_t st);
void launch_assemble_adj(const AssembleLinArgs& a, int nmembers, hipStream_t st);

// y[f] = a*x[f] + b*y[f] over nf arrays of np points per member (b == 0 writes a*x)
struct AxpbyArgs {
  long long np, x_stride, y_stride;
  int nf;
  const double* x[kMaxPtr];
  double* y[kMaxPtr];
  double a, b;
};
void launch_axpby(const AxpbyArgs& a, int nmembers, hipStream_t st);

// The field combinations between two adjoint RHS evaluations of the backward RK4
// pass, in one launch (the same operations, in the same order, as the separate
// y = a x + b y updates they replace):
//   mode 0:  acc = sb (+ acc unless first);  k = a2 sb + 1.0 (a1 qb)
//   mode 1:  qb = 1.0 sb + 1.0 (1.0 acc + 1.0 qb)
struct AdjCombineArgs {
  long long np, stride;   // member stride of all four arrays
  int nf;
  const double* sb[kMaxPtr];
  double* qb[kMaxPtr];
  double* acc[kMaxPtr];
  double* k[kMaxPtr];
  double a1, a2;
  int first, mode;
};
void launch_adj_combine(const AdjCombineArgs& a, int nmembers, hipStream_t st);

// ------------------------------------------------------- stability analysis
struct PerturbArgs {
  long long np, member_stride, nflow, ldv;
  double* state;        // state of the first member, 5 arrays
  double* mean;         // SFD mean of the first member or nullptr
  const double* base;   // base flow, 5 arrays
  const int* rmap;      // point -> flow-region index or -1 (nullptr: every point, identity)
  const double* V;      // Krylov vectors (column b for member b) or nullptr (state = base)
  double scale;
  int is3d;
  // singleBeta packing: rmap covers the z = 0 plane (plane points), and the value of
  // plane k is weighted by zw[k] = cos(z_k) (zw[nz + k] = sin(z_k) for W)
  int single_beta;
  long long plane;
  int nz;
  const double* zw;
};
void launch_perturb(const PerturbArgs& a, int count, hipStream_t st);

struct CollectArgs {
  long long np, member_stride, nflow, ldw;
  const double* plus;   // state of the first member
  const double* minus;  // state of the first subtracted member or nullptr
  const double* ref;    // subtracted vector (when minus is nullptr) or nullptr
  const int* idx;       // flow-region index -> point (nullptr: identity)
  double* W;
  double scale;
  int is3d;
  // singleBeta packing: idx points into the z = 0 plane, and a field is projected as
  // sum_k zw[2nz + k] q(plane k) (zw[3nz + k] for W), the weights cos(z_k) 2/nz, sin(z_k) 2/nz
  int single_beta;
  long long plane;
  int nz;
  const double* zw;
};
void launch_collect(const CollectArgs& a, int count, hipStream_t st);

struct BlockDotArgs {
  long long n, ldx, ldy, rows_per;
  int k, p;
  const double* X;
  const double* Y;
  double* partial;      // block_dot_chunks(n) x k x p
};
int block_dot_chunks(long long n);
void launch_block_dot(BlockDotArgs a, double* H, hipStream_t st);

struct BlockCombineArgs {
  long long n, ldx, ldy;
  int k, p;
  const double* X;
  const double* M;      // k x p, column-major, device
  double alpha, beta;
  double* Y;
};
void launch_block_combine(const BlockCombineArgs& a, hipStream_t st);

// --------------------------------------------------------------- boundaries
enum BcKind : int {
  BC_SET = 0,        // F = value
  BC_SET_R_FROM_P,   // R = value/(gamma-1)/E
  BC_NEU,            // F(plane) = sum c_k F(plane + s k)
  BC_NEU_P,          // R(plane) = (sum c_k (R E)(plane + s k)) / E(plane)
  BC_CORNER,         // corner averaging of extrapolations (E if adiabatic, then P)
  BC_LID,            // lidDrivenMovement on U
  BC_HOLD,           // holdInlet capture / restore
  BC_SINE,           // periodicSine on `field` (time dependent)
  BC_PACKET          // packet_2d on `field` (time dependent)
};

struct BcOp {
  int kind;
  int field;         // 0..4 (U V W R E)
  int box[6];        // 0-based inclusive
  int dir;           // 1..6 (Neumann) ; corners use cd
  int L;
  int coef;          // 0 neumannCoeffs, 1 neumann2Coeffs
  int cd[3];
  int adiabatic;
  int aux;           // offset into the aux table (lid x) / hold slot
  int hold_r;
  double value;
  long long npts;
  long long first_block;
  long long undo;    // offset of this op's 5 x npts saved values in the undo log
};

struct BcWave {
  int op_begin, op_end;
  long long nblocks;
};

struct BcArgs {
  const BcOp* ops;
  int op_begin, op_end;
  double* q[5];
  long long member_stride;
  const double* coef[2];
  const double* aux;
  double* hold;          // hold-inlet storage
  long long hold_member_stride;
  int nx, ny, nz;
  double gamma;
  int capture;           // holdInlet: 1 on the first call
  const double* clock;   // the time t of the step in progress (one double on the device)
  double toff;           // the ops see t = *clock + toff
  double* undo;          // when set, every op point saves its 5 field values here before the op
  long long undo_member_stride;
};
void launch_bc_wave(const BcArgs& a, long long nblocks, int nmembers, hipStream_t st);
// the device clock: set to t, or advanced by dt as main.F90:424 does (t = t + dt)
void launch_clock_set(double* clock, double t, hipStream_t st);
void launch_clock_add(double* clock, double dt, hipStream_t st);


// Tangent / adjoint of one wave of the boundary program (boundary.hip).
struct BcLinArgs {
  const BcOp* ops;
  int op_begin, op_end;
  double* q[5];          // base state (updated by the tangent kernel, read by the adjoint)
  long long member_stride;
  double* dq[5];         // tangent or adjoint fields
  long long lin_stride;
  const double* coef[2];
  const double* aux;
  const double* hold;
  long long hold_member_stride;
  int nx, ny, nz;
  double gamma;
  // adjoint only: when set, every op point first puts back the five values its wave
  // overwrote (undo log of BcArgs::undo), so the base is the state before the wave
  const double* undo;
  long long undo_member_stride;
};
void launch_bc_wave_tl(const BcLinArgs& a, long long nblocks, int nmembers, hipStream_t st);
void launch_bc_wave_adj(const BcLinArgs& a, long long nblocks, int nmembers, hipStream_t st);

}  // namespace gatt
