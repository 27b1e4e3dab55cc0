// C ABI of libgatt_dns (include/gatt_dns.h): context set-up, the boundary
// program compiler, and the orchestration of the time step on one HIP stream.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "gatt_internal.h"

using namespace gatt;

namespace gatt {
thread_local LaunchRecorder* g_recorder = nullptr;
}

// Scratch of the allocation-free block product (gatt_block_dot_noalloc): the per-chunk
// partial sums and the reduced k x p block.  Each device holds its own copy for the
// life of the process; host code reaches it only through the address hipGetSymbolAddress
// returns on that device (kept in g_scratch below), never through this name.
constexpr size_t kStaticScratchDoubles = size_t(1) << 20;   // 8 MiB
static __device__ __align__(256) double scratch[kStaticScratchDoubles];

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIPCHK(expr)                                                              \
  do {                                                                            \
    hipError_t e_ = (expr);                                                       \
    if (e_ != hipSuccess)                                                         \
      return fail(GATT_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

// device address of `scratch` per device, looked up once with that device current
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch[kMaxDevices] = {};

int init_static_scratch(int dev) {
  if (dev < 0 || dev >= kMaxDevices) return fail(GATT_E_UNSUPPORTED, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch[dev]) return GATT_OK;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p)
    return fail(GATT_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  g_scratch[dev] = static_cast<double*>(p);
  return GATT_OK;
}

// Per-member slab: arrays of np doubles.
enum : int {
  SL_QA = 0,    // state ping
  SL_QB = 5,    // state pong
  SL_S = 10,    // RK stage state
  SL_DL = 15,   // RK accumulator
  SL_MEAN = 20, // SFD mean (allocated only when SFD is on)
  SL_N_NOSFD = 20,
  SL_N_SFD = 25
};

// Derivative workspace of one member of a work group: arrays of np doubles.
enum : int {
  WK_D = 0,     // 15 first derivatives, (3f + d)
  WK_T = 15,    // 9 stresses / heat fluxes
  WK_G = 24,    // 12 second derivatives
  WK_N = 36
};

// Members whose RHS is evaluated in the same launches share the workspace;
// groups are sized so that the workspace covers about this many points.
constexpr long long kWorkPoints = 1LL << 23;

struct DirOps {
  LineOpDev op[2];           // deriv, filter
  double* coef[2] = {nullptr, nullptr};       // packed per-point coefficients
  double* coef_ends[2] = {nullptr, nullptr};  // per type A(1), C(1)
  double* fcoef[2] = {nullptr, nullptr};      // fused per-point coefficients (line_solve_wsp.h)
  double* fcoef_ends[2] = {nullptr, nullptr}; // per type A(1) B(2nd row), C(1) B(last row)
  double* fcoef_t[2] = {nullptr, nullptr};    // fused records coefficient-major (x lane-group solves)
  double* fcoef_p[2][3] = {{nullptr, nullptr, nullptr}, {nullptr, nullptr, nullptr}};  // lane-interleaved, M = 4, 8, 16
  double* fcoef_pt[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};  // transposed solve, lane-interleaved, M = 4, 8
  LineShape shape[2];
  int* types = nullptr;      // per line, 0-based
  bool uniform_tiles = false;  // every tile of 64 consecutive lines has one type
  bool identity[2] = {false, false};  // the operator is the identity (no line solve)
  int n1 = 0, nlines = 0;
  long long stride = 0;
};

struct Box { int f; int b[6]; };

bool overlap(const Box& a, const Box& b) {
  if (a.f != b.f) return false;
  for (int d = 0; d < 3; ++d)
    if (a.b[2 * d + 1] < b.b[2 * d] || b.b[2 * d + 1] < a.b[2 * d]) return false;
  return true;
}

}  // namespace

struct gatt_ctx {
  int device = 0;              // the device current at gatt_create
  int nx = 0, ny = 0, nz = 0, is3d = 0, batch = 1;
  long long np = 0;
  long long slab = 0;          // doubles per member
  double* mem = nullptr;       // batch * slab
  int wgroup = 1;              // members per work group
  long long work_stride = 0;   // doubles per work-group member
  double* work = nullptr;      // wgroup * work_stride
  FlowConst fc{};
  double gamma = 1.4;
  DirOps dir[3];
  std::vector<void*> dev_allocs;
  long long max_lds = 65536;
  int filterDir[3] = {1, 1, 1};
  double filterCharTime = -1;
  int stepping = 1;     // timeStepping: 1 RK4, 2 Euler, 3 SSPRK3
  double* d_clock = nullptr;   // time of the step in progress (main.F90:424 accumulates it)
  bool timed = false;          // the boundary program has time-dependent disturbances
  long long undo_total = 0;    // doubles per member of the boundary program's undo log
  double* d_undo = nullptr;    // undo log of the transposed boundary program (allocated on first use)
  // pressureGradient forcing (runForcings{2,3}D.F90)
  int pg = 0, pg_r = 0, pg_u = 0;
  int pg_box[6] = {0, 0, 0, 0, 0, 0};
  double pg_dpdx = 0;
  // boundary program
  BcOp* d_ops = nullptr;
  std::vector<BcOp> ops;
  std::vector<BcWave> waves;
  double* d_coef[2] = {nullptr, nullptr};
  double* d_aux = nullptr;
  double* d_hold = nullptr;
  long long hold_stride = 0;
  bool hold_captured = false;
  // forcing
  int forcing = 0;
  double f_eps = 0, f_e0 = 0;
  double* d_partial = nullptr;
  void* d_points = nullptr;       // gatt_get_points: point indices then gathered values
  int points_cap = 0;
  double* d_apartial = nullptr;    // forcing-sum partials written by the kernel that wrote the stage state
  int ablocks = 0;                 // assemble blocks per member (np % 256 == 0, else 0)
  int apart_blocks = 0;            // partials per member the last producer wrote
  bool fuse_rhs = false;           // y flux derivatives, assembly and stage update in one kernel
  bool sums_from_assemble = false; // the last assemble left the next stage's partials there
  double* d_sums = nullptr;
  double* d_lsums = nullptr;       // linear sums of the linearised forcing (8 per member)
  int sum_blocks = 0;
  // SFD
  int sfdType = 0;
  double sfdDelta = 0, sfdXval = 0;
  double* d_chi = nullptr;
  bool mean_ready = false;
  // gatt_checkpoint: the state and SFD mean of every member (allocated on first use)
  double* d_ckpt = nullptr;
  bool ckpt_mean_ready = false;
  // reductions
  unsigned long long* d_red = nullptr;  // batch * 8
  double* d_rec = nullptr;              // 8: member 0's maxima of the last recorded step
  int* d_flag = nullptr;
  int cur = 0;                          // 0: state in QA, 1: in QB
  hipStream_t stream = nullptr;
  // the three direction solves of an RHS pass run as parallel branches on these
  // streams, forked from and joined back into `stream`
  hipStream_t aux[3] = {nullptr, nullptr, nullptr};
  hipEvent_t ev_fork = nullptr;
  hipEvent_t ev_join[3] = {nullptr, nullptr, nullptr};
  bool fork_dirs = true;
  bool multi_dirs = true;   // serial direction solves of one lane-group shape as one launch
  // tangent / adjoint arrays, per member (allocated on first use)
  double* lin = nullptr;
  long long lin_stride = 0;
  // adjoint: derivatives, stresses and their derivatives of the four RK stages of
  // the base step, per member (allocated on first use)
  double* tape = nullptr;
  long long tape_stride = 0;
  // stability analysis
  long long nflow = 0;                  // flow-region points (0: not set, every point)
  int* d_idx = nullptr;                 // region index -> point
  int* d_rmap = nullptr;                // point -> region index or -1
  bool single_beta = false;             // singleBeta packing of the Krylov vectors
  double* d_zw = nullptr;               // its z weights: cos, sin, 2/nz cos, 2/nz sin
  double* d_base = nullptr;             // base flow, 5 arrays
  double* d_scratch = nullptr;          // block products: partial sums, H, M
  size_t scratch_doubles = 0;
  // captured double step
  hipGraphExec_t graph = nullptr;
  double graph_dt = 0;
  int graph_cur = -1;
  // captured adjoint step (its base state buffer, time and step size)
  hipGraphExec_t adj_graph = nullptr;
  double adj_graph_t = 0, adj_graph_dt = 0;
  int adj_graph_cur = -1;
  void drop_graphs() {
    if (graph) (void)hipGraphExecDestroy(graph);
    if (adj_graph) (void)hipGraphExecDestroy(adj_graph);
    graph = nullptr;
    adj_graph = nullptr;
  }

  double* slab_ptr(int idx) { return mem + (long long)idx * np; }
  double* work_ptr(int idx) { return work + (long long)idx * np; }
  double* state(int which) { return slab_ptr(which == 0 ? SL_QA : SL_QB); }
};

namespace {

template <typename T>
int dev_upload(gatt_ctx* c, const T* host, size_t count, T** out) {
  *out = nullptr;
  if (count == 0) return GATT_OK;
  void* p = nullptr;
  HIPCHK(hipMalloc(&p, count * sizeof(T)));
  c->dev_allocs.push_back(p);
  HIPCHK(hipMemcpy(p, host, count * sizeof(T), hipMemcpyHostToDevice));
  *out = static_cast<T*>(p);
  return GATT_OK;
}

// Default shape before tuning.  Many lines of at most 256 points: the fused
// wave-segment kernel (along x through the 16-line tile); long x lines: one wavefront
// per line; otherwise the lane-group kernel with the most lanes per line (the
// shortest stretch per lane, one wavefront per workgroup: the most workgroups), whose
// band sums follow the reference's order term by term.
LineShape choose_shape(int dir, int N, int nR, long long nlines, long long max_lds, bool uniform_tiles) {
  if (nlines >= 1024 && uniform_tiles) {
    // x: the 16-line LDS tile (unit-stride loads and stores; more workgroups per compute
    // unit than the 32-line tile) with 16 points per lane;
    // y and z: 32 points per lane for wide bands (fewer, fuller wavefronts), else 16
    const bool wide = dir == 0 || nR >= 5;
    for (int v : {3, 2, 1}) {
      if (v == 3 && dir != 0) continue;
      for (int M : {(wide && v != 3) ? 32 : 16, (wide && v != 3) ? 16 : 32}) {
        for (int wpb : {2, 1}) {
          if (v != 3 && wpb == 2) continue;
          LineShape s;
          s.variant = v;
          s.M = M;
          s.G = (N + M - 1) / M;
          if (v == 3) s.waves_per_block = wpb;
          if (line_shape_ok(s, N, nR, dir, max_lds, uniform_tiles)) return s;
        }
      }
    }
  }
  if (dir == 0 && N >= 256) {
    // long x lines whose tiles mix line types: one wavefront per line, the fewest points
    // per lane that cover it
    for (int M : {4, 8, 16}) {
      LineShape s;
      s.variant = 4;
      s.M = M;
      s.G = 64;
      s.waves_per_block = 4;
      if (line_shape_ok(s, N, nR, dir, max_lds, uniform_tiles)) return s;
    }
  }
  // the lane-group kernel with as many lanes per line as the line can use (the shortest
  // sequential stretch per lane) and one wavefront per workgroup (more workgroups)
  for (int M : {4, 8, 16, 32}) {
    for (int wpb : {1, 2, 4}) {
      LineShape s;
      s.M = M;
      s.G = pow2_at_least((N + M - 1) / M);
      if (s.G < 2) s.G = 2;
      s.waves_per_block = wpb;
      if (line_shape_ok(s, N, nR)) return s;
    }
  }
  LineShape bad;
  bad.M = 0;
  return bad;
}

// Per-point coefficients of the wave-segment kernel, [type][point][CS] with
// CS = 2nR-1+4 rounded up to even:
// the RHS band, then fa, fb of the forward sweep d = (r - fa d') fb, the
// backward factor ce of x = d - ce x', and D of the periodic first row.  The
// periodic index shift of B (derivs.F90:796-798) is resolved here.
int pack_coeffs(gatt_ctx* c, const gatt_lineop& h, double** coef, double** ends);

int upload_lineop(gatt_ctx* c, const gatt_lineop& h, int expectN, LineOpDev* out, const char* what) {
  if (h.n != expectN) return fail(GATT_E_INVALID, std::string(what) + ": line length does not match the grid");
  if (h.nR < 1 || h.nR > 11) return fail(GATT_E_UNSUPPORTED, std::string(what) + ": nR outside 1..11");
  if (h.n < 2 * h.nR - 1) return fail(GATT_E_INVALID, std::string(what) + ": line shorter than the stencil");
  if (!h.A || !h.B || !h.C || !h.R || (h.periodic && !h.D))
    return fail(GATT_E_INVALID, std::string(what) + ": missing coefficient array");
  out->N = h.n;
  out->ntypes = h.ntypes;
  out->nR = h.nR;
  out->periodic = h.periodic ? 1 : 0;
  const size_t nt = (size_t)h.ntypes, N = (size_t)h.n;
  double* p;
  int rc;
  if ((rc = dev_upload(c, h.A, (N - 1) * nt, &p))) return rc;
  out->A = p;
  if ((rc = dev_upload(c, h.B, (N - out->periodic) * nt, &p))) return rc;
  out->B = p;
  if ((rc = dev_upload(c, h.C, (N - 1) * nt, &p))) return rc;
  out->C = p;
  if (out->periodic) {
    if ((rc = dev_upload(c, h.D, N * nt, &p))) return rc;
    out->D = p;
  }
  if ((rc = dev_upload(c, h.R, N * (size_t)(2 * h.nR - 1) * nt, &p))) return rc;
  out->R = p;
  return GATT_OK;
}

int pack_coeffs(gatt_ctx* c, const gatt_lineop& h, double** coef, double** ends) {
  const int N = h.n, nt = h.ntypes, NW = 2 * h.nR - 1, CS = (NW + 4 + 1) & ~1, per = h.periodic ? 1 : 0;
  const int NB = N - per;
  std::vector<double> pk((size_t)nt * N * CS, 0.0), en((size_t)nt * 2, 0.0);
  for (int t = 0; t < nt; ++t) {
    for (int e = 0; e < N; ++e) {
      double* p = &pk[((size_t)t * N + e) * CS];
      for (int q = 0; q < NW; ++q) p[q] = h.R[(size_t)e + (size_t)N * (q + (size_t)NW * t)];
      p[NW] = e > per ? h.A[(size_t)(e - 1) + (size_t)(N - 1) * t] : 0.0;
      if (per) p[NW + 1] = e >= 1 ? h.B[(size_t)(e - 1) + (size_t)NB * t] : 0.0;
      else p[NW + 1] = h.B[(size_t)e + (size_t)NB * t];
      p[NW + 2] = e <= N - 2 ? h.C[(size_t)e + (size_t)(N - 1) * t] : 0.0;
      p[NW + 3] = per ? h.D[(size_t)e + (size_t)N * t] : 0.0;
    }
    en[2 * t] = h.A[(size_t)(N - 1) * t];
    en[2 * t + 1] = h.C[(size_t)(N - 1) * t];
  }
  int rc;
  if ((rc = dev_upload(c, pk.data(), pk.size(), coef))) return rc;
  return dev_upload(c, en.data(), en.size(), ends);
}

// Fused records of the branch-free wave-segment kernel, [type][point][FS] with
// FS = 2nR-1+3 rounded up to even: the band row times the forward factor B of
// the point, fa' = A B, ce = C, D' = D / B.  The periodic row 0 is not part of
// the Thomas system: its band row stays unscaled (only its D . r term is used)
// and its forward factor is zero, as is that of the first solved row; the last
// row's backward factor is zero.  The end factors are A(1) and C(1) times the B
// of the rows they correct (rows 2 and N).
int pack_coeffs_fused(gatt_ctx* c, const gatt_lineop& h, double** coef, double** ends, double** coef_t,
                      double** coef_p, double** coef_pt) {
  const int N = h.n, nt = h.ntypes, NW = 2 * h.nR - 1, FS = (NW + 3 + 1) & ~1, per = h.periodic ? 1 : 0;
  const int NB = N - per;
  std::vector<double> pk((size_t)nt * N * FS, 0.0), en((size_t)nt * 2, 0.0);
  for (int t = 0; t < nt; ++t) {
    auto Bof = [&](int e) {   // forward factor of row e (periodic: B is shifted by one row)
      if (per) return e >= 1 ? h.B[(size_t)(e - 1) + (size_t)NB * t] : 1.0;
      return h.B[(size_t)e + (size_t)NB * t];
    };
    for (int e = 0; e < N; ++e) {
      double* p = &pk[((size_t)t * N + e) * FS];
      const double b = Bof(e);
      for (int q = 0; q < NW; ++q) p[q] = h.R[(size_t)e + (size_t)N * (q + (size_t)NW * t)] * b;
      p[NW] = e > per ? h.A[(size_t)(e - 1) + (size_t)(N - 1) * t] * b : 0.0;
      p[NW + 1] = e <= N - 2 ? h.C[(size_t)e + (size_t)(N - 1) * t] : 0.0;
      p[NW + 2] = per ? h.D[(size_t)e + (size_t)N * t] / b : 0.0;
    }
    if (per) {
      en[2 * t] = h.A[(size_t)(N - 1) * t] * Bof(1);
      en[2 * t + 1] = h.C[(size_t)(N - 1) * t] * Bof(N - 1);
    }
  }
  // the same records coefficient-major, [type][k][point] for k < 2nR-1+3
  const int NK = NW + 3;
  std::vector<double> tr((size_t)nt * NK * N, 0.0);
  for (int t = 0; t < nt; ++t)
    for (int k = 0; k < NK; ++k)
      for (int e = 0; e < N; ++e) tr[((size_t)t * NK + k) * N + e] = pk[((size_t)t * N + e) * FS + k];
  int rc;
  if ((rc = dev_upload(c, pk.data(), pk.size(), coef))) return rc;
  if ((rc = dev_upload(c, tr.data(), tr.size(), coef_t))) return rc;
  // lane-interleaved copies for one wavefront per line with M points per lane
  // (M = 4, 8, 16): [type][k][tt][lane] holds point lane*M + tt, zero past the line,
  // so that a wavefront reads each coefficient of its tt-th points as 64 consecutive doubles
  const int Ms[3] = {4, 8, 16};
  for (int mi = 0; mi < 3; ++mi) {
    const int M = Ms[mi];
    coef_p[mi] = nullptr;
    if (64 * M < N) continue;
    std::vector<double> il((size_t)nt * NK * 64 * M, 0.0);
    for (int t = 0; t < nt; ++t)
      for (int k = 0; k < NK; ++k)
        for (int tt = 0; tt < M; ++tt)
          for (int l = 0; l < 64; ++l) {
            const int e = l * M + tt;
            if (e < N) il[(((size_t)t * NK + k) * M + tt) * 64 + l] = tr[((size_t)t * NK + k) * N + e];
          }
    if ((rc = dev_upload(c, il.data(), il.size(), &coef_p[mi]))) return rc;
  }
  // the records of the transposed solve, lane-interleaved for M = 4, 8: for output point
  // j, rows 0..2nR-2 hold the band column j of offset l = row - (nR-1), taken from row
  // j - l (circularly); then the shifted sweep factors ce(j-1) (zero for j <= periodic),
  // fa'(j+1), and D(j)
  const int H = h.nR - 1;
  for (int mi = 0; mi < 2; ++mi) {
    const int M = Ms[mi];
    coef_pt[mi] = nullptr;
    if (64 * M < N) continue;
    std::vector<double> il((size_t)nt * NK * 64 * M, 0.0);
    auto at = [&](int t, int k, int e) { return tr[((size_t)t * NK + k) * N + e]; };
    for (int t = 0; t < nt; ++t)
      for (int tt = 0; tt < M; ++tt)
        for (int l = 0; l < 64; ++l) {
          const int j = l * M + tt;
          if (j >= N) continue;
          auto put = [&](int k, double v) { il[(((size_t)t * NK + k) * M + tt) * 64 + l] = v; };
          for (int o = -H; o <= H; ++o) {
            int se = j - o;
            se = se < 0 ? se + N : (se >= N ? se - N : se);
            put(H + o, at(t, H + o, se));
          }
          put(NW, j > per ? at(t, NW + 1, j - 1) : 0.0);
          put(NW + 1, j + 1 < N ? at(t, NW, j + 1) : 0.0);
          put(NW + 2, at(t, NW + 2, j));
        }
    if ((rc = dev_upload(c, il.data(), il.size(), &coef_pt[mi]))) return rc;
  }
  return dev_upload(c, en.data(), en.size(), ends);
}

// ------------------------------------------------------------ BC compiler
struct OpSets {
  std::vector<Box> rd, wr;
};

int compile_bc(gatt_ctx* c, const gatt_bc_desc& bc) {
  const int nx = c->nx, ny = c->ny, nz = c->nz;
  const int dims[3] = {nx, ny, nz};
  std::vector<BcOp> ops;
  std::vector<OpSets> sets;
  std::vector<double> aux;
  long long hold_total = 0;
  auto in_range = [&](const int b[6]) {
    for (int d = 0; d < 3; ++d)
      if (b[2 * d] < 0 || b[2 * d + 1] >= dims[d] || b[2 * d] > b[2 * d + 1]) return false;
    return true;
  };
  auto add = [&](BcOp op, OpSets s) -> int {
    if (!in_range(op.box)) return fail(GATT_E_INVALID, "boundary box outside the grid");
    for (auto& b : s.rd)
      if (!in_range(b.b)) return fail(GATT_E_INVALID, "boundary stencil outside the grid");
    op.npts = (long long)(op.box[1] - op.box[0] + 1) * (op.box[3] - op.box[2] + 1) * (op.box[5] - op.box[4] + 1);
    op.undo = c->undo_total;
    c->undo_total += 5 * op.npts;
    ops.push_back(op);
    sets.push_back(s);
    return GATT_OK;
  };
  auto box_of = [](const gatt_box_list& bl, int i, int b[6]) {
    for (int q = 0; q < 6; ++q) b[q] = bl.box[i + bl.count * q] - 1;
  };
  // field index of the BC slots U,V,W,P,E (P acts on R)
  const int slot_field[5] = {GATT_U, GATT_V, GATT_W, GATT_R, GATT_E};
  auto skip_field = [&](int f) { return !c->is3d && f == GATT_W; };
  // stencil box along Neumann direction d from a plane box
  auto stencil = [](const int b[6], int d, int L, int out[6]) {
    for (int q = 0; q < 6; ++q) out[q] = b[q];
    switch (d) {
      case 1: out[1] = b[0] + L; break;
      case 2: out[0] = b[0] - L; out[1] = b[0]; break;
      case 3: out[3] = b[2] + L; break;
      case 4: out[2] = b[2] - L; out[3] = b[2]; break;
      case 5: out[5] = b[4] + L; break;
      case 6: out[4] = b[4] - L; out[5] = b[4]; break;
    }
  };
  // the points a one-sided extrapolation from a plane box reads: 1..L steps along direction d
  // (extrap in boundary.hip; the plane itself is not read)
  auto stencil_ex = [](const int b[6], int d, int L, int out[6]) {
    for (int q = 0; q < 6; ++q) out[q] = b[q];
    switch (d) {
      case 1: out[0] = b[0] + 1; out[1] = b[0] + L; break;
      case 2: out[0] = b[0] - L; out[1] = b[0] - 1; break;
      case 3: out[2] = b[2] + 1; out[3] = b[2] + L; break;
      case 4: out[2] = b[2] - L; out[3] = b[2] - 1; break;
      case 5: out[4] = b[4] + 1; out[5] = b[4] + L; break;
      case 6: out[4] = b[4] - L; out[5] = b[4] - 1; break;
    }
  };
  auto plane = [](const int b[6], int d, int out[6]) {
    for (int q = 0; q < 6; ++q) out[q] = b[q];
    const int ax = (d - 1) / 2;
    out[2 * ax + 1] = out[2 * ax];
  };
  auto thin = [](const int b[6], int d) {
    const int ax = (d - 1) / 2;
    return b[2 * ax] == b[2 * ax + 1];
  };
  int rc;
  // Dirichlet: U, V, W, E (nseq.F90:359-370)
  const int dord[4] = {GATT_BC_U, GATT_BC_V, GATT_BC_W, GATT_BC_E};
  for (int s : dord) {
    const gatt_box_list& bl = bc.dir[s];
    const int f = slot_field[s];
    if (skip_field(f)) continue;
    for (int i = 0; i < bl.count; ++i) {
      BcOp op{};
      op.kind = BC_SET;
      op.field = f;
      box_of(bl, i, op.box);
      op.value = bl.value[i];
      OpSets st;
      Box w{f, {}};
      std::memcpy(w.b, op.box, sizeof(w.b));
      st.wr.push_back(w);
      if ((rc = add(op, st))) return rc;
    }
  }
  // Dirichlet pressure (nseq.F90:371-373)
  {
    const gatt_box_list& bl = bc.dir[GATT_BC_P];
    for (int i = 0; i < bl.count; ++i) {
      BcOp op{};
      op.kind = BC_SET_R_FROM_P;
      op.field = GATT_R;
      box_of(bl, i, op.box);
      op.value = bl.value[i];
      OpSets st;
      Box r{GATT_E, {}}, w{GATT_R, {}};
      std::memcpy(r.b, op.box, sizeof(r.b));
      std::memcpy(w.b, op.box, sizeof(w.b));
      st.rd.push_back(r);
      st.wr.push_back(w);
      if ((rc = add(op, st))) return rc;
    }
  }
  auto neumann_lists = [&](const gatt_box_list* lists, int L, int coef) -> int {
    for (int s : dord) {
      const gatt_box_list& bl = lists[s];
      const int f = slot_field[s];
      if (skip_field(f)) continue;
      for (int i = 0; i < bl.count; ++i) {
        int b[6];
        box_of(bl, i, b);
        BcOp op{};
        op.kind = BC_NEU;
        op.field = f;
        op.dir = bl.dir[i];
        if (op.dir < 1 || op.dir > 6) return fail(GATT_E_INVALID, "Neumann direction outside 1..6");
        op.L = L;
        op.coef = coef;
        plane(b, op.dir, op.box);
        OpSets st;
        Box r{f, {}}, w{f, {}};
        stencil_ex(op.box, op.dir, L, r.b);
        std::memcpy(w.b, op.box, sizeof(w.b));
        st.rd.push_back(r);
        st.wr.push_back(w);
        if ((rc = add(op, st))) return rc;
      }
    }
    const gatt_box_list& bl = lists[GATT_BC_P];
    for (int i = 0; i < bl.count; ++i) {
      int b[6];
      box_of(bl, i, b);
      BcOp op{};
      op.kind = BC_NEU_P;
      op.field = GATT_R;
      op.dir = bl.dir[i];
      if (op.dir < 1 || op.dir > 6) return fail(GATT_E_INVALID, "Neumann direction outside 1..6");
      if (!thin(b, op.dir)) return fail(GATT_E_UNSUPPORTED, "pressure Neumann box thicker than one plane");
      op.L = L;
      op.coef = coef;
      plane(b, op.dir, op.box);
      OpSets st;
      // P = R E along the stencil, then R = P / E at the plane
      Box r1{GATT_R, {}}, r2{GATT_E, {}}, w{GATT_R, {}};
      stencil_ex(op.box, op.dir, L, r1.b);
      stencil(op.box, op.dir, L, r2.b);
      std::memcpy(w.b, op.box, sizeof(w.b));
      st.rd.push_back(r1);
      st.rd.push_back(r2);
      st.wr.push_back(w);
      if ((rc = add(op, st))) return rc;
    }
    return GATT_OK;
  };
  // Neumann (nseq.F90:377-399)
  if ((rc = neumann_lists(bc.neu, bc.neumannLength, 0))) return rc;
  // corners (nseq.F90:402-488)
  for (int i = 0; i < bc.nCorners; ++i) {
    BcOp op{};
    op.kind = BC_CORNER;
    op.field = GATT_R;
    for (int q = 0; q < 6; ++q) op.box[q] = bc.cL[i + bc.nCorners * q] - 1;
    for (int q = 0; q < 3; ++q) op.cd[q] = bc.cD[i + bc.nCorners * q];
    op.adiabatic = bc.cAdiabatic[i] == 1;
    op.L = bc.neumannLength;
    if (op.cd[0] == 0 && op.cd[1] == 0 && op.cd[2] == 0) return fail(GATT_E_INVALID, "corner without direction");
    OpSets st;
    int bb[6];
    std::memcpy(bb, op.box, sizeof(bb));
    for (int ax = 0; ax < 3; ++ax) {
      if (op.cd[ax] == 0) continue;
      const int d = op.cd[ax] == 1 ? 2 * ax + 1 : 2 * ax + 2;
      if (!thin(op.box, d)) return fail(GATT_E_UNSUPPORTED, "corner box thicker than one point across a wall");
      int sb[6];
      stencil(op.box, d, op.L, sb);
      for (int q = 0; q < 3; ++q) {
        bb[2 * q] = std::min(bb[2 * q], sb[2 * q]);
        bb[2 * q + 1] = std::max(bb[2 * q + 1], sb[2 * q + 1]);
      }
    }
    Box r1{GATT_R, {}}, r2{GATT_E, {}}, w1{GATT_R, {}}, w2{GATT_E, {}};
    std::memcpy(r1.b, bb, sizeof(bb));
    std::memcpy(r2.b, bb, sizeof(bb));
    std::memcpy(w1.b, op.box, sizeof(bb));
    std::memcpy(w2.b, op.box, sizeof(bb));
    st.rd.push_back(r1);
    st.rd.push_back(r2);
    st.wr.push_back(w1);
    if (op.adiabatic) st.wr.push_back(w2);
    if ((rc = add(op, st))) return rc;
  }
  // null second derivative (nseq.F90:491-513)
  if ((rc = neumann_lists(bc.sec, bc.neumann2Length, 1))) return rc;
  // disturbances (runDisturbances.F90)
  for (int i = 0; i < bc.nDisturb; ++i) {
    const gatt_disturbance& ds = bc.disturb[i];
    BcOp op{};
    for (int q = 0; q < 6; ++q) op.box[q] = ds.box[q] - 1;
    OpSets st;
    if (ds.kind == GATT_DIST_LID_DRIVEN) {
      op.kind = BC_LID;
      op.field = GATT_U;
      op.value = ds.par[0];
      op.aux = (int)aux.size();
      const int nb = op.box[1] - op.box[0] + 1;
      if (!ds.x) return fail(GATT_E_INVALID, "lid disturbance without x coordinates");
      for (int q = 0; q < nb; ++q) aux.push_back(ds.x[q]);
      Box w{GATT_U, {}};
      std::memcpy(w.b, op.box, sizeof(w.b));
      st.wr.push_back(w);
    } else if (ds.kind == GATT_DIST_HOLD_INLET) {
      op.kind = BC_HOLD;
      op.hold_r = ds.par[0] > 0.5;
      op.aux = (int)hold_total;
      const long long npts =
          (long long)(op.box[1] - op.box[0] + 1) * (op.box[3] - op.box[2] + 1) * (op.box[5] - op.box[4] + 1);
      hold_total += 5 * npts;
      for (int f = 0; f < 5; ++f) {
        Box b{f, {}};
        std::memcpy(b.b, op.box, sizeof(b.b));
        st.rd.push_back(b);
        st.wr.push_back(b);
      }
    } else if (ds.kind == GATT_DIST_PERIODIC_SINE || ds.kind == GATT_DIST_PACKET_2D) {
      // periodicSine.F90 / packet_2d.F90 on the field passed as u.  Their pi, alpha, x0,
      // Tend, Tp, sigmaT and sigmaX are default (single precision) reals: the values
      // are rounded to float here, as the assignments in those routines round them.
      if (ds.var < 0 || ds.var > 4) return fail(GATT_E_INVALID, "disturbance variable outside U..E");
      if (!ds.x) return fail(GATT_E_INVALID, "disturbance without x coordinates");
      const int nb = op.box[1] - op.box[0] + 1;
      op.field = ds.var;
      op.aux = (int)aux.size();
      for (int q = 0; q < nb; ++q) aux.push_back(ds.x[q]);
      const float pi = (float)(4.0 * std::atan(1.0));
      const double xl = ds.x[nb - 1], xr = ds.x[0];
      if (ds.kind == GATT_DIST_PERIODIC_SINE) {
        op.kind = BC_SINE;
        op.value = ds.par[1];
        aux.push_back((double)(float)(2.0 * (double)pi / (xl - xr)));
        aux.push_back(ds.par[0]);
      } else {
        op.kind = BC_PACKET;
        const double omega = ds.par[0], nmodes = ds.par[1];
        op.value = ds.par[2];
        op.L = (int)nmodes;
        const float Tend = (float)((2.0 * (double)pi / omega) * 2.5e-1);
        aux.push_back((double)(float)((xl + xr) / 2));
        aux.push_back((double)Tend);
        aux.push_back((double)(float)((double)Tend * 5e-1));
        aux.push_back((double)(float)((double)Tend * 2.5e-1));
        aux.push_back((double)(float)((xl - xr) / 8.0));
        aux.push_back(omega);
        aux.push_back(nmodes);
      }
      Box w{ds.var, {}};
      std::memcpy(w.b, op.box, sizeof(w.b));
      st.wr.push_back(w);
      st.rd.push_back(w);
      c->timed = true;
    } else {
      return fail(GATT_E_UNSUPPORTED, "unknown disturbance kind");
    }
    if ((rc = add(op, st))) return rc;
  }
  // Dead writes: where a later op B writes field f over points an earlier single-field op A
  // (a Dirichlet value, R from P, a Neumann or null-second-derivative extrapolation) wrote,
  // and neither B nor any op between them reads f there, A's values there are overwritten
  // unread (a corner point of an inflow plane and a wall row, say).  A's box loses those
  // points where what remains is a box (the overlap covers A's box in two axes from one
  // end), or A goes entirely; the program's result is unchanged, and the write-write
  // conflicts between boundary planes that meet at a corner no longer order the waves.
  auto reads_of = [&](const BcOp& op, const OpSets& st, std::vector<Box>& out) {
    out.clear();
    if (op.kind == BC_SET) return;
    if (op.kind == BC_SET_R_FROM_P) {
      Box r{GATT_E, {}};
      std::memcpy(r.b, op.box, sizeof(r.b));
      out.push_back(r);
      return;
    }
    if (op.kind == BC_NEU) {
      Box r{op.field, {}};
      stencil_ex(op.box, op.dir, op.L, r.b);
      out.push_back(r);
      return;
    }
    if (op.kind == BC_NEU_P) {
      Box r1{GATT_R, {}}, r2{GATT_E, {}};
      stencil_ex(op.box, op.dir, op.L, r1.b);
      stencil(op.box, op.dir, op.L, r2.b);
      out.push_back(r1);
      out.push_back(r2);
      return;
    }
    out = st.rd;
  };
  auto trimmable = [](const BcOp& op) {
    return op.kind == BC_SET || op.kind == BC_SET_R_FROM_P || op.kind == BC_NEU || op.kind == BC_NEU_P;
  };
  auto intersect = [](const int a[6], const int b[6], int out[6]) {
    for (int d = 0; d < 3; ++d) {
      out[2 * d] = std::max(a[2 * d], b[2 * d]);
      out[2 * d + 1] = std::min(a[2 * d + 1], b[2 * d + 1]);
      if (out[2 * d] > out[2 * d + 1]) return false;
    }
    return true;
  };
  // wave of each op (see below); returns the number of waves
  auto conflicts = [&](const OpSets& a, const OpSets& b) {
    for (auto& w : a.wr) {
      for (auto& x : b.wr)
        if (overlap(w, x)) return true;
      for (auto& x : b.rd)
        if (overlap(w, x)) return true;
    }
    for (auto& r : a.rd)
      for (auto& x : b.wr)
        if (overlap(r, x)) return true;
    return false;
  };
  auto levels_of = [&](const std::vector<OpSets>& st, std::vector<int>& level) {
    const int n = (int)st.size();
    level.assign(n, 0);
    int nl = 0;
    for (int i = 0; i < n; ++i) {
      for (int j = 0; j < i; ++j)
        if (level[j] + 1 > level[i] && conflicts(st[i], st[j])) level[i] = level[j] + 1;
      nl = std::max(nl, level[i] + 1);
    }
    return nl;
  };
  {
    const std::vector<BcOp> ops0 = ops;
    const std::vector<OpSets> sets0 = sets;
    std::vector<char> alive(ops.size(), 1);
    std::vector<Box> rd;
    // GATT_BC_TRIM=0 keeps every op as listed, for comparisons.  The search is cubic in the
    // op count, so programs of more than 400 ops (many corner points) are kept as listed.
    const char* trim_env = std::getenv("GATT_BC_TRIM");
    bool changed = !(trim_env && trim_env[0] == '0') && ops.size() <= 400;
    while (changed) {
      changed = false;
      for (size_t bi = 0; bi < ops.size(); ++bi) {
        if (!alive[bi]) continue;
        for (size_t ai = 0; ai < bi; ++ai) {
          if (!alive[ai] || !trimmable(ops[ai])) continue;
          BcOp& A = ops[ai];
          for (const Box& w : sets[bi].wr) {
            int I[6];
            if (w.f != A.field || !intersect(A.box, w.b, I)) continue;
            bool read = false;
            for (size_t ci = ai + 1; ci <= bi && !read; ++ci) {
              if (!alive[ci]) continue;
              reads_of(ops[ci], sets[ci], rd);
              int J[6];
              for (const Box& r : rd)
                if (r.f == A.field && intersect(r.b, I, J)) read = true;
            }
            if (read) continue;
            int full = 0, part = -1;
            for (int d = 0; d < 3; ++d) {
              if (I[2 * d] == A.box[2 * d] && I[2 * d + 1] == A.box[2 * d + 1]) ++full;
              else part = d;
            }
            if (full == 3) {
              alive[ai] = 0;
            } else if (full == 2 && I[2 * part] == A.box[2 * part]) {
              A.box[2 * part] = I[2 * part + 1] + 1;
            } else if (full == 2 && I[2 * part + 1] == A.box[2 * part + 1]) {
              A.box[2 * part + 1] = I[2 * part] - 1;
            } else {
              continue;
            }
            // the trimmed op's own sets follow its box
            std::memcpy(sets[ai].wr[0].b, A.box, sizeof(A.box));
            reads_of(A, sets[ai], sets[ai].rd);
            changed = true;
            break;
          }
        }
      }
    }
    std::vector<BcOp> kept;
    std::vector<OpSets> kept_sets;
    for (size_t i = 0; i < ops.size(); ++i) {
      if (!alive[i]) continue;
      kept.push_back(ops[i]);
      kept_sets.push_back(sets[i]);
    }
    // the trimmed program is kept only where it takes fewer waves (with as many waves, the
    // program as listed measured faster on the lid-driven cavity)
    std::vector<int> lv;
    if (levels_of(kept_sets, lv) < levels_of(sets0, lv)) {
      ops.swap(kept);
      sets.swap(kept_sets);
    } else {
      ops = ops0;
      sets = sets0;
    }
    c->undo_total = 0;
    for (BcOp& op : ops) {
      op.npts = (long long)(op.box[1] - op.box[0] + 1) * (op.box[3] - op.box[2] + 1) * (op.box[5] - op.box[4] + 1);
      op.undo = c->undo_total;
      c->undo_total += 5 * op.npts;
    }
  }
  // Waves: each op goes one wave after the last earlier op it conflicts with (it reads what
  // that op writes, or writes what that op reads or writes), so the ops of one wave are
  // mutually independent, every pair of conflicting ops keeps its order, and independent
  // ops of the reference's list share a wave even when a conflicting op stands between
  // them.  The ops are then stored wave after wave (stable within a wave).
  const int nops = (int)ops.size();
  std::vector<int> level;
  const int nlev = levels_of(sets, level);
  std::vector<BcWave> waves;
  std::vector<BcOp> sorted;
  sorted.reserve(nops);
  for (int l = 0; l < nlev; ++l) {
    const int begin = (int)sorted.size();
    long long blocks = 0;
    for (int i = 0; i < nops; ++i) {
      if (level[i] != l) continue;
      BcOp op = ops[i];
      op.first_block = blocks;
      blocks += (op.npts + 255) / 256;
      sorted.push_back(op);
    }
    waves.push_back({begin, (int)sorted.size(), blocks});
  }
  ops.swap(sorted);
  c->ops = ops;
  c->waves = waves;
  if (!ops.empty() && (rc = dev_upload(c, ops.data(), ops.size(), &c->d_ops))) return rc;
  if (!aux.empty() && (rc = dev_upload(c, aux.data(), aux.size(), &c->d_aux))) return rc;
  if (bc.neumannLength > 0 && (rc = dev_upload(c, bc.neumannCoeffs, (size_t)bc.neumannLength, &c->d_coef[0])))
    return rc;
  if (bc.neumann2Length > 0 && (rc = dev_upload(c, bc.neumann2Coeffs, (size_t)bc.neumann2Length, &c->d_coef[1])))
    return rc;
  if (hold_total > 0) {
    void* p = nullptr;
    HIPCHK(hipMalloc(&p, (size_t)hold_total * c->batch * sizeof(double)));
    c->dev_allocs.push_back(p);
    c->d_hold = static_cast<double*>(p);
    c->hold_stride = hold_total;
  }
  return GATT_OK;
}

// ------------------------------------------------------------- execution
// applyBoundCond at time *clock + toff (the clock holds the t of the step in progress);
// with `undo`, every op point first saves its five values there (member stride undo_total)
void run_bc(gatt_ctx* c, double* q0, double toff, double* undo = nullptr) {
  if (g_recorder) {
    // a recorded step sequence marks where each boundary application starts, so that the
    // members' applications line up when their numbers of waves differ
    g_recorder->bc = g_recorder->nbc++;
    LaunchRecord mark;
    mark.bc = g_recorder->bc;
    g_recorder->recs.push_back(std::move(mark));
  }
  for (const BcWave& w : c->waves) {
    BcArgs a{};
    a.ops = c->d_ops;
    a.op_begin = w.op_begin;
    a.op_end = w.op_end;
    for (int f = 0; f < 5; ++f) a.q[f] = q0 + (long long)f * c->np;
    a.member_stride = c->slab;
    a.coef[0] = c->d_coef[0];
    a.coef[1] = c->d_coef[1];
    a.aux = c->d_aux;
    a.hold = c->d_hold;
    a.hold_member_stride = c->hold_stride;
    a.nx = c->nx;
    a.ny = c->ny;
    a.nz = c->nz;
    a.gamma = c->gamma;
    a.capture = c->hold_captured ? 0 : 1;
    a.clock = c->d_clock;
    a.toff = toff;
    a.undo = undo;
    a.undo_member_stride = c->undo_total;
    launch_bc_wave(a, w.nblocks, c->batch, c->stream);
  }
  if (g_recorder) g_recorder->bc = -1;
  c->hold_captured = true;
}

// Operator, line geometry and coefficient tables of the solve of `kind` along `dir`.
void fill_solve_args(const gatt_ctx* c, int dir, int kind, SolveArgs& a) {
  const DirOps& d = c->dir[dir];
  a.op = d.op[kind];
  a.types = d.types;
  a.dir = dir;
  a.nx = c->nx;
  a.ny = c->ny;
  a.nz = c->nz;
  a.n1 = d.n1;
  a.nlines = d.nlines;
  a.stride = d.stride;
  a.G = d.shape[kind].G;
  a.coef = d.coef[kind];
  a.coef_ends = d.coef_ends[kind];
  a.fcoef = d.fcoef[kind];
  a.fcoef_ends = d.fcoef_ends[kind];
  a.fcoef_t = d.fcoef_t[kind];
  a.fcoef_p = d.shape[kind].variant == 4 ? d.fcoef_p[kind][d.shape[kind].M == 4 ? 0 : (d.shape[kind].M == 8 ? 1 : 2)]
                                          : nullptr;
}

// One line solve along `dir` for nf fields (member 0 pointers, member stride = slab).
void run_solve(gatt_ctx* c, int dir, int kind, int nf, double* const* in, double* const* out, double theta,
               int mode, int nmembers, long long in_stride, long long out_stride, hipStream_t st = nullptr) {
  DirOps& d = c->dir[dir];
  if (d.identity[kind]) {
    // the identity operator: a derivative or replacing filter copies, a blend
    // f + theta (f - f) leaves the field as it is
    if (mode == 1) return;
    for (int m = 0; m < nmembers; ++m)
      for (int f = 0; f < nf; ++f)
        if (in[f] + m * in_stride != out[f] + m * out_stride)
          launch_copy(out[f] + m * out_stride, in[f] + m * in_stride, c->np, st ? st : c->stream);
    return;
  }
  SolveArgs a{};
  fill_solve_args(c, dir, kind, a);
  a.nf = nf;
  a.in_stride = in_stride;
  a.out_stride = out_stride;
  for (int f = 0; f < nf; ++f) {
    a.in[f] = in[f];
    a.out[f] = out[f];
  }
  a.theta = theta;
  a.mode = mode;
  launch_line_solve(a, d.shape[kind], nmembers, st ? st : c->stream);
}

// The derivative solves of one RHS pass along every direction (in[d], out[d]: the nf fields of
// direction d) as one launch, where the directions run on one stream with lane-group shapes of
// one point count per lane and workgroup size (the lanes per line may differ) and one band
// width: each workgroup solves what it would in its direction's own launch.
// Returns false, with nothing launched, otherwise (also while a ragged group records launches,
// the combined kernel having no ragged twin).
bool run_solve_dirs_once(gatt_ctx* c, int ndir, int nf, double* const (*in)[kMaxPtr],
                         double* const (*out)[kMaxPtr], int nmembers, long long in_stride, long long out_stride) {
  if (c->fork_dirs || !c->multi_dirs || g_recorder || ndir < 2) return false;
  const LineShape& s0 = c->dir[0].shape[GATT_DERIV];
  SolveArgs parts[3];
  for (int d = 0; d < ndir; ++d) {
    const DirOps& dd = c->dir[d];
    const LineShape& s = dd.shape[GATT_DERIV];
    if (dd.identity[GATT_DERIV] || s.variant != 0 || s.M != s0.M ||
        s.waves_per_block != s0.waves_per_block || dd.op[GATT_DERIV].nR != c->dir[0].op[GATT_DERIV].nR)
      return false;
    SolveArgs& a = parts[d];
    a = SolveArgs{};
    fill_solve_args(c, d, GATT_DERIV, a);
    a.nf = nf;
    a.in_stride = in_stride;
    a.out_stride = out_stride;
    for (int f = 0; f < nf; ++f) {
      a.in[f] = in[d][f];
      a.out[f] = out[d][f];
    }
    a.mode = 0;
  }
  return launch_line_solve_multi(parts, ndir, s0, nmembers, c->stream);
}

// The derivative solve along `dir` can take its input as viscous fluxes computed on the fly:
// a fused wave-segment shape (variants 2 and 3) and a band the flux kernels are built for.
bool flux_solve_ok(const gatt_ctx* c, int dir) {
  const DirOps& d = c->dir[dir];
  const int v = d.shape[GATT_DERIV].variant;
  return !d.identity[GATT_DERIV] && (v == 2 || v == 3) && d.op[GATT_DERIV].nR <= kFluxMaxNR;
}

// Derivatives along `dir` of the viscous fluxes q[0..nf) of nmembers members: the inputs are
// computed from E (member stride e_stride) and the first derivatives Dw (member stride ws) as
// the kernel loads them (visc_flux); results in out (member stride ws).
void run_solve_flux(gatt_ctx* c, int dir, int nf, const int* q, double* const* out, int nmembers, const double* E,
                    long long e_stride, const double* Dw, long long ws, hipStream_t st) {
  SolveArgs a{};
  fill_solve_args(c, dir, GATT_DERIV, a);
  a.nf = nf;
  a.in_stride = e_stride;
  a.out_stride = ws;
  for (int f = 0; f < nf; ++f) {
    a.in[f] = E;   // not read in flux mode
    a.out[f] = out[f];
    a.flux_q[f] = q[f];
  }
  a.has_flux = 1;
  a.flux_E = E;
  a.flux_D = Dw;
  a.flux_d_stride = ws;
  a.fc = c->fc;
  a.is3d = c->is3d;
  a.mode = 0;
  launch_line_solve(a, c->dir[dir].shape[GATT_DERIV], nmembers, st);
}

// the pressureGradient forcing of the assemble kernel
void set_pressure_gradient(const gatt_ctx* c, AssembleArgs& aa) {
  aa.pg = c->pg;
  aa.pg_r = c->pg_r;
  aa.pg_u = c->pg_u;
  for (int q = 0; q < 6; ++q) aa.pg_box[q] = c->pg_box[q];
  aa.pg_dpdx = c->pg_dpdx;
  aa.nx = c->nx;
  aa.ny = c->ny;
}

// Navier-Stokes RHS of the stage state S and one RK stage (nseq3d/nseq2d + RK4 update),
// for all members, one work group after the other.
void run_rhs_stage(gatt_ctx* c, double* S, double* Q, double* Out, int stage, double dt) {
  const long long np = c->np;
  double* D = c->work_ptr(WK_D);
  double* T = c->work_ptr(WK_T);
  double* G = c->work_ptr(WK_G);
  const long long ws = c->work_stride;
  std::vector<int> flds = c->is3d ? std::vector<int>{0, 1, 2, 3, 4} : std::vector<int>{0, 1, 3, 4};
  const int ndir = c->is3d ? 3 : 2;
  // the direction solves are independent: fork them onto the auxiliary streams
  auto fork = [&]() {
    if (!c->fork_dirs) return;
    (void)hipEventRecord(c->ev_fork, c->stream);
    for (int d = 0; d < ndir; ++d) (void)hipStreamWaitEvent(c->aux[d], c->ev_fork, 0);
  };
  auto join = [&]() {
    if (!c->fork_dirs) return;
    for (int d = 0; d < ndir; ++d) {
      (void)hipEventRecord(c->ev_join[d], c->aux[d]);
      (void)hipStreamWaitEvent(c->stream, c->ev_join[d], 0);
    }
  };
  auto dir_stream = [&](int d) { return c->fork_dirs ? c->aux[d] : c->stream; };
  fork();   // the forcing sums below run beside the first-derivative solves
  if (c->forcing && c->is3d) {
    SumArgs sa{};
    sa.np = np;
    sa.member_stride = c->slab;
    sa.S = S;
    sa.sums = c->d_sums;
    if (c->sums_from_assemble && S == c->slab_ptr(SL_S)) {
      // the previous stage's assemble kernel summed the values it wrote here
      sa.partial = c->d_apartial;
      sa.nblocks = c->apart_blocks;
      launch_forcing_final(sa, c->batch, c->stream);
    } else {
      sa.partial = c->d_partial;
      sa.nblocks = c->sum_blocks;
      launch_forcing_sums(sa, c->batch, c->stream);
    }
  }
  c->sums_from_assemble = false;
  const bool fused = c->fuse_rhs;
  const int pblocks = fused ? rhs_line_blocks(c->dir[1].nlines) : c->ablocks;
  // the stage states of RK4 (stages 1-3) and SSPRK3 (6-7) feed the next stage's sums when no
  // boundary op touches them in between
  const bool fuse_sums = c->forcing && c->is3d && pblocks > 0 && c->waves.empty() && Out == c->slab_ptr(SL_S) &&
                         (stage == 1 || stage == 2 || stage == 3 || stage == 6 || stage == 7);
  // second derivatives: x: txx,txy,txz,qx  y: txy,tyy,tyz,qy  z: txz,tyz,tzz,qz
  const int tin3[3][4] = {{0, 3, 4, 6}, {3, 1, 5, 7}, {4, 5, 2, 8}};
  for (int m0 = 0; m0 < c->batch; m0 += c->wgroup) {
    const int nm = std::min(c->wgroup, c->batch - m0);
    const long long mo = (long long)m0 * c->slab;
    if (m0 > 0) fork();
    {
      double* in[3][kMaxPtr];
      double* out[3][kMaxPtr];
      int nf = 0;
      for (int d = 0; d < ndir; ++d) {
        nf = 0;
        for (int f : flds) {
          in[d][nf] = S + mo + (long long)f * np;
          out[d][nf] = D + (long long)(3 * f + d) * np;
          ++nf;
        }
      }
      if (!run_solve_dirs_once(c, ndir, nf, in, out, nm, c->slab, ws))
        for (int d = 0; d < ndir; ++d) run_solve(c, d, GATT_DERIV, nf, in[d], out[d], 0.0, 0, nm, c->slab, ws, dir_stream(d));
    }
    join();
    AssembleArgs aa{};
    aa.c = c->fc;
    aa.np = np;
    aa.member_stride = c->slab;
    aa.work_stride = ws;
    aa.S = S + mo;
    aa.Q = Q + mo;
    aa.Out = Out + mo;
    aa.Delta = c->slab_ptr(SL_DL) + mo;
    aa.D = D;
    aa.T = T;
    aa.G = G;
    aa.sums = (c->forcing && c->is3d) ? c->d_sums + 8LL * m0 : nullptr;
    aa.forcing_eps = c->f_eps;
    aa.forcing_e0 = c->f_e0;
    aa.stage = stage;
    aa.dt = dt;
    aa.is3d = c->is3d;
    set_pressure_gradient(c, aa);
    aa.sum_partial = fuse_sums ? c->d_apartial + (long long)m0 * pblocks * 5 : nullptr;
    if (fused) {
      // x (and z) flux derivatives into G, the fluxes computed as the solves load them where
      // the shape allows (otherwise from the stresses of visc_kernel), then the y flux
      // derivatives, the assembly and the stage update in one kernel along y
      const int other[2] = {0, 2};
      const int nother = c->is3d ? 2 : 1;
      bool need_t = false;
      for (int k = 0; k < nother; ++k) need_t = need_t || !flux_solve_ok(c, other[k]);
      if (need_t) {
        ViscArgs va{};
        va.c = c->fc;
        va.np = np;
        va.member_stride = c->slab;
        va.work_stride = ws;
        va.E = S + mo + 4 * np;
        va.D = D;
        va.T = T;
        va.is3d = c->is3d;
        launch_visc(va, nm, c->stream);
      }
      fork();
      for (int k = 0; k < nother; ++k) {
        const int d = other[k];
        double* in[kMaxPtr];
        double* out[kMaxPtr];
        int q[kMaxPtr];
        int nf = 0;
        for (int j = 0; j < 4; ++j) {
          if (!c->is3d && j == 2) continue;
          q[nf] = tin3[d][j];
          in[nf] = T + (long long)tin3[d][j] * np;
          out[nf] = G + (long long)(4 * d + j) * np;
          ++nf;
        }
        if (flux_solve_ok(c, d))
          run_solve_flux(c, d, nf, q, out, nm, S + mo + 4 * np, c->slab, D, ws, dir_stream(d));
        else
          run_solve(c, d, GATT_DERIV, nf, in, out, 0.0, 0, nm, ws, ws, dir_stream(d));
      }
      join();
      RhsLineArgs r{};
      fill_solve_args(c, 1, GATT_DERIV, r.s);
      r.a = aa;
      launch_rhs_line(r, nm, c->stream);
      continue;
    }
    ViscArgs va{};
    va.c = c->fc;
    va.np = np;
    va.member_stride = c->slab;
    va.work_stride = ws;
    va.E = S + mo + 4 * np;
    va.D = D;
    va.T = T;
    va.is3d = c->is3d;
    launch_visc(va, nm, c->stream);
    fork();
    {
      double* in[3][kMaxPtr];
      double* out[3][kMaxPtr];
      int nf = 0;
      for (int d = 0; d < ndir; ++d) {
        nf = 0;
        for (int q = 0; q < 4; ++q) {
          if (!c->is3d && q == 2) continue;
          in[d][nf] = T + (long long)tin3[d][q] * np;
          out[d][nf] = G + (long long)(4 * d + q) * np;
          ++nf;
        }
      }
      if (!run_solve_dirs_once(c, ndir, nf, in, out, nm, ws, ws))
        for (int d = 0; d < ndir; ++d) run_solve(c, d, GATT_DERIV, nf, in[d], out[d], 0.0, 0, nm, ws, ws, dir_stream(d));
    }
    join();
    launch_assemble(aa, nm, c->stream);
  }
  c->sums_from_assemble = fuse_sums;
  c->apart_blocks = pblocks;
}

// RK4_3D / RK4_2D: state `from` -> `to` (timeIntegration.F90:382-496); the stage
// boundary conditions act at t + dt/2, t + dt/2, t + dt, t + dt with t on the clock.
void run_rk4(gatt_ctx* c, double* from, double* to, double dt) {
  double* S = c->slab_ptr(SL_S);
  run_rhs_stage(c, from, from, S, 1, dt);
  run_bc(c, S, dt / 2.0);
  run_rhs_stage(c, S, from, S, 2, dt);
  run_bc(c, S, dt / 2.0);
  run_rhs_stage(c, S, from, S, 3, dt);
  run_bc(c, S, dt);
  run_rhs_stage(c, S, from, to, 4, dt);
  run_bc(c, to, dt);
}

// EULER_3D / EULER_2D (timeIntegration.F90:500-566, 158-222): EULER_3D applies the
// boundary conditions at t, EULER_2D at t + dt.
void run_euler(gatt_ctx* c, double* from, double* to, double dt) {
  run_rhs_stage(c, from, from, to, 5, dt);
  run_bc(c, to, c->is3d ? 0.0 : dt);
}

// SSPRK3_3D / SSPRK3_2D (timeIntegration.F90:569-685, 225-335); the Delta slab holds
// k1 and then k1 + k2.
void run_ssprk3(gatt_ctx* c, double* from, double* to, double dt) {
  double* S = c->slab_ptr(SL_S);
  run_rhs_stage(c, from, from, S, 6, dt);
  run_bc(c, S, dt);
  run_rhs_stage(c, S, from, S, 7, dt);
  run_bc(c, S, dt / 2.0);
  run_rhs_stage(c, S, from, to, 8, dt);
  run_bc(c, to, dt);
}

// The integrator timeStepping selects (main.F90:255-350).
void run_time_step(gatt_ctx* c, double* from, double* to, double dt) {
  switch (c->stepping) {
    case 2: run_euler(c, from, to, dt); break;
    case 3: run_ssprk3(c, from, to, dt); break;
    default: run_rk4(c, from, to, dt); break;
  }
}

void run_filters(gatt_ctx* c, double* q, double theta) {
  std::vector<int> flds = c->is3d ? std::vector<int>{0, 1, 2, 3, 4} : std::vector<int>{0, 1, 3, 4};
  const int ndir = c->is3d ? 3 : 2;
  for (int d = 0; d < ndir; ++d) {
    if (!c->filterDir[d]) continue;
    double* ptr[kMaxPtr];
    int nf = 0;
    for (int f : flds) ptr[nf++] = q + (long long)f * c->np;
    run_solve(c, d, GATT_FILTER, nf, ptr, ptr, theta, 1, c->batch, c->slab, c->slab);
  }
}

// The filters from src into dst (member stride slab for both): the first direction that
// filters reads src, the others act in place on dst; src is copied when none filters.
void run_filters_from(gatt_ctx* c, const double* src, double* dst, double theta) {
  std::vector<int> flds = c->is3d ? std::vector<int>{0, 1, 2, 3, 4} : std::vector<int>{0, 1, 3, 4};
  const int ndir = c->is3d ? 3 : 2;
  bool first = true;
  for (int d = 0; d < ndir; ++d) {
    if (!c->filterDir[d] || c->dir[d].identity[GATT_FILTER]) continue;
    double* in[kMaxPtr];
    double* out[kMaxPtr];
    int nf = 0;
    for (int f : flds) {
      in[nf] = const_cast<double*>(first ? src : dst) + (long long)f * c->np;
      out[nf] = dst + (long long)f * c->np;
      ++nf;
    }
    run_solve(c, d, GATT_FILTER, nf, in, out, theta, 1, c->batch, c->slab, c->slab);
    first = false;
  }
  if (first)
    for (int b = 0; b < c->batch; ++b) launch_copy(dst + b * c->slab, src + b * c->slab, 5 * c->np, c->stream);
}

// mean_ro: the mean is read, not updated (the adjoint's recompute of a base step about a given mean)
void run_sfd(gatt_ctx* c, double* q, double dt, bool mean_ro = false) {
  if (c->sfdType <= 0) return;
  SfdArgs s{};
  s.np = c->np;
  s.member_stride = c->slab;
  s.Q = q;
  s.Mean = c->slab_ptr(SL_MEAN);
  s.chi = c->sfdType == 2 ? c->d_chi : nullptr;
  s.chi_val = c->sfdXval;
  s.dt = dt;
  s.delta = c->sfdDelta;
  s.nf_mask = c->is3d ? 0x1f : 0x1b;
  s.mean_ro = mean_ro ? 1 : 0;
  launch_sfd(s, c->batch, c->stream);
}

void ensure_mean(gatt_ctx* c) {
  if (c->sfdType > 0 && !c->mean_ready) {
    for (int b = 0; b < c->batch; ++b)
      launch_copy(c->slab_ptr(SL_MEAN) + b * c->slab, c->state(c->cur) + b * c->slab, 5 * c->np, c->stream);
    c->mean_ready = true;
  }
}

// main.F90:255-424 for all members: from state `cur` into the other buffer, at the
// time on the clock, which then advances by dt.
void enqueue_step(gatt_ctx* c, int cur, double dt) {
  double* from = c->state(cur);
  double* to = c->state(1 - cur);
  run_time_step(c, from, to, dt);
  run_filters(c, to, dt / c->filterCharTime);
  run_sfd(c, to, dt);
  run_bc(c, to, dt);
  ReduceArgs r{};
  r.np = c->np;
  r.member_stride = c->slab;
  r.Qn = to;
  r.Q = from;
  r.out = c->d_red;
  r.is3d = c->is3d;
  launch_step_reduce(r, c->batch, c->stream);
  // the record also clears the maxima for the next step and advances the clock (main.F90:424)
  launch_record(c->d_red, c->d_rec, c->d_flag, c->batch, c->stream, c->d_clock, dt);
}

int check_ctx(gatt_ctx* c) {
  if (!c) return fail(GATT_E_INVALID, "null context");
  return GATT_OK;
}

}  // namespace

extern "C" {

const char* gatt_last_error(void) { return g_err.c_str(); }

int gatt_device_info(char* buf, int32_t len) {
  hipDeviceProp_t prop;
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  HIPCHK(hipGetDeviceProperties(&prop, dev));
  if (buf && len > 0) snprintf(buf, (size_t)len, "%s (%s), %d CUs", prop.name, prop.gcnArchName, prop.multiProcessorCount);
  return GATT_OK;
}

int gatt_create(const gatt_problem_desc* d, int32_t batch, gatt_ctx** out) {
  if (!d || !out) return fail(GATT_E_INVALID, "null argument");
  *out = nullptr;
  if (batch < 1) return fail(GATT_E_INVALID, "batch must be >= 1");
  if (d->nx < 2 || d->ny < 2 || d->nz < 1) return fail(GATT_E_INVALID, "grid must be at least 2x2x1");
  std::unique_ptr<gatt_ctx> c(new gatt_ctx);
  HIPCHK(hipGetDevice(&c->device));
  {
    const int rc = init_static_scratch(c->device);
    if (rc) return rc;
  }
  c->nx = d->nx;
  c->ny = d->ny;
  c->nz = d->nz;
  c->is3d = d->nz > 1;
  c->batch = batch;
  c->np = (long long)d->nx * d->ny * d->nz;
  c->slab = (long long)(d->sfdType > 0 ? SL_N_SFD : SL_N_NOSFD) * c->np;
  c->wgroup = (int)std::max(1LL, std::min<long long>(batch, kWorkPoints / c->np));
  c->work_stride = (long long)WK_N * c->np;
  c->gamma = d->gamma;
  c->fc.gm1 = d->gamma - 1.0;
  c->fc.invRe = 1.0 / d->Re;
  c->fc.muA = 1.0 + 110.0 / d->T0;
  c->fc.muB = 110.0 / d->T0;
  c->fc.muC = d->gamma * (d->gamma - 1.0) * d->Ma * d->Ma;
  c->fc.qk = d->gamma / d->Re / d->Pr;
  for (int k = 0; k < 3; ++k) c->filterDir[k] = d->filterDir[k];
  c->filterCharTime = d->filterCharTime;
  c->stepping = d->timeStepping == 2 || d->timeStepping == 3 ? d->timeStepping : 1;
  HIPCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  for (int k = 0; k < 3; ++k) {
    HIPCHK(hipStreamCreateWithFlags(&c->aux[k], hipStreamNonBlocking));
    HIPCHK(hipEventCreateWithFlags(&c->ev_join[k], hipEventDisableTiming));
  }
  HIPCHK(hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming));
  // The direction solves of a pass run as parallel branches where the pass is large enough
  // for them to overlap (at least 2^20 points over the members); on smaller grids the fork and
  // join cost more than the overlap saves.  GATT_SERIAL_DIRECTIONS=1 (or 0) forces one stream
  // (or the branches) for comparisons.
  c->fork_dirs = c->np * batch >= (1LL << 20);
  if (const char* e = std::getenv("GATT_SERIAL_DIRECTIONS")) c->fork_dirs = e[0] != '1';
  // on one stream, the lane-group solves of all directions go out as one launch where they
  // share a shape; GATT_MULTI_DIRECTIONS=0 launches each direction on its own
  c->multi_dirs = true;
  if (const char* e = std::getenv("GATT_MULTI_DIRECTIONS")) c->multi_dirs = e[0] != '0';
  {
    int dev = 0, v = 0;
    HIPCHK(hipGetDevice(&dev));
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeSharedMemPerBlockOptin, dev) != hipSuccess || v <= 0)
      HIPCHK(hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    c->max_lds = v > 0 ? v : 65536;
  }
  const int dims[3] = {d->nx, d->ny, d->nz};
  const int ndir = c->is3d ? 3 : 2;
  int rc;
  for (int k = 0; k < ndir; ++k) {
    DirOps& o = c->dir[k];
    if ((rc = upload_lineop(c.get(), d->deriv[k], dims[k], &o.op[0], "deriv"))) return rc;
    if ((rc = upload_lineop(c.get(), d->filter[k], dims[k], &o.op[1], "filter"))) return rc;
    if ((rc = pack_coeffs(c.get(), d->deriv[k], &o.coef[0], &o.coef_ends[0]))) return rc;
    if ((rc = pack_coeffs(c.get(), d->filter[k], &o.coef[1], &o.coef_ends[1]))) return rc;
    if ((rc = pack_coeffs_fused(c.get(), d->deriv[k], &o.fcoef[0], &o.fcoef_ends[0], &o.fcoef_t[0], o.fcoef_p[0],
                                o.fcoef_pt[0])))
      return rc;
    if ((rc = pack_coeffs_fused(c.get(), d->filter[k], &o.fcoef[1], &o.fcoef_ends[1], &o.fcoef_t[1], o.fcoef_p[1],
                                o.fcoef_pt[1])))
      return rc;
    if (k == 0) { o.n1 = d->ny; o.nlines = d->ny * d->nz; o.stride = 1; }
    if (k == 1) { o.n1 = d->nx; o.nlines = d->nx * d->nz; o.stride = d->nx; }
    if (k == 2) { o.n1 = d->nx; o.nlines = d->nx * d->ny; o.stride = (long long)d->nx * d->ny; }
    if (!d->types[k]) return fail(GATT_E_INVALID, "missing type map");
    std::vector<int> tm(d->types[k], d->types[k] + o.nlines);
    for (auto& v : tm) {
      v -= 1;
      if (v < 0 || v >= o.op[0].ntypes || v >= o.op[1].ntypes) return fail(GATT_E_INVALID, "type id outside 1..ntypes");
    }
    if ((rc = dev_upload(c.get(), tm.data(), tm.size(), &o.types))) return rc;
    o.uniform_tiles = true;
    for (size_t l = 0; l < tm.size(); ++l)
      if (tm[l] != tm[l - l % 64]) o.uniform_tiles = false;
    for (int kind = 0; kind < 2; ++kind) {
      // a one-point stencil: makeMatrices.m:14-20 replaces the filter of a direction too
      // short for its stencil by LHS = RHS = 1, the identity, which needs no solve
      const gatt_lineop& h = kind == 0 ? d->deriv[k] : d->filter[k];
      if (h.nR == 1) {
        // R = 1, A = C = 0, B = 1, and for the periodic layout D = e_1 (first row of the
        // inverse): derivs.F90:556-948 then returns F exactly
        const long long nb = (long long)(h.n - (h.periodic ? 1 : 0)) * h.ntypes;
        bool ident = true;
        for (long long e = 0; ident && e < (long long)h.n * h.ntypes; ++e) ident = h.R[e] == 1.0;
        for (long long e = 0; ident && e < nb; ++e) ident = h.B[e] == 1.0;
        for (long long e = 0; ident && e < (long long)(h.n - 1) * h.ntypes; ++e) ident = h.A[e] == 0.0 && h.C[e] == 0.0;
        for (long long e = 0; ident && h.periodic && e < (long long)h.n * h.ntypes; ++e)
          ident = h.D[e] == (e % h.n == 0 ? 1.0 : 0.0);
        if (!ident) return fail(GATT_E_UNSUPPORTED, "one-point operators other than the identity");
        o.identity[kind] = true;
        continue;
      }
      o.shape[kind] = choose_shape(k, dims[k], o.op[kind].nR, o.nlines, c->max_lds, o.uniform_tiles);
      if (!line_shape_ok(o.shape[kind], dims[k], o.op[kind].nR, k, c->max_lds, o.uniform_tiles))
        return fail(GATT_E_UNSUPPORTED, "no line-solve shape fits this line length");
    }
  }
  // The y flux derivatives fused with the assembly (rhs_line.hip) where the y operator allows
  // it (one line type per 64-line tile, lines of at most 256 points), only when
  // GATT_FUSED_RHS=1: it moves fewer arrays but runs slower than the separate visc /
  // second-derivative / assemble passes (two wavefronts per SIMD and a serial point loop
  // in the fused kernel, the viscosity evaluated per flux), so those stay the default.
  {
    const DirOps& o = c->dir[1];
    const char* e = std::getenv("GATT_FUSED_RHS");
    c->fuse_rhs = e && e[0] == '1' && !o.identity[0] && o.uniform_tiles && rhs_line_fits(d->ny, o.op[0].nR);
  }
  if ((rc = compile_bc(c.get(), d->bc))) return rc;
  for (int i = 0; i < d->nForcing; ++i) {
    if (d->forcing[i].kind == GATT_FORCING_PRESSURE_GRADIENT) {
      // pressureGradient.F90 (runForcings{2,3}D.F90)
      const gatt_forcing& f = d->forcing[i];
      if (c->pg) return fail(GATT_E_UNSUPPORTED, "at most one pressureGradient forcing");
      if (f.var[0] < 0 || f.var[0] > 4 || f.var[1] < 0 || f.var[1] > 4)
        return fail(GATT_E_INVALID, "pressureGradient variables outside U..E");
      const int lim[3] = {d->nx, d->ny, d->nz};
      for (int q = 0; q < 3; ++q)
        if (f.box[2 * q] < 1 || f.box[2 * q + 1] > lim[q] || f.box[2 * q] > f.box[2 * q + 1])
          return fail(GATT_E_INVALID, "pressureGradient box outside the grid");
      c->pg = 1;
      c->pg_r = f.var[0];
      c->pg_u = f.var[1];
      for (int q = 0; q < 6; ++q) c->pg_box[q] = f.box[q] - 1;
      c->pg_dpdx = f.par[0];
      continue;
    }
    if (d->forcing[i].kind != GATT_FORCING_TURBULENCE) return fail(GATT_E_UNSUPPORTED, "unknown forcing kind");
    if (d->nz == 1) return fail(GATT_E_UNSUPPORTED, "the turbulence forcing is three-dimensional");
    const int* b = d->forcing[i].box;
    if (b[0] != 1 || b[1] != d->nx || b[2] != 1 || b[3] != d->ny || b[4] != 1 || b[5] != d->nz)
      return fail(GATT_E_UNSUPPORTED, "turbulence forcing must cover the whole grid");
    c->forcing = 1;
    c->f_eps = d->forcing[i].par[0];
    c->f_e0 = d->forcing[i].par[1];
  }
  if (c->forcing) {
    c->sum_blocks = 1024;
    void* p = nullptr;
    HIPCHK(hipMalloc(&p, sizeof(double) * 5 * c->sum_blocks * batch));
    c->dev_allocs.push_back(p);
    c->d_partial = static_cast<double*>(p);
    // partial sums written by the producer of a stage state: assemble_kernel (256-point
    // blocks) or the fused y-line kernel (one per 32 lines)
    c->ablocks = c->np % 256 == 0 ? (int)(c->np / 256) : 0;
    const int cap = std::max(c->ablocks, rhs_line_blocks(c->dir[1].nlines));
    HIPCHK(hipMalloc(&p, sizeof(double) * 5 * (size_t)cap * batch));
    c->dev_allocs.push_back(p);
    c->d_apartial = static_cast<double*>(p);
    HIPCHK(hipMalloc(&p, sizeof(double) * 8 * batch));
    c->dev_allocs.push_back(p);
    c->d_sums = static_cast<double*>(p);
    HIPCHK(hipMalloc(&p, sizeof(double) * 8 * batch));
    c->dev_allocs.push_back(p);
    c->d_lsums = static_cast<double*>(p);
  }
  c->sfdType = d->sfdType;
  c->sfdDelta = d->sfdDelta;
  c->sfdXval = d->sfdXval;
  if (c->sfdType == 2) {
    if (!d->sfdX) return fail(GATT_E_INVALID, "SFD type 2 needs the chi field");
    if ((rc = dev_upload(c.get(), d->sfdX, (size_t)c->np, &c->d_chi))) return rc;
  }
  void* p = nullptr;
  HIPCHK(hipMalloc(&p, sizeof(unsigned long long) * 8 * batch));
  c->dev_allocs.push_back(p);
  c->d_red = static_cast<unsigned long long*>(p);
  // zero before the first step; every step record clears it for the next
  HIPCHK(hipMemsetAsync(c->d_red, 0, sizeof(unsigned long long) * 8 * batch, c->stream));
  HIPCHK(hipMalloc(&p, sizeof(double) * 8));
  c->dev_allocs.push_back(p);
  c->d_rec = static_cast<double*>(p);
  HIPCHK(hipMalloc(&p, sizeof(int)));
  c->dev_allocs.push_back(p);
  c->d_flag = static_cast<int*>(p);
  HIPCHK(hipMalloc(&p, sizeof(double)));
  c->dev_allocs.push_back(p);
  c->d_clock = static_cast<double*>(p);
  launch_clock_set(c->d_clock, 0.0, c->stream);
  HIPCHK(hipMemsetAsync(c->d_flag, 0, sizeof(int), c->stream));
  hipError_t e = hipMalloc(&p, sizeof(double) * (size_t)c->slab * batch);
  if (e != hipSuccess) return fail(GATT_E_NOMEM, "state allocation failed");
  c->mem = static_cast<double*>(p);
  HIPCHK(hipMemsetAsync(c->mem, 0, sizeof(double) * (size_t)c->slab * batch, c->stream));
  e = hipMalloc(&p, sizeof(double) * (size_t)c->work_stride * c->wgroup);
  if (e != hipSuccess) return fail(GATT_E_NOMEM, "workspace allocation failed");
  c->dev_allocs.push_back(p);
  c->work = static_cast<double*>(p);
  HIPCHK(hipMemsetAsync(c->work, 0, sizeof(double) * (size_t)c->work_stride * c->wgroup, c->stream));
  // the clears run on the context's (non-blocking) stream: finish them before any other stream's work
  HIPCHK(hipStreamSynchronize(c->stream));
  *out = c.release();
  return GATT_OK;
}

int gatt_destroy(gatt_ctx* c) {
  if (!c) return GATT_OK;
  if (c->stream) hipStreamSynchronize(c->stream);
  c->drop_graphs();
  for (void* p : c->dev_allocs) hipFree(p);
  if (c->d_idx) hipFree(c->d_idx);
  if (c->d_rmap) hipFree(c->d_rmap);
  if (c->d_base) hipFree(c->d_base);
  if (c->d_scratch) hipFree(c->d_scratch);
  if (c->d_ckpt) hipFree(c->d_ckpt);
  if (c->d_points) hipFree(c->d_points);
  if (c->lin) hipFree(c->lin);
  if (c->tape) hipFree(c->tape);
  if (c->mem) hipFree(c->mem);
  for (int k = 0; k < 3; ++k) {
    if (c->aux[k]) (void)hipStreamSynchronize(c->aux[k]);
    if (c->aux[k]) (void)hipStreamDestroy(c->aux[k]);
    if (c->ev_join[k]) (void)hipEventDestroy(c->ev_join[k]);
  }
  if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
  if (c->stream) hipStreamDestroy(c->stream);
  delete c;
  return GATT_OK;
}

static constexpr int32_t kMaxPoints = 1 << 20;

static int copy_fields(gatt_ctx* c, int32_t member, double* dev, const double* const* hq, bool to_dev) {
  if (member < 0 || member >= c->batch) return fail(GATT_E_INVALID, "member out of range");
  double* base = dev + (long long)member * c->slab;
  for (int f = 0; f < 5; ++f) {
    if (!c->is3d && f == GATT_W) continue;
    if (!hq[f]) return fail(GATT_E_INVALID, "null field pointer");
    double* dptr = base + (long long)f * c->np;
    if (to_dev) HIPCHK(hipMemcpyAsync(dptr, hq[f], sizeof(double) * c->np, hipMemcpyHostToDevice, c->stream));
    else HIPCHK(hipMemcpyAsync((void*)hq[f], dptr, sizeof(double) * c->np, hipMemcpyDeviceToHost, c->stream));
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  return GATT_OK;
}

int gatt_set_state(gatt_ctx* c, int32_t member, const double* const* q) {
  if (check_ctx(c) || !q) return fail(GATT_E_INVALID, "null argument");
  return copy_fields(c, member, c->state(c->cur), q, true);
}

int gatt_get_state(gatt_ctx* c, int32_t member, double* const* q) {
  if (check_ctx(c) || !q) return fail(GATT_E_INVALID, "null argument");
  int rc = copy_fields(c, member, c->state(c->cur), (const double* const*)q, false);
  if (rc) return rc;
  if (!c->is3d && q[GATT_W]) std::memset(q[GATT_W], 0, sizeof(double) * c->np);
  return GATT_OK;
}

int gatt_get_previous_state(gatt_ctx* c, int32_t member, double* const* q) {
  if (check_ctx(c) || !q) return fail(GATT_E_INVALID, "null argument");
  int rc = copy_fields(c, member, c->state(1 - c->cur), (const double* const*)q, false);
  if (rc) return rc;
  if (!c->is3d && q[GATT_W]) std::memset(q[GATT_W], 0, sizeof(double) * c->np);
  return GATT_OK;
}

int gatt_checkpoint(gatt_ctx* c, int32_t op) {
  if (check_ctx(c)) return GATT_E_INVALID;
  if (op != 0 && op != 1) return fail(GATT_E_INVALID, "checkpoint op must be 0 (save) or 1 (restore)");
  // the state and, with SFD, its mean: ten arrays per member
  const long long per = 10 * c->np;
  if (!c->d_ckpt) {
    if (op == 1) return fail(GATT_E_INVALID, "no checkpoint saved");
    void* p = nullptr;
    if (hipMalloc(&p, sizeof(double) * (size_t)per * c->batch) != hipSuccess)
      return fail(GATT_E_NOMEM, "checkpoint allocation failed");
    c->d_ckpt = static_cast<double*>(p);
  }
  const bool mean = c->sfdType > 0;
  for (int m = 0; m < c->batch; ++m) {
    double* st = c->state(c->cur) + (long long)m * c->slab;
    double* mn = c->slab_ptr(SL_MEAN) + (long long)m * c->slab;
    double* ck = c->d_ckpt + (long long)m * per;
    if (op == 0) {
      launch_copy(ck, st, 5 * c->np, c->stream);
      if (mean) launch_copy(ck + 5 * c->np, mn, 5 * c->np, c->stream);
    } else {
      launch_copy(st, ck, 5 * c->np, c->stream);
      if (mean) launch_copy(mn, ck + 5 * c->np, 5 * c->np, c->stream);
    }
  }
  if (op == 0) c->ckpt_mean_ready = c->mean_ready;
  else c->mean_ready = c->ckpt_mean_ready;
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  return GATT_OK;
}

int gatt_get_points(gatt_ctx* c, int32_t member, int32_t npts, const int64_t* idx, double* out) {
  if (check_ctx(c)) return GATT_E_INVALID;
  if (member < 0 || member >= c->batch) return fail(GATT_E_INVALID, "member out of range");
  if (npts < 0 || npts > kMaxPoints) return fail(GATT_E_INVALID, "point count out of range");
  if (npts == 0) return GATT_OK;
  if (!idx || !out) return fail(GATT_E_INVALID, "null argument");
  for (int32_t i = 0; i < npts; ++i)
    if (idx[i] < 0 || idx[i] >= c->np) return fail(GATT_E_INVALID, "point index outside the grid");
  if (npts > c->points_cap) {
    HIPCHK(hipStreamSynchronize(c->stream));
    if (c->d_points) HIPCHK(hipFree(c->d_points));
    c->d_points = nullptr;
    c->points_cap = 0;
    HIPCHK(hipMalloc(&c->d_points, (size_t)npts * (sizeof(long long) + 5 * sizeof(double))));
    c->points_cap = npts;
  }
  long long* d_idx = static_cast<long long*>(c->d_points);
  double* d_out = reinterpret_cast<double*>(d_idx + c->points_cap);
  HIPCHK(hipMemcpyAsync(d_idx, idx, sizeof(long long) * npts, hipMemcpyHostToDevice, c->stream));
  launch_gather_points(c->state(c->cur) + (long long)member * c->slab, c->np, c->is3d, npts, d_idx, d_out, c->stream);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(out, d_out, sizeof(double) * 5 * npts, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return GATT_OK;
}

int gatt_set_mean(gatt_ctx* c, int32_t member, const double* const* q) {
  if (check_ctx(c) || !q) return fail(GATT_E_INVALID, "null argument");
  if (c->sfdType <= 0) return fail(GATT_E_INVALID, "SFD is off: no mean to set");
  // the other members keep the state as their mean, as at the first step
  ensure_mean(c);
  return copy_fields(c, member, c->slab_ptr(SL_MEAN), q, true);
}

int gatt_get_mean(gatt_ctx* c, int32_t member, double* const* q) {
  if (check_ctx(c) || !q) return fail(GATT_E_INVALID, "null argument");
  if (c->sfdType <= 0) return fail(GATT_E_INVALID, "SFD is off: no mean to read");
  // before the mean is set or first used it is the state
  return copy_fields(c, member, c->mean_ready ? c->slab_ptr(SL_MEAN) : c->state(c->cur), (const double* const*)q,
                     false);
}

int gatt_rk4(gatt_ctx* c, double t, double dt) {
  if (check_ctx(c)) return GATT_E_INVALID;
  launch_clock_set(c->d_clock, t, c->stream);
  run_rk4(c, c->state(c->cur), c->state(1 - c->cur), dt);
  c->cur = 1 - c->cur;
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  return GATT_OK;
}

int gatt_euler(gatt_ctx* c, double t, double dt) {
  if (check_ctx(c)) return GATT_E_INVALID;
  launch_clock_set(c->d_clock, t, c->stream);
  run_euler(c, c->state(c->cur), c->state(1 - c->cur), dt);
  c->cur = 1 - c->cur;
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  return GATT_OK;
}

int gatt_ssprk3(gatt_ctx* c, double t, double dt) {
  if (check_ctx(c)) return GATT_E_INVALID;
  launch_clock_set(c->d_clock, t, c->stream);
  run_ssprk3(c, c->state(c->cur), c->state(1 - c->cur), dt);
  c->cur = 1 - c->cur;
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  return GATT_OK;
}

int gatt_apply_filters(gatt_ctx* c, double theta) {
  if (check_ctx(c)) return GATT_E_INVALID;
  run_filters(c, c->state(c->cur), theta);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  return GATT_OK;
}

int gatt_apply_bound_cond(gatt_ctx* c, double t) {
  if (check_ctx(c)) return GATT_E_INVALID;
  launch_clock_set(c->d_clock, t, c->stream);
  run_bc(c, c->state(c->cur), 0.0);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  return GATT_OK;
}

int gatt_sfd(gatt_ctx* c, double dt) {
  if (check_ctx(c)) return GATT_E_INVALID;
  ensure_mean(c);
  run_sfd(c, c->state(c->cur), dt);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  return GATT_OK;
}

int gatt_rhs(gatt_ctx* c, double* const* dq) {
  if (check_ctx(c) || !dq) return fail(GATT_E_INVALID, "null argument");
  double* S = c->state(c->cur);
  double* O = c->slab_ptr(SL_S);
  run_rhs_stage(c, S, S, O, 0, 0.0);
  HIPCHK(hipGetLastError());
  for (int f = 0; f < 5; ++f) {
    if (!dq[f]) continue;
    if (!c->is3d && f == GATT_W) {
      std::memset(dq[f], 0, sizeof(double) * c->np);
      continue;
    }
    HIPCHK(hipMemcpyAsync(dq[f], O + (long long)f * c->np, sizeof(double) * c->np, hipMemcpyDeviceToHost, c->stream));
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  return GATT_OK;
}

static int read_scalars(gatt_ctx* c, double* umax, double* maxchange, int32_t* diverged) {
  double v[8];
  int flag = 0;
  HIPCHK(hipMemcpyAsync(v, c->d_rec, sizeof(v), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipMemcpyAsync(&flag, c->d_flag, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (maxchange)
    for (int i = 0; i < 5; ++i) maxchange[i] = v[i];
  if (umax)
    for (int i = 0; i < 3; ++i) umax[i] = v[5 + i];
  if (diverged) *diverged = flag;
  return GATT_OK;
}

int gatt_step(gatt_ctx* c, double t, double dt, double* umax, double* maxchange, int32_t* diverged) {
  if (check_ctx(c)) return GATT_E_INVALID;
  ensure_mean(c);
  launch_clock_set(c->d_clock, t, c->stream);
  // the divergence flag describes this call's steps only
  launch_clear_i32(c->d_flag, 1, c->stream);
  // holdInlet captures the inflow at the first boundary application (main.F90:112)
  if (c->d_hold && !c->hold_captured) run_bc(c, c->state(c->cur), 0.0);
  enqueue_step(c, c->cur, dt);
  c->cur = 1 - c->cur;
  HIPCHK(hipGetLastError());
  return read_scalars(c, umax, maxchange, diverged);
}

int gatt_advance(gatt_ctx* c, int32_t nsteps, double t0, double dt, double* maxchange, int32_t* diverged) {
  if (check_ctx(c)) return GATT_E_INVALID;
  if (nsteps < 0) return fail(GATT_E_INVALID, "nsteps < 0");
  ensure_mean(c);
  // holdInlet captures the inflow at the first boundary application; a problem that holds its
  // inflow and has not had one gets it here (main.F90:112); otherwise the state is left as given
  launch_clock_set(c->d_clock, t0, c->stream);
  // cleared outside the captured graph: the flag then accumulates over this call's steps
  launch_clear_i32(c->d_flag, 1, c->stream);
  if (c->d_hold && !c->hold_captured) run_bc(c, c->state(c->cur), 0.0);
  int s = 0;
  if (c->graph && (c->graph_dt != dt || c->graph_cur != c->cur)) {
    hipGraphExecDestroy(c->graph);
    c->graph = nullptr;
  }
  if (nsteps >= 2 && !c->graph) {
    hipGraph_t g;
    HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    enqueue_step(c, c->cur, dt);
    enqueue_step(c, 1 - c->cur, dt);
    HIPCHK(hipStreamEndCapture(c->stream, &g));
    HIPCHK(hipGraphInstantiate(&c->graph, g, nullptr, nullptr, 0));
    HIPCHK(hipGraphDestroy(g));
    c->graph_dt = dt;
    c->graph_cur = c->cur;
  }
  for (; s + 2 <= nsteps; s += 2) HIPCHK(hipGraphLaunch(c->graph, c->stream));
  if (s < nsteps) {
    enqueue_step(c, c->cur, dt);
    c->cur = 1 - c->cur;
    if (c->graph) {
      hipGraphExecDestroy(c->graph);
      c->graph = nullptr;
    }
  }
  HIPCHK(hipGetLastError());
  return read_scalars(c, nullptr, maxchange, diverged);
}

int gatt_line_solve(gatt_ctx* c, int32_t dir, int32_t kind, int32_t nfields, const double* const* in,
                    double* const* out) {
  if (check_ctx(c)) return GATT_E_INVALID;
  if (dir < 0 || dir > 2 || (!c->is3d && dir == 2)) return fail(GATT_E_INVALID, "direction out of range");
  if (kind != GATT_DERIV && kind != GATT_FILTER) return fail(GATT_E_INVALID, "unknown operator kind");
  if (nfields < 1 || nfields > kMaxPtr || !in || !out) return fail(GATT_E_INVALID, "bad field list");
  double* ip[kMaxPtr];
  double* op[kMaxPtr];
  for (int f = 0; f < nfields; ++f) {
    if (!in[f] || !out[f] || in[f] == out[f]) return fail(GATT_E_INVALID, "null or aliasing field pointer");
    ip[f] = const_cast<double*>(in[f]);
    op[f] = out[f];
  }
  run_solve(c, dir, kind, nfields, ip, op, 0.0, 0, 1, 0, 0);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  return GATT_OK;
}

int gatt_sync(gatt_ctx* c) {
  if (check_ctx(c)) return GATT_E_INVALID;
  HIPCHK(hipStreamSynchronize(c->stream));
  return GATT_OK;
}

int gatt_set_line_shape(gatt_ctx* c, int32_t dir, int32_t kind, int32_t variant, int32_t param, int32_t wpb) {
  if (check_ctx(c)) return GATT_E_INVALID;
  if (dir < 0 || dir > 2 || (!c->is3d && dir == 2) || (kind != GATT_DERIV && kind != GATT_FILTER))
    return fail(GATT_E_INVALID, "direction or kind out of range");
  DirOps& d = c->dir[dir];
  if (d.identity[kind]) return fail(GATT_E_UNSUPPORTED, "the identity operator has no line solve");
  const int N = d.op[kind].N;
  LineShape s;
  s.variant = variant;
  if (variant == 0) {
    if (param < 2 || param > 64 || (param & (param - 1)) || wpb < 1 || wpb > 4)
      return fail(GATT_E_INVALID, "bad lane-group shape");
    s.G = param;
    s.M = std::max(4, pow2_at_least((N + param - 1) / param));
    s.waves_per_block = wpb;
  } else if (variant == 4) {
    s.M = param;
    s.G = 64;
    s.waves_per_block = 4;
  } else if (variant >= 1 && variant <= 3) {
    s.M = param;
    s.G = param > 0 ? (N + param - 1) / param : 0;
    // variant 3: waves_per_block selects the tile (1: 32 lines, 2: 16 lines)
    s.waves_per_block = variant == 3 ? (wpb == 2 ? 2 : 1) : s.G;
  } else {
    return fail(GATT_E_INVALID, "unknown line-solve variant");
  }
  if (!line_shape_ok(s, N, d.op[kind].nR, dir, c->max_lds, d.uniform_tiles))
    return fail(GATT_E_UNSUPPORTED, "line shape does not fit this line");
  d.shape[kind] = s;
  c->drop_graphs();
  return GATT_OK;
}

int gatt_get_line_shape(gatt_ctx* c, int32_t dir, int32_t kind, int32_t* variant, int32_t* G, int32_t* M,
                        int32_t* wpb) {
  if (check_ctx(c)) return GATT_E_INVALID;
  if (dir < 0 || dir > 2 || (!c->is3d && dir == 2) || (kind != GATT_DERIV && kind != GATT_FILTER))
    return fail(GATT_E_INVALID, "direction or kind out of range");
  const LineShape& s = c->dir[dir].shape[kind];
  if (variant) *variant = s.variant;
  if (G) *G = s.G;
  if (M) *M = s.M;
  if (wpb) *wpb = s.waves_per_block;
  return GATT_OK;
}

int gatt_time_kernel(gatt_ctx* c, int32_t kernel, int32_t reps, double* avg_ms, double* bytes) {
  if (check_ctx(c)) return GATT_E_INVALID;
  if (reps < 1) return fail(GATT_E_INVALID, "reps < 1");
  const long long np = c->np;
  const int ndir = c->is3d ? 3 : 2;
  std::vector<int> flds = c->is3d ? std::vector<int>{0, 1, 2, 3, 4} : std::vector<int>{0, 1, 3, 4};
  double* S = c->state(c->cur);
  double* D = c->work_ptr(WK_D);
  double* T = c->work_ptr(WK_T);
  double* G = c->work_ptr(WK_G);
  const long long ws = c->work_stride;
  const int nm = c->wgroup;  // the first work group
  int members = nm;
  int launches = 0;
  double arrays = 0;  // arrays touched per rep
  auto body = [&]() {
    if (kernel >= GATT_K_DERIV_X && kernel <= GATT_K_DERIV_Z) {
      const int d = kernel - GATT_K_DERIV_X;
      double* in[kMaxPtr];
      double* out[kMaxPtr];
      int nf = 0;
      for (int f : flds) { in[nf] = S + (long long)f * np; out[nf] = D + (long long)(3 * f + d) * np; ++nf; }
      run_solve(c, d, GATT_DERIV, nf, in, out, 0.0, 0, nm, c->slab, ws);
      launches = 1;
      arrays = 2.0 * nf;
    } else if (kernel == GATT_K_FLUX_X || kernel == GATT_K_FLUX_Z) {
      // flux derivatives, the fluxes computed from E and the first derivatives as they are loaded
      const int d = kernel == GATT_K_FLUX_X ? 0 : 2;
      const int tin3[3][4] = {{0, 3, 4, 6}, {3, 1, 5, 7}, {4, 5, 2, 8}};
      double* out[kMaxPtr];
      int q[kMaxPtr];
      int nf = 0;
      for (int j = 0; j < 4; ++j) {
        if (!c->is3d && j == 2) continue;
        q[nf] = tin3[d][j];
        out[nf] = G + (long long)(4 * d + j) * np;
        ++nf;
      }
      run_solve_flux(c, d, nf, q, out, nm, S + 4 * np, c->slab, D, ws, c->stream);
      launches = 1;
      // distinct inputs (E and 8 first derivatives; 2-D: E and 5) and the outputs
      arrays = c->is3d ? 9 + 4 : 6 + 3;
    } else if (kernel == GATT_K_RHS_Y) {
      AssembleArgs aa{};
      aa.c = c->fc; aa.np = np; aa.member_stride = c->slab; aa.work_stride = ws; aa.S = c->slab_ptr(SL_S); aa.Q = S;
      aa.Out = c->slab_ptr(SL_S); aa.Delta = c->slab_ptr(SL_DL); aa.D = D; aa.T = T; aa.G = G;
      aa.sums = nullptr; aa.stage = 2; aa.dt = 0.0; aa.is3d = c->is3d;
      set_pressure_gradient(c, aa);
      RhsLineArgs r{};
      fill_solve_args(c, 1, GATT_DERIV, r.s);
      r.a = aa;
      launch_rhs_line(r, nm, c->stream);
      launches = 1;
      // reads S(5) D(15) Gx(4) Gz(4) Q(5) Delta(5), writes Delta(5) Out(5); 2-D: 4, 8, 3, 0, 4, 4, 4, 4
      arrays = c->is3d ? 48 : 31;
    } else if (kernel >= GATT_K_FILTER_X && kernel <= GATT_K_FILTER_Z) {
      const int d = kernel - GATT_K_FILTER_X;
      double* ptr[kMaxPtr];
      int nf = 0;
      for (int f : flds) ptr[nf++] = S + (long long)f * np;
      run_solve(c, d, GATT_FILTER, nf, ptr, ptr, -1.0, 1, c->batch, c->slab, c->slab);
      members = c->batch;
      launches = 1;
      arrays = 2.0 * nf;
    } else if (kernel == GATT_K_VISC) {
      ViscArgs va{};
      va.c = c->fc; va.np = np; va.member_stride = c->slab; va.work_stride = ws; va.E = S + 4 * np;
      va.D = D; va.T = T; va.is3d = c->is3d;
      launch_visc(va, nm, c->stream);
      launches = 1;
      arrays = c->is3d ? 13 + 9 : 7 + 5;
    } else {
      AssembleArgs aa{};
      aa.c = c->fc; aa.np = np; aa.member_stride = c->slab; aa.work_stride = ws; aa.S = c->slab_ptr(SL_S); aa.Q = S;
      aa.Out = c->slab_ptr(SL_S); aa.Delta = c->slab_ptr(SL_DL); aa.D = D; aa.T = T; aa.G = G;
      aa.sums = nullptr; aa.stage = 2; aa.dt = 0.0; aa.is3d = c->is3d;
      set_pressure_gradient(c, aa);
      launch_assemble(aa, nm, c->stream);
      launches = 1;
      // reads S(5) D(15) G(12) Q(5) Delta(5), writes Delta(5) Out(5); 2-D: 4, 8, 6, 4, 4, 4, 4
      arrays = c->is3d ? 52 : 34;
    }
  };
  if (kernel == GATT_K_STEP) {
    // the graph of two steps gatt_advance replays; each replay leaves cur where it was
    if (!c->graph) return fail(GATT_E_INVALID, "advance at least two steps before timing the step graph");
    HIPCHK(hipGraphLaunch(c->graph, c->stream));
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    HIPCHK(hipEventRecord(e0, c->stream));
    for (int r = 0; r < reps; ++r) HIPCHK(hipGraphLaunch(c->graph, c->stream));
    HIPCHK(hipEventRecord(e1, c->stream));
    HIPCHK(hipEventSynchronize(e1));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    HIPCHK(hipGetLastError());
    int sfd = 0;
    if (c->sfdType == 2 && c->sfdDelta > 0) sfd = c->is3d ? 11 : 9;
    else if (c->sfdType == 1 && c->sfdDelta > 0) sfd = c->is3d ? 10 : 8;
    else if (c->sfdType > 0) sfd = c->is3d ? 6 : 5;
    if (avg_ms) *avg_ms = (double)ms / (2.0 * reps);
    if (bytes) *bytes = 8.0 * ((c->is3d ? 377 : 196) + sfd) * (double)np * c->batch;
    return GATT_OK;
  }
  if (kernel < 0 || kernel > GATT_K_RHS_Y ||
      (!c->is3d && (kernel == GATT_K_DERIV_Z || kernel == GATT_K_FILTER_Z || kernel == GATT_K_FLUX_Z)))
    return fail(GATT_E_INVALID, "unknown kernel class");
  if ((kernel == GATT_K_FLUX_X && !flux_solve_ok(c, 0)) || (kernel == GATT_K_FLUX_Z && !flux_solve_ok(c, 2)) ||
      (kernel == GATT_K_RHS_Y && !c->fuse_rhs))
    return fail(GATT_E_UNSUPPORTED, "this kernel class is not part of this context's step");
  (void)ndir;
  body();  // warm-up
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  HIPCHK(hipEventRecord(e0, c->stream));
  for (int r = 0; r < reps; ++r) body();
  HIPCHK(hipEventRecord(e1, c->stream));
  HIPCHK(hipEventSynchronize(e1));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, e0, e1));
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  HIPCHK(hipGetLastError());
  if (avg_ms) *avg_ms = (double)ms / (reps * launches);
  if (bytes) *bytes = arrays / launches * 8.0 * (double)np * members;
  return GATT_OK;
}


// ------------------------------------------------------------ stability analysis

static long long packed_points(const gatt_ctx* c) {
  return c->single_beta ? (long long)c->nx * c->ny : c->np;
}
static long long flow_points(const gatt_ctx* c) { return c->nflow > 0 ? c->nflow : packed_points(c); }
static int slots(const gatt_ctx* c) { return c->is3d ? 5 : 4; }

static int ensure_scratch(gatt_ctx* c, size_t doubles) {
  if (doubles <= c->scratch_doubles) return GATT_OK;
  HIPCHK(hipStreamSynchronize(c->stream));
  if (c->d_scratch) HIPCHK(hipFree(c->d_scratch));
  c->d_scratch = nullptr;
  c->scratch_doubles = 0;
  void* p = nullptr;
  if (hipMalloc(&p, sizeof(double) * doubles) != hipSuccess) return fail(GATT_E_NOMEM, "scratch allocation failed");
  c->d_scratch = static_cast<double*>(p);
  c->scratch_doubles = doubles;
  return GATT_OK;
}

static bool ranges_overlap(const double* a, long long na, const double* b, long long nb) {
  return a < b + nb && b < a + na;
}

int gatt_set_flow_region(gatt_ctx* c, int64_t count, const int64_t* idx) {
  if (check_ctx(c)) return GATT_E_INVALID;
  HIPCHK(hipStreamSynchronize(c->stream));
  if (c->d_idx) HIPCHK(hipFree(c->d_idx));
  if (c->d_rmap) HIPCHK(hipFree(c->d_rmap));
  c->d_idx = nullptr;
  c->d_rmap = nullptr;
  c->nflow = 0;
  const long long npk = packed_points(c);
  if (!idx || count == npk) {
    if (idx)
      for (int64_t r = 0; r < count; ++r)
        if (idx[r] != r) goto general;
    return GATT_OK;
  }
general:
  if (count < 1 || count > npk) return fail(GATT_E_INVALID, "flow region size out of range");
  if (c->np > 0x7fffffffLL) return fail(GATT_E_UNSUPPORTED, "flow regions need fewer than 2^31 points");
  {
    std::vector<int> fwd((size_t)count), inv((size_t)npk, -1);
    for (int64_t r = 0; r < count; ++r) {
      if (idx[r] < 0 || idx[r] >= npk) return fail(GATT_E_INVALID, "flow region index outside the grid");
      if (r > 0 && idx[r] <= idx[r - 1]) return fail(GATT_E_INVALID, "flow region indices must increase");
      fwd[(size_t)r] = (int)idx[r];
      inv[(size_t)idx[r]] = (int)r;
    }
    void* p = nullptr;
    HIPCHK(hipMalloc(&p, sizeof(int) * fwd.size()));
    c->d_idx = static_cast<int*>(p);
    HIPCHK(hipMalloc(&p, sizeof(int) * inv.size()));
    c->d_rmap = static_cast<int*>(p);
    HIPCHK(hipMemcpy(c->d_idx, fwd.data(), sizeof(int) * fwd.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->d_rmap, inv.data(), sizeof(int) * inv.size(), hipMemcpyHostToDevice));
  }
  c->nflow = count;
  return GATT_OK;
}

int gatt_set_single_beta(gatt_ctx* c, int32_t on) {
  if (check_ctx(c)) return GATT_E_INVALID;
  if (on && !c->is3d) return fail(GATT_E_INVALID, "the singleBeta packing needs a 3-D grid");
  if ((bool)on == c->single_beta) return GATT_OK;
  // the region refers to the other layout: back to every point
  int rc = gatt_set_flow_region(c, 0, nullptr);
  if (rc) return rc;
  c->single_beta = on != 0;
  if (c->single_beta && !c->d_zw) {
    // Ztemp = linspace(0, 2 pi, nz+1) without its end (calcInstability.m:495-497, 566-568)
    const int nz = c->nz;
    std::vector<double> w(4 * (size_t)nz);
    for (int k = 0; k < nz; ++k) {
      const double z = k * (2.0 * M_PI) / nz;
      w[k] = std::cos(z);
      w[nz + k] = std::sin(z);
      w[2 * nz + k] = std::cos(z) * 2 / nz;
      w[3 * nz + k] = std::sin(z) * 2 / nz;
    }
    if ((rc = dev_upload(c, w.data(), w.size(), &c->d_zw))) return rc;
  }
  return GATT_OK;
}

int gatt_vector_length(gatt_ctx* c, int64_t* n) {
  if (check_ctx(c) || !n) return fail(GATT_E_INVALID, "null argument");
  *n = (int64_t)slots(c) * flow_points(c);
  return GATT_OK;
}

int gatt_set_base(gatt_ctx* c, const double* const* q) {
  if (check_ctx(c) || !q) return fail(GATT_E_INVALID, "null argument");
  if (!c->d_base) {
    void* p = nullptr;
    if (hipMalloc(&p, sizeof(double) * 5 * (size_t)c->np) != hipSuccess)
      return fail(GATT_E_NOMEM, "base flow allocation failed");
    c->d_base = static_cast<double*>(p);
    HIPCHK(hipMemsetAsync(c->d_base, 0, sizeof(double) * 5 * (size_t)c->np, c->stream));
  }
  for (int f = 0; f < 5; ++f) {
    if (!c->is3d && f == GATT_W) continue;
    if (!q[f]) return fail(GATT_E_INVALID, "null field pointer");
    HIPCHK(hipMemcpyAsync(c->d_base + (long long)f * c->np, q[f], sizeof(double) * c->np, hipMemcpyHostToDevice,
                          c->stream));
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  return GATT_OK;
}

int gatt_perturb(gatt_ctx* c, int32_t first, int32_t count, const double* V, int64_t ldv, double scale) {
  if (check_ctx(c)) return GATT_E_INVALID;
  if (!c->d_base) return fail(GATT_E_INVALID, "no base flow set");
  if (first < 0 || count < 1 || first + count > c->batch) return fail(GATT_E_INVALID, "member range out of bounds");
  const long long nvec = (long long)slots(c) * flow_points(c);
  if (V && ldv < nvec) return fail(GATT_E_INVALID, "ldv smaller than the vector length");
  PerturbArgs a{};
  a.np = c->np;
  a.member_stride = c->slab;
  a.nflow = flow_points(c);
  a.ldv = ldv;
  a.state = c->state(c->cur) + (long long)first * c->slab;
  a.mean = c->sfdType > 0 ? c->slab_ptr(SL_MEAN) + (long long)first * c->slab : nullptr;
  a.base = c->d_base;
  a.rmap = c->d_rmap;
  a.V = V;
  a.scale = scale;
  a.is3d = c->is3d;
  a.single_beta = c->single_beta;
  a.plane = (long long)c->nx * c->ny;
  a.nz = c->nz;
  a.zw = c->d_zw;
  launch_perturb(a, count, c->stream);
  if (c->sfdType > 0) c->mean_ready = true;
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  return GATT_OK;
}

int gatt_collect(gatt_ctx* c, int32_t first, int32_t count, int32_t minus_first, const double* ref, double scale,
                 double* W, int64_t ldw) {
  if (check_ctx(c) || !W) return fail(GATT_E_INVALID, "null argument");
  if (first < 0 || count < 1 || first + count > c->batch) return fail(GATT_E_INVALID, "member range out of bounds");
  if (minus_first >= 0 && minus_first + count > c->batch) return fail(GATT_E_INVALID, "subtracted members out of bounds");
  const long long nvec = (long long)slots(c) * flow_points(c);
  if (ldw < nvec) return fail(GATT_E_INVALID, "ldw smaller than the vector length");
  CollectArgs a{};
  a.np = c->np;
  a.member_stride = c->slab;
  a.nflow = flow_points(c);
  a.ldw = ldw;
  a.plus = c->state(c->cur) + (long long)first * c->slab;
  a.minus = minus_first >= 0 ? c->state(c->cur) + (long long)minus_first * c->slab : nullptr;
  a.ref = minus_first >= 0 ? nullptr : ref;
  a.idx = c->d_idx;
  a.W = W;
  a.scale = scale;
  a.is3d = c->is3d;
  a.single_beta = c->single_beta;
  a.plane = (long long)c->nx * c->ny;
  a.nz = c->nz;
  a.zw = c->d_zw;
  launch_collect(a, count, c->stream);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  return GATT_OK;
}

int gatt_block_dot(gatt_ctx* c, int64_t n, int32_t k, int32_t p, const double* X, int64_t ldx, const double* Y,
                   int64_t ldy, double* H) {
  if (check_ctx(c) || !X || !Y || !H) return fail(GATT_E_INVALID, "null argument");
  if (n < 1 || k < 1 || p < 1 || ldx < n || ldy < n) return fail(GATT_E_INVALID, "bad block dimensions");
  const int nch = block_dot_chunks(n);
  const size_t kp = (size_t)k * p;
  int rc = ensure_scratch(c, kp * (nch + 1));
  if (rc) return rc;
  BlockDotArgs a{};
  a.n = n;
  a.ldx = ldx;
  a.ldy = ldy;
  a.k = k;
  a.p = p;
  a.X = X;
  a.Y = Y;
  a.partial = c->d_scratch + kp;
  launch_block_dot(a, c->d_scratch, c->stream);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(H, c->d_scratch, sizeof(double) * kp, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return GATT_OK;
}

int gatt_block_dot_noalloc(gatt_ctx* c, int64_t n, int32_t k, int32_t p, const double* X, int64_t ldx,
                           const double* Y, int64_t ldy, double* H) {
  if (check_ctx(c) || !X || !Y || !H) return fail(GATT_E_INVALID, "null argument");
  if (n < 1 || k < 1 || p < 1 || ldx < n || ldy < n) return fail(GATT_E_INVALID, "bad block dimensions");
  // the same chunking and kernels as gatt_block_dot, the partials in the static scratch
  const int nch = block_dot_chunks(n);
  const size_t kp = (size_t)k * p;
  if (kp * ((size_t)nch + 1) > kStaticScratchDoubles)
    return fail(GATT_E_INVALID, "block product needs " + std::to_string(kp * ((size_t)nch + 1)) +
                                    " scratch doubles, more than the allocation-free limit of " +
                                    std::to_string(kStaticScratchDoubles));
  double* sp = g_scratch[c->device];
  if (!sp) return fail(GATT_E_INVALID, "static scratch not set up on this device");
  // every partial the reduction reads is written by the product kernel first, so the
  // scratch needs no clearing between calls
  BlockDotArgs a{};
  a.n = n;
  a.ldx = ldx;
  a.ldy = ldy;
  a.k = k;
  a.p = p;
  a.X = X;
  a.Y = Y;
  a.partial = sp + kp;
  launch_block_dot(a, sp, c->stream);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(H, sp, sizeof(double) * kp, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return GATT_OK;
}

int gatt_block_update(gatt_ctx* c, int64_t n, int32_t k, int32_t p, const double* X, int64_t ldx, const double* M,
                      double alpha, double beta, double* Y, int64_t ldy) {
  if (check_ctx(c) || !X || !M || !Y) return fail(GATT_E_INVALID, "null argument");
  if (n < 1 || k < 1 || p < 1 || ldx < n || ldy < n) return fail(GATT_E_INVALID, "bad block dimensions");
  if (ranges_overlap(X, (long long)(k - 1) * ldx + n, Y, (long long)(p - 1) * ldy + n))
    return fail(GATT_E_INVALID, "X and Y overlap");
  // rows of M in chunks of 256 so that a chunk of 32 columns fits 64 KiB of LDS;
  // chunk 0 applies alpha, later chunks accumulate
  constexpr int kChunk = 256;
  const size_t kp = (size_t)k * p;
  const int rc = ensure_scratch(c, k > kChunk ? (size_t)kChunk * p : kp);
  if (rc) return rc;
  for (int i0 = 0; i0 < k; i0 += kChunk) {
    const int kc = std::min(kChunk, k - i0);
    if (kc == k) {
      HIPCHK(hipMemcpyAsync(c->d_scratch, M, sizeof(double) * kp, hipMemcpyHostToDevice, c->stream));
    } else {
      // the chunk's rows of M as a kc x p column-major block
      std::vector<double> sub((size_t)kc * p);
      for (int j = 0; j < p; ++j)
        for (int i = 0; i < kc; ++i) sub[(size_t)j * kc + i] = M[(size_t)j * k + i0 + i];
      HIPCHK(hipMemcpyAsync(c->d_scratch, sub.data(), sizeof(double) * sub.size(), hipMemcpyHostToDevice,
                            c->stream));
      HIPCHK(hipStreamSynchronize(c->stream));
    }
    BlockCombineArgs a{};
    a.n = n;
    a.ldx = ldx;
    a.ldy = ldy;
    a.k = kc;
    a.p = p;
    a.X = X + (long long)i0 * ldx;
    a.M = c->d_scratch;
    a.alpha = i0 == 0 ? alpha : 1.0;
    a.beta = beta;
    a.Y = Y;
    launch_block_combine(a, c->stream);
    // the next chunk overwrites the staged M
    if (i0 + kChunk < k) HIPCHK(hipStreamSynchronize(c->stream));
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  return GATT_OK;
}


// ------------------------------------------------------------ ragged groups

}  // extern "C"

struct gatt_group {
  std::vector<gatt_ctx*> ctx;
  int mode = 0;                         // 0: one graph with a branch per member; 1: per-member graphs;
                                        // 2: one graph of ragged launches (every member in each launch)
  hipGraphExec_t graph = nullptr;
  void* ragged_buf = nullptr;           // mode 2: argument blocks and block tables of the graph's launches
  int ragged_launches = 0;              // mode 2: launches of one double step
  std::vector<std::array<gatt::LineShape, 6>> shapes0;   // mode 2: the members' own launch shapes
  std::vector<hipGraphExec_t> mgraph;   // mode 1: one double-step graph per member
  std::vector<double> graph_dt;
  std::vector<int> graph_cur;
  std::vector<hipEvent_t> ev;           // fork/join events (members + 1)
};

namespace {

void group_drop_graph(gatt_group* g) {
  if (g->graph) {
    (void)hipGraphExecDestroy(g->graph);
    g->graph = nullptr;
  }
  if (g->ragged_buf) {
    // the graph's launches read the argument blocks: none may be in flight
    for (gatt_ctx* c : g->ctx) (void)hipStreamSynchronize(c->stream);
    (void)hipFree(g->ragged_buf);
    g->ragged_buf = nullptr;
  }
  for (hipGraphExec_t& x : g->mgraph)
    if (x) (void)hipGraphExecDestroy(x);
  g->mgraph.clear();
}

// One double step of every member, as parallel branches forked from the first
// member's stream and joined back into it.  Inside the group graph every member
// launches its direction solves in order on its own stream (the group already
// overlaps whole members); the members' own setting is restored afterwards, so
// their gatt_advance graphs keep the parallel direction branches.
int group_capture(gatt_group* g, const double* dt) {
  gatt_ctx* c0 = g->ctx[0];
  hipGraph_t graph;
  std::vector<char> fork(g->ctx.size());
  for (size_t m = 0; m < g->ctx.size(); ++m) {
    fork[m] = g->ctx[m]->fork_dirs;
    g->ctx[m]->fork_dirs = false;
  }
  auto restore = [&]() {
    for (size_t m = 0; m < g->ctx.size(); ++m) g->ctx[m]->fork_dirs = fork[m] != 0;
  };
  hipError_t e = hipStreamBeginCapture(c0->stream, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) {
    restore();
    return fail(GATT_E_HIP, std::string("hipStreamBeginCapture: ") + hipGetErrorString(e));
  }
  (void)hipEventRecord(g->ev[0], c0->stream);
  for (size_t m = 1; m < g->ctx.size(); ++m) (void)hipStreamWaitEvent(g->ctx[m]->stream, g->ev[0], 0);
  for (size_t m = 0; m < g->ctx.size(); ++m) {
    gatt_ctx* c = g->ctx[m];
    enqueue_step(c, c->cur, dt[m]);
    enqueue_step(c, 1 - c->cur, dt[m]);
  }
  for (size_t m = 1; m < g->ctx.size(); ++m) {
    (void)hipEventRecord(g->ev[m], g->ctx[m]->stream);
    (void)hipStreamWaitEvent(c0->stream, g->ev[m], 0);
  }
  e = hipStreamEndCapture(c0->stream, &graph);
  restore();
  if (e != hipSuccess) return fail(GATT_E_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
  e = hipGraphInstantiate(&g->graph, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (e != hipSuccess) return fail(GATT_E_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
  g->graph_dt.assign(dt, dt + g->ctx.size());
  g->graph_cur.clear();
  for (gatt_ctx* c : g->ctx) g->graph_cur.push_back(c->cur);
  return GATT_OK;
}

// Mode 2: one launch shape per direction and operator that every member can run with the
// same kernel instantiation and workgroup size.  The member with the longest lines picks
// the shape; the one-wavefront x solve (variant 4) and the lane-group solve (variant 0)
// take any shorter line with the same points per lane, the wave-segment variants (1-3)
// size their workgroups by the line length and are replaced by the lane-group solve
// when the members' lines differ in length.
int group_harmonise_shapes(gatt_group* g) {
  const gatt_ctx* c0 = g->ctx[0];
  const int ndir = c0->is3d ? 3 : 2;
  for (int d = 0; d < ndir; ++d)
    for (int kind : {GATT_DERIV, GATT_FILTER}) {
      // y and z lines: the wave-segment solve (lane = line) with one workgroup size for all,
      // the segment count of the longest line, when every member's tiles carry one operator
      // type, no line is periodic and the members bring enough lines between them to fill it
      if (d > 0) {
        bool ok = true;
        long long lines = 0;
        int nmax = 0;
        for (gatt_ctx* c : g->ctx) {
          const DirOps& o = c->dir[d];
          ok = ok && !o.identity[kind] && o.uniform_tiles && !o.op[kind].periodic;
          lines += o.nlines;
          nmax = std::max(nmax, o.op[kind].N);
        }
        bool done = false;
        for (int M : {16, 32}) {
          if (!ok || lines < 1024 || done) break;
          LineShape t;
          t.variant = 2;
          t.M = M;
          t.G = (nmax + M - 1) / M;
          t.waves_per_block = t.G;
          bool fits = true;
          for (gatt_ctx* c : g->ctx)
            fits = fits && line_shape_ok(t, c->dir[d].op[kind].N, c->dir[d].op[kind].nR, d, c->max_lds, true);
          if (!fits) continue;
          for (gatt_ctx* c : g->ctx) {
            c->dir[d].shape[kind] = t;
            c->drop_graphs();
          }
          done = true;
        }
        if (done) continue;
      }
      int big = -1;
      bool same_n = true;
      for (size_t m = 0; m < g->ctx.size(); ++m) {
        const DirOps& o = g->ctx[m]->dir[d];
        if (o.identity[kind]) continue;
        if (big >= 0 && o.op[kind].N != g->ctx[big]->dir[d].op[kind].N) same_n = false;
        if (big < 0 || o.op[kind].N > g->ctx[big]->dir[d].op[kind].N) big = (int)m;
      }
      if (big < 0) continue;
      const DirOps& ob = g->ctx[big]->dir[d];
      LineShape s = ob.shape[kind];
      if (s.variant >= 1 && s.variant <= 3 && !same_n) {
        s.variant = -1;
        const int N = ob.op[kind].N;
        // a lane-group shape whose points per lane every member's lines map back to (each
        // member then takes its own lanes per line, below)
        auto common = [&](const LineShape& t) {
          for (gatt_ctx* c : g->ctx) {
            const DirOps& o = c->dir[d];
            if (o.identity[kind]) continue;
            const int Nm = o.op[kind].N;
            LineShape u = t;
            u.G = std::min(64, std::max(2, pow2_at_least((Nm + t.M - 1) / t.M)));
            if (std::max(4, pow2_at_least((Nm + u.G - 1) / u.G)) != t.M) return false;
            if (!line_shape_ok(u, Nm, o.op[kind].nR, d, c->max_lds, o.uniform_tiles)) return false;
          }
          return true;
        };
        for (int M : {16, 8, 4, 32}) {
          for (int wpb : {4, 2, 1}) {
            LineShape t;
            t.M = M;
            t.G = std::max(2, pow2_at_least((N + M - 1) / M));
            t.waves_per_block = wpb;
            if (line_shape_ok(t, N, ob.op[kind].nR) && common(t)) {
              s = t;
              break;
            }
          }
          if (s.variant == 0) break;
        }
        if (s.variant != 0) return fail(GATT_E_UNSUPPORTED, "ragged group: no common line shape");
      }
      for (gatt_ctx* c : g->ctx) {
        DirOps& o = c->dir[d];
        if (o.identity[kind]) continue;
        const int N = o.op[kind].N;
        LineShape t = s;
        if (s.variant == 4) {
          // one wavefront per line: the ragged form takes each member's points per lane, so
          // every member gets the fewest that cover its lines (all 64 lanes busy)
          for (int M : {4, 8, 16}) {
            t.M = M;
            if (64 * M >= N && line_shape_ok(t, N, o.op[kind].nR, d, c->max_lds, o.uniform_tiles)) break;
          }
        } else if (s.variant == 0) {
          t.G = std::min(64, std::max(2, pow2_at_least((N + s.M - 1) / s.M)));
          if (std::max(4, pow2_at_least((N + t.G - 1) / t.G)) != s.M)
            return fail(GATT_E_UNSUPPORTED, "ragged group: line lengths too far apart for one lane-group shape");
        } else if (s.variant >= 1 && s.variant <= 3) {
          t.G = (N + s.M - 1) / s.M;
        }
        if (!line_shape_ok(t, N, o.op[kind].nR, d, c->max_lds, o.uniform_tiles))
          return fail(GATT_E_UNSUPPORTED, "ragged group: a member cannot take the common line shape");
        o.shape[kind] = t;
        c->drop_graphs();
      }
    }
  return GATT_OK;
}

// Mode 2: record the double step of every member (the launches it would make, with their
// argument blocks and grids) and merge the i-th launches of all members into one ragged
// launch.  Boundary applications are merged wave by wave (a member with fewer waves takes
// no blocks in the later ones).  The merged launches are captured into one graph on the
// first member's stream; their argument blocks and block tables live in one device buffer
// owned by the group.
int group_capture_ragged(gatt_group* g, const double* dt) {
  const size_t n = g->ctx.size();
  std::vector<LaunchRecorder> rec(n);
  for (size_t m = 0; m < n; ++m) {
    gatt_ctx* c = g->ctx[m];
    const bool fork = c->fork_dirs;
    c->fork_dirs = false;
    g_recorder = &rec[m];
    enqueue_step(c, c->cur, dt[m]);
    enqueue_step(c, 1 - c->cur, dt[m]);
    g_recorder = nullptr;
    c->fork_dirs = fork;
  }
  // each member's sequence as segments: a single launch, or the waves of one boundary application
  struct Seg {
    int bc;
    std::vector<const LaunchRecord*> r;
  };
  std::vector<std::vector<Seg>> segs(n);
  for (size_t m = 0; m < n; ++m)
    for (const LaunchRecord& r : rec[m].recs) {
      if (r.bc == -2) return fail(GATT_E_UNSUPPORTED, "ragged group: a member's step has a launch without a ragged form");
      if (!r.fn) segs[m].push_back(Seg{r.bc, {}});
      else if (r.bc >= 0) segs[m].back().r.push_back(&r);
      else segs[m].push_back(Seg{-1, {&r}});
    }
  for (size_t m = 1; m < n; ++m) {
    if (segs[m].size() != segs[0].size()) return fail(GATT_E_UNSUPPORTED, "ragged group: the members' steps differ");
    for (size_t k = 0; k < segs[0].size(); ++k) {
      const Seg& a = segs[0][k];
      const Seg& b = segs[m][k];
      if ((a.bc >= 0) != (b.bc >= 0)) return fail(GATT_E_UNSUPPORTED, "ragged group: the members' steps differ");
      if (a.bc < 0 && (a.r[0]->fn != b.r[0]->fn || a.r[0]->args.size() != b.r[0]->args.size() ||
                       a.r[0]->block.x != b.r[0]->block.x))
        return fail(GATT_E_UNSUPPORTED, "ragged group: the members' launches differ (kernel or workgroup size) at launch " +
                                            std::to_string(k) + " of member " + std::to_string(m));
    }
  }
  // the merged launches: per launch one record (or none) per member
  std::vector<std::vector<const LaunchRecord*>> launches;
  for (size_t k = 0; k < segs[0].size(); ++k) {
    size_t waves = 0;
    for (size_t m = 0; m < n; ++m) waves = std::max(waves, segs[m][k].r.size());
    for (size_t w = 0; w < waves; ++w) {
      std::vector<const LaunchRecord*> l(n, nullptr);
      for (size_t m = 0; m < n; ++m)
        if (w < segs[m][k].r.size()) l[m] = segs[m][k].r[w];
      launches.push_back(l);
    }
  }
  // device layout: per launch the n argument blocks, then start[n + 1] and gx[n]
  auto align = [](size_t v) { return (v + 255) & ~(size_t)255; };
  struct Plan {
    const void* fn;
    dim3 block;
    size_t shmem;
    size_t args_off, start_off, gx_off;
    int total;
  };
  std::vector<Plan> plan;
  size_t bytes = 0;
  for (const auto& l : launches) {
    Plan p{};
    size_t asz = 0;
    for (const LaunchRecord* r : l)
      if (r) {
        if (p.fn && (r->fn != p.fn || r->block.x != p.block.x))
          return fail(GATT_E_UNSUPPORTED, "ragged group: the members' launches differ (kernel or workgroup size)");
        p.fn = r->fn;
        p.block = r->block;
        p.shmem = std::max(p.shmem, r->shmem);
        asz = r->args.size();
      }
    p.args_off = bytes;
    bytes = align(bytes + asz * n);
    p.start_off = bytes;
    bytes = align(bytes + sizeof(int) * (n + 1));
    p.gx_off = bytes;
    bytes = align(bytes + sizeof(int) * n);
    plan.push_back(p);
  }
  std::vector<char> host(bytes, 0);
  for (size_t i = 0; i < launches.size(); ++i) {
    Plan& p = plan[i];
    int* start = reinterpret_cast<int*>(host.data() + p.start_off);
    int* gx = reinterpret_cast<int*>(host.data() + p.gx_off);
    long long total = 0;
    for (size_t m = 0; m < n; ++m) {
      const LaunchRecord* r = launches[i][m];
      start[m] = (int)total;
      gx[m] = 1;
      if (!r) continue;
      std::memcpy(host.data() + p.args_off + m * r->args.size(), r->args.data(), r->args.size());
      gx[m] = (int)r->gx;
      total += (long long)r->gx * r->gy;
    }
    if (total <= 0 || total > 0x7fffffffLL) return fail(GATT_E_UNSUPPORTED, "ragged group: launch size out of range");
    start[n] = (int)total;
    p.total = (int)total;
  }
  group_drop_graph(g);
  HIPCHK(hipMalloc(&g->ragged_buf, std::max<size_t>(bytes, 256)));
  HIPCHK(hipMemcpy(g->ragged_buf, host.data(), bytes, hipMemcpyHostToDevice));
  char* base = static_cast<char*>(g->ragged_buf);
  for (const Plan& p : plan)
    if (p.shmem > 65536)
      HIPCHK(hipFuncSetAttribute(p.fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)p.shmem));
  gatt_ctx* c0 = g->ctx[0];
  hipGraph_t graph;
  HIPCHK(hipStreamBeginCapture(c0->stream, hipStreamCaptureModeThreadLocal));
  hipError_t e = hipSuccess;
  for (const Plan& p : plan) {
    const void* args = base + p.args_off;
    RaggedMap map{reinterpret_cast<const int*>(base + p.start_off), reinterpret_cast<const int*>(base + p.gx_off),
                  (int)n};
    void* params[] = {&args, &map};
    e = hipLaunchKernel(p.fn, dim3((unsigned)p.total), p.block, params, p.shmem, c0->stream);
    if (e != hipSuccess) break;
  }
  const hipError_t e2 = hipStreamEndCapture(c0->stream, &graph);
  if (e != hipSuccess || e2 != hipSuccess)
    return fail(GATT_E_HIP, std::string("ragged group capture: ") + hipGetErrorString(e != hipSuccess ? e : e2));
  e = hipGraphInstantiate(&g->graph, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (e != hipSuccess) return fail(GATT_E_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
  g->ragged_launches = (int)plan.size();
  g->graph_dt.assign(dt, dt + n);
  g->graph_cur.clear();
  for (gatt_ctx* c : g->ctx) g->graph_cur.push_back(c->cur);
  return GATT_OK;
}

}  // namespace

extern "C" {

int gatt_group_create(gatt_ctx* const* ctxs, int32_t n, int32_t mode, gatt_group** out) {
  if (!ctxs || !out || n < 1) return fail(GATT_E_INVALID, "bad group arguments");
  if (mode < 0 || mode > 2) return fail(GATT_E_INVALID, "group mode must be 0, 1 or 2");
  *out = nullptr;
  std::unique_ptr<gatt_group> g(new gatt_group);
  for (int32_t m = 0; m < n; ++m) {
    if (!ctxs[m]) return fail(GATT_E_INVALID, "null context in group");
    for (int32_t k = 0; k < m; ++k)
      if (ctxs[k] == ctxs[m]) return fail(GATT_E_INVALID, "a context appears twice in the group");
    g->ctx.push_back(ctxs[m]);
  }
  g->mode = mode;
  if (mode == 2) {
    for (gatt_ctx* c : g->ctx) {
      if (c->is3d != g->ctx[0]->is3d) return fail(GATT_E_UNSUPPORTED, "ragged group: 2-D and 3-D members mixed");
      if (c->fuse_rhs) return fail(GATT_E_UNSUPPORTED, "ragged group: the fused stage kernel has no ragged form");
      std::array<LineShape, 6> sh;
      for (int d = 0; d < 3; ++d)
        for (int k = 0; k < 2; ++k) sh[2 * d + k] = c->dir[d].shape[k];
      g->shapes0.push_back(sh);
    }
    const int rc = group_harmonise_shapes(g.get());
    if (rc) {
      for (size_t m = 0; m < g->ctx.size(); ++m)
        for (int d = 0; d < 3; ++d)
          for (int k = 0; k < 2; ++k) g->ctx[m]->dir[d].shape[k] = g->shapes0[m][2 * d + k];
      return rc;
    }
  }
  g->ev.resize(n);
  for (auto& e : g->ev) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  *out = g.release();
  return GATT_OK;
}

int gatt_group_destroy(gatt_group* g) {
  if (!g) return GATT_OK;
  for (gatt_ctx* c : g->ctx) (void)hipStreamSynchronize(c->stream);
  group_drop_graph(g);
  // mode 2: the members get their own launch shapes back
  for (size_t m = 0; m < g->shapes0.size(); ++m) {
    gatt_ctx* c = g->ctx[m];
    for (int d = 0; d < 3; ++d)
      for (int k = 0; k < 2; ++k) c->dir[d].shape[k] = g->shapes0[m][2 * d + k];
    c->drop_graphs();
  }
  for (auto& e : g->ev) (void)hipEventDestroy(e);
  delete g;
  return GATT_OK;
}

int gatt_group_advance(gatt_group* g, int32_t nsteps, const double* t0, const double* dt, double* maxchange,
                       int32_t* diverged) {
  if (!g || !t0 || !dt) return fail(GATT_E_INVALID, "null argument");
  if (nsteps < 0) return fail(GATT_E_INVALID, "nsteps < 0");
  const size_t n = g->ctx.size();
  for (size_t m = 0; m < n; ++m) {
    launch_clock_set(g->ctx[m]->d_clock, t0[m], g->ctx[m]->stream);
    launch_clear_i32(g->ctx[m]->d_flag, 1, g->ctx[m]->stream);
  }
  for (gatt_ctx* c : g->ctx) {
    ensure_mean(c);
    if (c->d_hold && !c->hold_captured) run_bc(c, c->state(c->cur), 0.0);
  }
  if (g->mode == 1) {
    // per-member double-step graphs owned by the group (direction solves in order on
    // the member's stream, as in the group graph), launched round-robin on the members' streams
    bool stale = g->mgraph.size() != n;
    for (size_t m = 0; m < n && !stale; ++m)
      stale = !g->mgraph[m] || g->graph_dt[m] != dt[m] || g->graph_cur[m] != g->ctx[m]->cur;
    if (stale) group_drop_graph(g);
    if (nsteps >= 2 && g->mgraph.empty()) {
      g->mgraph.assign(n, nullptr);
      for (size_t m = 0; m < n; ++m) {
        gatt_ctx* c = g->ctx[m];
        const bool fork = c->fork_dirs;
        c->fork_dirs = false;
        hipGraph_t gr;
        hipError_t e = hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal);
        if (e == hipSuccess) {
          enqueue_step(c, c->cur, dt[m]);
          enqueue_step(c, 1 - c->cur, dt[m]);
          e = hipStreamEndCapture(c->stream, &gr);
        }
        c->fork_dirs = fork;
        if (e == hipSuccess) {
          e = hipGraphInstantiate(&g->mgraph[m], gr, nullptr, nullptr, 0);
          (void)hipGraphDestroy(gr);
        }
        if (e != hipSuccess) {
          group_drop_graph(g);
          return fail(GATT_E_HIP, std::string("member graph capture: ") + hipGetErrorString(e));
        }
      }
      g->graph_dt.assign(dt, dt + n);
      g->graph_cur.clear();
      for (gatt_ctx* c : g->ctx) g->graph_cur.push_back(c->cur);
    }
    for (int s = 0; s + 2 <= nsteps; s += 2)
      for (size_t m = 0; m < n; ++m) HIPCHK(hipGraphLaunch(g->mgraph[m], g->ctx[m]->stream));
  } else {
    bool stale = !g->graph;
    for (size_t m = 0; m < n && !stale; ++m)
      stale = g->graph_dt[m] != dt[m] || g->graph_cur[m] != g->ctx[m]->cur;
    if (stale) group_drop_graph(g);
    if (nsteps >= 2 && !g->graph) {
      int rc = g->mode == 2 ? group_capture_ragged(g, dt) : group_capture(g, dt);
      if (rc) return rc;
    }
    // the graph runs on the first member's stream: it follows work pending on every member's stream
    if (nsteps >= 2)
      for (size_t m = 1; m < n; ++m) {
        HIPCHK(hipEventRecord(g->ev[m], g->ctx[m]->stream));
        HIPCHK(hipStreamWaitEvent(g->ctx[0]->stream, g->ev[m], 0));
      }
    for (int s = 0; s + 2 <= nsteps; s += 2) HIPCHK(hipGraphLaunch(g->graph, g->ctx[0]->stream));
    // later work on a member's own stream must follow the group's launches
    if (nsteps >= 2) {
      HIPCHK(hipEventRecord(g->ev[0], g->ctx[0]->stream));
      for (size_t m = 1; m < n; ++m) HIPCHK(hipStreamWaitEvent(g->ctx[m]->stream, g->ev[0], 0));
    }
  }
  if (nsteps % 2) {
    for (size_t m = 0; m < n; ++m) {
      gatt_ctx* c = g->ctx[m];
      enqueue_step(c, c->cur, dt[m]);
      c->cur = 1 - c->cur;
    }
    // the state buffers swapped: the captured double steps start from the other one
    group_drop_graph(g);
  }
  HIPCHK(hipGetLastError());
  for (size_t m = 0; m < n; ++m) {
    int rc = read_scalars(g->ctx[m], nullptr, maxchange ? maxchange + 5 * m : nullptr, diverged ? diverged + m : nullptr);
    if (rc) return rc;
  }
  return GATT_OK;
}


// ------------------------------------------------------ tangent and adjoint

}  // extern "C"

namespace {

// Per-member linear slab: arrays of np doubles.
enum : int {
  LN_Q = 0,     // tangent / adjoint state (the vector the caller sets and reads)
  LN_S = 5,     // tangent stage state / adjoint of a stage state
  LN_DL = 10,   // tangent RK accumulator / adjoint of a stage slope
  LN_B1 = 15,   // adjoint: stage states S1, S2, S3 of the base step
  LN_B2 = 20,
  LN_B3 = 25,
  LN_D = 30,    // 15 first-derivative tangents / adjoints
  LN_T = 45,    // 9 stress and heat-flux tangents / adjoints
  LN_G = 54,    // 12 second-derivative tangents / adjoints
  LN_B4 = 66,   // adjoint: the base step's state after the stage-4 boundary conditions
  LN_M = 71,    // tangent / adjoint of the SFD mean (evolving mean, Delta > 0)
  LN_N = 76
};

// the boundary programs of one base step write five undo logs (stages 1..4 and the
// final application), each undo_total doubles per member
constexpr int kUndoLogs = 5;

int ensure_lin(gatt_ctx* c) {
  if (c->lin) return GATT_OK;
  c->lin_stride = (long long)LN_N * c->np;
  void* p = nullptr;
  if (hipMalloc(&p, sizeof(double) * (size_t)c->lin_stride * c->batch) != hipSuccess)
    return fail(GATT_E_NOMEM, "tangent/adjoint allocation failed");
  c->lin = static_cast<double*>(p);
  if (c->undo_total > 0) {
    if (hipMalloc(&p, sizeof(double) * (size_t)kUndoLogs * c->undo_total * c->batch) != hipSuccess)
      return fail(GATT_E_NOMEM, "boundary undo log allocation failed");
    c->d_undo = static_cast<double*>(p);
    c->dev_allocs.push_back(p);
  }
  // on the context's stream, so that later copies into the slab are ordered after the clear
  HIPCHK(hipMemsetAsync(c->lin, 0, sizeof(double) * (size_t)c->lin_stride * c->batch, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return GATT_OK;
}

double* lin_ptr(gatt_ctx* c, int idx) { return c->lin + (long long)idx * c->np; }

int ensure_tape(gatt_ctx* c) {
  if (c->tape) return GATT_OK;
  c->tape_stride = 4LL * WK_N * c->np;
  void* p = nullptr;
  if (hipMalloc(&p, sizeof(double) * (size_t)c->tape_stride * c->batch) != hipSuccess)
    return fail(GATT_E_NOMEM, "adjoint stage tape allocation failed");
  c->tape = static_cast<double*>(p);
  return GATT_OK;
}

// derivative array `wk` (WK_D, WK_T, WK_G) of RK stage s (0..3) in the tape, member 0
double* tape_ptr(gatt_ctx* c, int s, int wk) { return c->tape + ((long long)s * WK_N + wk) * c->np; }

int lin_supported(gatt_ctx* c) {
  if (c->stepping != 1) return fail(GATT_E_UNSUPPORTED, "tangent and adjoint steps follow the RK4 integrator only");
  if (c->timed) return fail(GATT_E_UNSUPPORTED, "tangent and adjoint steps do not cover time-dependent disturbances");
  if (c->pg) return fail(GATT_E_UNSUPPORTED, "tangent and adjoint steps do not cover the pressureGradient forcing");
  for (int d = 0; d < (c->is3d ? 3 : 2); ++d)
    if (c->dir[d].identity[0] || c->dir[d].identity[1])
      return fail(GATT_E_UNSUPPORTED, "tangent and adjoint steps need a line operator in every direction");
  const int ndir = c->is3d ? 3 : 2;
  for (int d = 0; d < ndir; ++d)
    if (line_solve_t_points_per_lane(c->dir[d].op[0].N) == 0)
      return fail(GATT_E_UNSUPPORTED, "transposed line solves are compiled for lines of at most 512 points");
  return GATT_OK;
}

std::vector<int> state_fields(const gatt_ctx* c) {
  return c->is3d ? std::vector<int>{0, 1, 2, 3, 4} : std::vector<int>{0, 1, 3, 4};
}

// y = a x + b y over the state fields of every member
void fields_axpby(gatt_ctx* c, double* y, long long ys, const double* x, long long xs, double a, double b) {
  AxpbyArgs ar{};
  ar.np = c->np;
  ar.x_stride = xs;
  ar.y_stride = ys;
  ar.nf = 0;
  for (int f : state_fields(c)) {
    ar.x[ar.nf] = x + (long long)f * c->np;
    ar.y[ar.nf] = y + (long long)f * c->np;
    ++ar.nf;
  }
  ar.a = a;
  ar.b = b;
  launch_axpby(ar, c->batch, c->stream);
}

// the combinations of the backward RK4 pass in one launch (AdjCombineArgs)
void adj_combine(gatt_ctx* c, double* QB, const double* SB, double* ACC, double* KB, double a1, double a2, int first,
                 int mode) {
  AdjCombineArgs ar{};
  ar.np = c->np;
  ar.stride = c->lin_stride;
  for (int f : state_fields(c)) {
    const long long o = (long long)f * c->np;
    ar.sb[ar.nf] = SB + o;
    ar.qb[ar.nf] = QB + o;
    ar.acc[ar.nf] = ACC + o;
    ar.k[ar.nf] = KB ? KB + o : nullptr;
    ++ar.nf;
  }
  ar.a1 = a1;
  ar.a2 = a2;
  ar.first = first;
  ar.mode = mode;
  launch_adj_combine(ar, c->batch, c->stream);
}

// Base derivatives, stresses and their derivatives at S (member m0 pointer,
// member stride ss) into D, T, G (member stride ws), for nm members.
// The derivative solves of every direction (in[d], out[d]: nf fields of direction d) of the
// linearised steps' base passes, on the context's stream: one launch where the directions share
// a lane-group shape, otherwise a launch per direction.
void run_deriv_dirs(gatt_ctx* c, int ndir, int nf, double* const (*in)[kMaxPtr], double* const (*out)[kMaxPtr],
                    int nm, long long in_stride, long long out_stride) {
  if (run_solve_dirs_once(c, ndir, nf, in, out, nm, in_stride, out_stride)) return;
  for (int d = 0; d < ndir; ++d)
    run_solve(c, d, GATT_DERIV, nf, in[d], out[d], 0.0, 0, nm, in_stride, out_stride);
}

void base_dtg(gatt_ctx* c, const double* S, long long ss, int nm, double* D, double* T, double* G, long long ws) {
  const long long np = c->np;
  const int ndir = c->is3d ? 3 : 2;
  {
    double* in[3][kMaxPtr];
    double* out[3][kMaxPtr];
    int nf = 0;
    for (int d = 0; d < ndir; ++d) {
      nf = 0;
      for (int f : state_fields(c)) {
        in[d][nf] = const_cast<double*>(S) + (long long)f * np;
        out[d][nf] = D + (long long)(3 * f + d) * np;
        ++nf;
      }
    }
    run_deriv_dirs(c, ndir, nf, in, out, nm, ss, ws);
  }
  ViscArgs va{};
  va.c = c->fc;
  va.np = np;
  va.member_stride = ss;
  va.work_stride = ws;
  va.E = S + 4 * np;
  va.D = D;
  va.T = T;
  va.is3d = c->is3d;
  launch_visc(va, nm, c->stream);
  const int tin3[3][4] = {{0, 3, 4, 6}, {3, 1, 5, 7}, {4, 5, 2, 8}};
  {
    double* in[3][kMaxPtr];
    double* out[3][kMaxPtr];
    int nf = 0;
    for (int d = 0; d < ndir; ++d) {
      nf = 0;
      for (int q = 0; q < 4; ++q) {
        if (!c->is3d && q == 2) continue;
        in[d][nf] = T + (long long)tin3[d][q] * np;
        out[d][nf] = G + (long long)(4 * d + q) * np;
        ++nf;
      }
    }
    run_deriv_dirs(c, ndir, nf, in, out, nm, ws, ws);
  }
}

// The turbulence forcing of the linearised steps (turbulenceGenerator, 3-D).
bool lin_forcing(const gatt_ctx* c) { return c->forcing && c->is3d; }

// The forcing sums of the base stage state S (member stride ss) of every member into d_sums.
void base_forcing_sums(gatt_ctx* c, const double* S, long long ss) {
  if (!lin_forcing(c)) return;
  SumArgs sa{};
  sa.np = c->np;
  sa.member_stride = ss;
  sa.S = S;
  sa.sums = c->d_sums;
  sa.partial = c->d_partial;
  sa.nblocks = c->sum_blocks;
  launch_forcing_sums(sa, c->batch, c->stream);
}

// The linear sums of X (lin stride) against the base S (stride ss) of every member into d_lsums.
void lin_forcing_sums(gatt_ctx* c, const double* S, long long ss, const double* X) {
  if (!lin_forcing(c)) return;
  LinSumArgs la{};
  la.np = c->np;
  la.base_stride = ss;
  la.lin_stride = c->lin_stride;
  la.S = S;
  la.X = X;
  la.partial = c->d_partial;
  la.sums = c->d_lsums;
  la.nblocks = c->sum_blocks;
  launch_lin_sums(la, c->batch, c->stream);
}

// forward assembly of one stage for nm members from member m0 (the tail of
// run_rhs_stage) from the derivative arrays D, T, G (member stride ws)
void forward_assemble(gatt_ctx* c, double* S, double* Q, double* Out, int stage, double dt, int m0, int nm,
                      const double* D, const double* T, const double* G, long long ws) {
  AssembleArgs aa{};
  const long long mo = (long long)m0 * c->slab;
  aa.c = c->fc;
  aa.np = c->np;
  aa.member_stride = c->slab;
  aa.work_stride = ws;
  aa.S = S + mo;
  aa.Q = Q + mo;
  aa.Out = Out + mo;
  aa.Delta = c->slab_ptr(SL_DL) + mo;
  aa.D = D;
  aa.T = T;
  aa.G = G;
  // the turbulence forcing with the sums base_forcing_sums left for this stage state
  aa.sums = lin_forcing(c) ? c->d_sums + 8LL * m0 : nullptr;
  aa.forcing_eps = c->f_eps;
  aa.forcing_e0 = c->f_e0;
  aa.stage = stage;
  aa.dt = dt;
  aa.is3d = c->is3d;
  set_pressure_gradient(c, aa);
  launch_assemble(aa, nm, c->stream);
}

// Tangent of one RK stage for members m0..m0+nm-1 at the base stage state
// S (slab stride); base D, T, G must be in the work arrays.
void tl_stage(gatt_ctx* c, const double* S, double* dS, double* dQ, double* dOut, int stage, double dt, int m0,
              int nm) {
  const long long np = c->np, ls = c->lin_stride, lo = (long long)m0 * ls;
  double* dD = lin_ptr(c, LN_D) + lo;
  double* dT = lin_ptr(c, LN_T) + lo;
  double* dG = lin_ptr(c, LN_G) + lo;
  const int ndir = c->is3d ? 3 : 2;
  for (int d = 0; d < ndir; ++d) {
    double* in[kMaxPtr];
    double* out[kMaxPtr];
    int nf = 0;
    for (int f : state_fields(c)) {
      in[nf] = dS + lo + (long long)f * np;
      out[nf] = dD + (long long)(3 * f + d) * np;
      ++nf;
    }
    run_solve(c, d, GATT_DERIV, nf, in, out, 0.0, 0, nm, ls, ls);
  }
  ViscLinArgs v{};
  v.c = c->fc;
  v.np = np;
  v.member_stride = c->slab;
  v.work_stride = c->work_stride;
  v.lin_stride = ls;
  v.E = S + (long long)m0 * c->slab + 4 * np;
  v.D = c->work_ptr(WK_D);
  v.dE = dS + lo + 4 * np;
  v.dD = dD;
  v.dT = dT;
  v.is3d = c->is3d;
  launch_visc_tl(v, nm, c->stream);
  const int tin3[3][4] = {{0, 3, 4, 6}, {3, 1, 5, 7}, {4, 5, 2, 8}};
  for (int d = 0; d < ndir; ++d) {
    double* in[kMaxPtr];
    double* out[kMaxPtr];
    int nf = 0;
    for (int q = 0; q < 4; ++q) {
      if (!c->is3d && q == 2) continue;
      in[nf] = dT + (long long)tin3[d][q] * np;
      out[nf] = dG + (long long)(4 * d + q) * np;
      ++nf;
    }
    run_solve(c, d, GATT_DERIV, nf, in, out, 0.0, 0, nm, ls, ls);
  }
  AssembleLinArgs a{};
  a.c = c->fc;
  a.np = np;
  a.member_stride = c->slab;
  a.work_stride = c->work_stride;
  a.lin_stride = ls;
  a.S = S + (long long)m0 * c->slab;
  a.D = c->work_ptr(WK_D);
  a.T = c->work_ptr(WK_T);
  a.G = c->work_ptr(WK_G);
  a.dS = dS + lo;
  a.dD = dD;
  a.dT = dT;
  a.dG = dG;
  a.dQ = dQ + lo;
  a.dOut = dOut + lo;
  a.dDelta = lin_ptr(c, LN_DL) + lo;
  a.stage = stage;
  a.dt = dt;
  a.is3d = c->is3d;
  if (lin_forcing(c)) {
    a.sums = c->d_sums + 8LL * m0;
    a.lsums = c->d_lsums + 8LL * m0;
    a.forcing_eps = c->f_eps;
  }
  launch_assemble_tl(a, nm, c->stream);
}

// Transposed operator of `kind` along `dir` applied to nf fields: out = beta*src
// + (bit f of acc_mask ? out : 0) + alpha * R^T (Thomas^T in) for field f; `in` is
// workspace (the two-launch path runs the Thomas part in place on it, see
// solve_t_one_launch).
constexpr int kAccAll = 0xffff;
void run_solve_t(gatt_ctx* c, int dir, int kind, int nf, double* const* in, double* const* out,
                 double* const* src, double alpha, double beta, int acc_mask, int nm, long long in_stride,
                 long long out_stride) {
  DirOps& d = c->dir[dir];
  BandTArgs b{};
  b.dir = dir;
  b.nx = c->nx;
  b.ny = c->ny;
  b.nz = c->nz;
  b.N = d.op[kind].N;
  b.nR = d.op[kind].nR;
  b.cs = (2 * d.op[kind].nR - 1 + 4 + 1) & ~1;
  b.np = c->np;
  b.coef = d.coef[kind];
  b.R = d.op[kind].R;
  b.fcoef_t = d.fcoef_t[kind];
  b.fcoef_pt = d.fcoef_pt[kind][d.op[kind].N <= 256 ? 0 : 1];
  b.uniform = d.uniform_tiles ? 1 : 0;
  b.fcoef_ends = d.fcoef_ends[kind];
  b.periodic = d.op[kind].periodic;
  b.nlines = d.nlines;
  b.types = d.types;
  b.nf = nf;
  b.in_stride = in_stride;
  b.out_stride = out_stride;
  for (int f = 0; f < nf; ++f) {
    b.in[f] = in[f];
    b.out[f] = out[f];
    b.src[f] = src ? src[f] : nullptr;
  }
  b.alpha = alpha;
  b.beta = beta;
  b.accumulate = acc_mask;
  // along x: one launch, one wavefront per line
  if (dir == 0 && launch_line_solve_xgt(b, nm, c->stream)) return;
  SolveArgs a{};
  a.op = d.op[kind];
  a.types = d.types;
  a.dir = dir;
  a.nx = c->nx;
  a.ny = c->ny;
  a.nz = c->nz;
  a.n1 = d.n1;
  a.nlines = d.nlines;
  a.stride = d.stride;
  a.cs = b.cs;
  a.nf = nf;
  a.in_stride = in_stride;
  a.out_stride = in_stride;
  for (int f = 0; f < nf; ++f) {
    a.in[f] = in[f];
    a.out[f] = in[f];
  }
  a.coef = d.coef[kind];
  a.coef_ends = d.coef_ends[kind];
  a.uniform = d.uniform_tiles ? 1 : 0;
  if (launch_line_solve_band_t(a, b, nm, c->max_lds, c->stream)) return;
  launch_line_solve_t(a, nm, c->stream);
  launch_band_t(b, nm, c->stream);
}

// The transposed solve of `kind` along `dir` for nf fields runs in one launch that reads
// its whole input before writing (so it leaves `in` as it was, and may act in place).
bool solve_t_one_launch(const gatt_ctx* c, int dir, int kind, int nf) {
  const LineOpDev& o = c->dir[dir].op[kind];
  if (dir == 0) return o.N <= 512 && o.nR >= 2 && o.nR <= 11 && o.nR - 1 <= o.N;
  return line_solve_band_t_fits(dir, o.N, o.nR, nf, c->batch, c->max_lds);
}

// Sbar (lin slab, all members) = J(S)^T kbar for the base stage state S
// (member 0 pointer, member stride ss) whose derivatives, stresses and their
// derivatives are in D, T, G (member stride ws).
// kscale: the RHS adjoint is kscale * kbar (kscale = 1 unless the turbulence forcing, whose
// sums read kbar as it is, is off)
void adj_rhs(gatt_ctx* c, const double* S, long long ss, double* kbar, double* sbar, const double* Dw,
             const double* Tw, const double* Gw, long long ws, double kscale = 1.0) {
  const long long np = c->np, ls = c->lin_stride;
  const int ndir = c->is3d ? 3 : 2;
  const int tin3[3][4] = {{0, 3, 4, 6}, {3, 1, 5, 7}, {4, 5, 2, 8}};
  {
    const int m0 = 0, nm = c->batch;
    const long long lo = 0;
    double* Db = lin_ptr(c, LN_D) + lo;
    double* Tb = lin_ptr(c, LN_T) + lo;
    double* Gb = lin_ptr(c, LN_G) + lo;
    double* sb = sbar + lo;   // set (not accumulated) by the point-wise adjoint below
    AssembleLinArgs a{};
    a.c = c->fc;
    a.np = np;
    a.member_stride = ss;
    a.work_stride = ws;
    a.lin_stride = ls;
    a.S = S + (long long)m0 * ss;
    a.D = Dw;
    a.T = Tw;
    a.G = Gw;
    a.dbar = kbar + lo;
    a.dbar_scale = kscale;
    a.Sbar = sb;
    a.Dbar = Db;
    a.Tbar = Tb;
    a.Gbar = Gb;
    a.is3d = c->is3d;
    // the flux-derivative adjoints of the three directions are the same four values
    // (vtx, vty, vtz and -Phi each enter once per direction): one copy serves every
    // direction when no transposed solve works in place on its input
    const int nflux = c->is3d ? 4 : 3;
    bool shared = true;
    for (int d = 0; d < ndir; ++d) shared = shared && solve_t_one_launch(c, d, GATT_DERIV, nflux);
    a.gbar_copies = shared ? 1 : ndir;
    if (lin_forcing(c)) {
      // the forcing's sums of the base stage state and of the RHS adjoint against it
      base_forcing_sums(c, S, ss);
      lin_forcing_sums(c, S, ss, kbar + lo);
      a.sums = c->d_sums;
      a.lsums = c->d_lsums;
      a.forcing_eps = c->f_eps;
    }
    launch_assemble_adj(a, nm, c->stream);
    // Tbar += d^T Gbar, direction by direction; the heat fluxes get their one contribution
    // here (assemble_adj leaves them unset)
    for (int d = 0; d < ndir; ++d) {
      double* in[kMaxPtr];
      double* out[kMaxPtr];
      int nf = 0, acc = 0;
      for (int q = 0; q < 4; ++q) {
        if (!c->is3d && q == 2) continue;
        in[nf] = Gb + (long long)((shared ? 0 : 4 * d) + q) * np;
        out[nf] = Tb + (long long)tin3[d][q] * np;
        if (q != 3) acc |= 1 << nf;
        ++nf;
      }
      run_solve_t(c, d, GATT_DERIV, nf, in, out, nullptr, 1.0, 0.0, acc, nm, ls, ls);
    }
    ViscLinArgs v{};
    v.c = c->fc;
    v.np = np;
    v.member_stride = ss;
    v.work_stride = ws;
    v.lin_stride = ls;
    v.E = S + (long long)m0 * ss + 4 * np;
    v.D = Dw;
    v.dT = Tb;
    v.Dbar = Db;
    v.dEbar = sb + 4 * np;
    v.is3d = c->is3d;
    launch_visc_adj(v, nm, c->stream);
    // Sbar += d^T Dbar
    for (int d = 0; d < ndir; ++d) {
      double* in[kMaxPtr];
      double* out[kMaxPtr];
      int nf = 0;
      for (int f : state_fields(c)) {
        in[nf] = Db + (long long)(3 * f + d) * np;
        out[nf] = sb + (long long)f * np;
        ++nf;
      }
      run_solve_t(c, d, GATT_DERIV, nf, in, out, nullptr, 1.0, 0.0, kAccAll, nm, ls, ls);
    }
  }
}

BcLinArgs bc_lin_args(gatt_ctx* c, const BcWave& w, double* q, long long qs, double* d, long long ds) {
  BcLinArgs a{};
  a.ops = c->d_ops;
  a.op_begin = w.op_begin;
  a.op_end = w.op_end;
  for (int f = 0; f < 5; ++f) {
    a.q[f] = q + (long long)f * c->np;
    a.dq[f] = d + (long long)f * c->np;
  }
  a.member_stride = qs;
  a.lin_stride = ds;
  a.coef[0] = c->d_coef[0];
  a.coef[1] = c->d_coef[1];
  a.aux = c->d_aux;
  a.hold = c->d_hold;
  a.hold_member_stride = c->hold_stride;
  a.nx = c->nx;
  a.ny = c->ny;
  a.nz = c->nz;
  a.gamma = c->gamma;
  return a;
}

// Boundary program on the base q (member stride qs) with its tangent d (stride ds).
void run_bc_tl(gatt_ctx* c, double* q, long long qs, double* d, long long ds) {
  for (const BcWave& w : c->waves)
    launch_bc_wave_tl(bc_lin_args(c, w, q, qs, d, ds), w.nblocks, c->batch, c->stream);
}

// undo log k (0..4) of the base step's boundary applications
double* undo_log(gatt_ctx* c, int k) {
  return c->d_undo ? c->d_undo + (long long)k * c->undo_total * c->batch : nullptr;
}

// lam <- (d BC / d q)^T lam, waves in reverse order.  `post` holds the state after the
// boundary application that wrote undo log `undo`; walking back, each wave's saved
// values are put back first, so the wave's transpose sees the state before it.
// `post` ends as the state before the application.
void run_bc_adj(gatt_ctx* c, double* post, long long ps, double* lam, const double* undo) {
  const long long ls = c->lin_stride;
  for (int wi = (int)c->waves.size() - 1; wi >= 0; --wi) {
    const BcWave& w = c->waves[(size_t)wi];
    BcLinArgs la = bc_lin_args(c, w, post, ps, lam, ls);
    la.undo = undo;
    la.undo_member_stride = c->undo_total;
    launch_bc_wave_adj(la, w.nblocks, c->batch, c->stream);
  }
}

// The SFD blend of main.F90:382-401 on the tangents (transpose = false) or the adjoints
// (transpose = true) of the state and, where the mean evolves (Delta > 0), of the mean.
SfdArgs sfd_lin_args(gatt_ctx* c, double dt, bool transpose) {
  SfdArgs sa{};
  sa.np = c->np;
  sa.member_stride = c->lin_stride;
  sa.Q = lin_ptr(c, LN_Q);
  sa.Mean = c->sfdDelta > 0 ? lin_ptr(c, LN_M) : nullptr;
  sa.chi = c->sfdType == 2 ? c->d_chi : nullptr;
  sa.chi_val = c->sfdXval;
  sa.dt = dt;
  sa.delta = c->sfdDelta > 0 ? c->sfdDelta : 0.0;
  sa.nf_mask = c->is3d ? 0x1f : 0x1b;
  sa.transpose = transpose ? 1 : 0;
  return sa;
}

// One step of the base state with its tangent (main.F90:255-420 and its derivative).
void enqueue_tlm_step(gatt_ctx* c, double t, double dt) {
  (void)t;
  double* q = c->state(c->cur);
  double* qn = c->state(1 - c->cur);
  double* S = c->slab_ptr(SL_S);
  double* dq = lin_ptr(c, LN_Q);
  double* dS = lin_ptr(c, LN_S);
  for (int stage = 1; stage <= 4; ++stage) {
    double* Sin = stage == 1 ? q : S;
    double* Out = stage == 4 ? qn : S;
    double* dSin = stage == 1 ? dq : dS;
    double* dOut = stage == 4 ? dq : dS;
    // the forcing's sums of the base stage state and of the perturbation against it
    base_forcing_sums(c, Sin, c->slab);
    lin_forcing_sums(c, Sin, c->slab, dSin);
    for (int m0 = 0; m0 < c->batch; m0 += c->wgroup) {
      const int nm = std::min(c->wgroup, c->batch - m0);
      base_dtg(c, Sin + (long long)m0 * c->slab, c->slab, nm, c->work_ptr(WK_D), c->work_ptr(WK_T),
               c->work_ptr(WK_G), c->work_stride);
      tl_stage(c, Sin, dSin, dq, dOut, stage, dt, m0, nm);   // before the base stage state is overwritten
      forward_assemble(c, Sin, q, Out, stage, dt, m0, nm, c->work_ptr(WK_D), c->work_ptr(WK_T), c->work_ptr(WK_G),
                       c->work_stride);
    }
    run_bc_tl(c, Out, c->slab, dOut, c->lin_stride);    // timeIntegration.F90: BCs after every stage
  }
  const double theta = dt / c->filterCharTime;
  run_filters(c, qn, theta);
  const int ndir = c->is3d ? 3 : 2;
  for (int d = 0; d < ndir; ++d) {
    if (!c->filterDir[d]) continue;
    double* ptr[kMaxPtr];
    int nf = 0;
    for (int f : state_fields(c)) ptr[nf++] = dq + (long long)f * c->np;
    run_solve(c, d, GATT_FILTER, nf, ptr, ptr, theta, 1, c->batch, c->lin_stride, c->lin_stride);
  }
  if (c->sfdType > 0) {
    run_sfd(c, qn, dt);
    // the blend is linear in (q, mean): the same kernel on the tangents, with the mean's
    // tangent where the mean evolves (Delta > 0), a zero mean tangent where it is frozen
    launch_sfd(sfd_lin_args(c, dt, false), c->batch, c->stream);
  }
  run_bc_tl(c, qn, c->slab, dq, c->lin_stride);
  ReduceArgs r{};
  r.np = c->np;
  r.member_stride = c->slab;
  r.Qn = qn;
  r.Q = q;
  r.out = c->d_red;
  r.is3d = c->is3d;
  launch_step_reduce(r, c->batch, c->stream);
  launch_record(c->d_red, c->d_rec, c->d_flag, c->batch, c->stream);
  c->cur = 1 - c->cur;
}

// Adjoint of one step about the base state q (unchanged): qbar <- J(q)^T qbar.
void enqueue_adj_step(gatt_ctx* c, double t, double dt) {
  (void)t;
  const long long ls = c->lin_stride, sl = c->slab;
  double* q = c->state(c->cur);
  double* S = c->slab_ptr(SL_S);
  double* QB = lin_ptr(c, LN_Q);
  double* SB = lin_ptr(c, LN_S);
  double* KB = lin_ptr(c, LN_DL);
  double* B1 = lin_ptr(c, LN_B1);
  double* B2 = lin_ptr(c, LN_B2);
  double* B3 = lin_ptr(c, LN_B3);
  // stage states of the base step (timeIntegration.F90:382-487); the derivatives,
  // stresses and their derivatives of every stage go to the tape for the adjoint
  const long long ts = c->tape_stride;
  const int nb = c->batch;
  double* qn = c->state(1 - c->cur);     // scratch for the last stage; the base q is not modified
  double* B4 = lin_ptr(c, LN_B4);
  // every boundary application logs what it overwrites, for the transposed program
  if (nb == 1) {
    // one member: the member strides of the state slab and the linearised slab do not
    // matter, so the stage states go straight into B1..B4 (no copies), and the filters
    // read B4 and write the step's result into qn
    base_dtg(c, q, sl, nb, tape_ptr(c, 0, WK_D), tape_ptr(c, 0, WK_T), tape_ptr(c, 0, WK_G), ts);
    base_forcing_sums(c, q, sl);
    forward_assemble(c, q, q, B1, 1, dt, 0, nb, tape_ptr(c, 0, WK_D), tape_ptr(c, 0, WK_T), tape_ptr(c, 0, WK_G), ts);
    run_bc(c, B1, dt / 2.0, undo_log(c, 0));
    base_dtg(c, B1, ls, nb, tape_ptr(c, 1, WK_D), tape_ptr(c, 1, WK_T), tape_ptr(c, 1, WK_G), ts);
    base_forcing_sums(c, B1, ls);
    forward_assemble(c, B1, q, B2, 2, dt, 0, nb, tape_ptr(c, 1, WK_D), tape_ptr(c, 1, WK_T), tape_ptr(c, 1, WK_G), ts);
    run_bc(c, B2, dt / 2.0, undo_log(c, 1));
    base_dtg(c, B2, ls, nb, tape_ptr(c, 2, WK_D), tape_ptr(c, 2, WK_T), tape_ptr(c, 2, WK_G), ts);
    base_forcing_sums(c, B2, ls);
    forward_assemble(c, B2, q, B3, 3, dt, 0, nb, tape_ptr(c, 2, WK_D), tape_ptr(c, 2, WK_T), tape_ptr(c, 2, WK_G), ts);
    run_bc(c, B3, dt, undo_log(c, 2));
    base_dtg(c, B3, ls, nb, tape_ptr(c, 3, WK_D), tape_ptr(c, 3, WK_T), tape_ptr(c, 3, WK_G), ts);
    base_forcing_sums(c, B3, ls);
    forward_assemble(c, B3, q, B4, 4, dt, 0, nb, tape_ptr(c, 3, WK_D), tape_ptr(c, 3, WK_T), tape_ptr(c, 3, WK_G), ts);
    run_bc(c, B4, dt, undo_log(c, 3));
    run_filters_from(c, B4, qn, dt / c->filterCharTime);
  } else {
    base_dtg(c, q, sl, nb, tape_ptr(c, 0, WK_D), tape_ptr(c, 0, WK_T), tape_ptr(c, 0, WK_G), ts);
    base_forcing_sums(c, q, sl);
    forward_assemble(c, q, q, S, 1, dt, 0, nb, tape_ptr(c, 0, WK_D), tape_ptr(c, 0, WK_T), tape_ptr(c, 0, WK_G), ts);
    run_bc(c, S, dt / 2.0, undo_log(c, 0));
    fields_axpby(c, B1, ls, S, sl, 1.0, 0.0);
    base_dtg(c, S, sl, nb, tape_ptr(c, 1, WK_D), tape_ptr(c, 1, WK_T), tape_ptr(c, 1, WK_G), ts);
    base_forcing_sums(c, S, sl);
    forward_assemble(c, S, q, S, 2, dt, 0, nb, tape_ptr(c, 1, WK_D), tape_ptr(c, 1, WK_T), tape_ptr(c, 1, WK_G), ts);
    run_bc(c, S, dt / 2.0, undo_log(c, 1));
    fields_axpby(c, B2, ls, S, sl, 1.0, 0.0);
    base_dtg(c, S, sl, nb, tape_ptr(c, 2, WK_D), tape_ptr(c, 2, WK_T), tape_ptr(c, 2, WK_G), ts);
    base_forcing_sums(c, S, sl);
    forward_assemble(c, S, q, S, 3, dt, 0, nb, tape_ptr(c, 2, WK_D), tape_ptr(c, 2, WK_T), tape_ptr(c, 2, WK_G), ts);
    run_bc(c, S, dt, undo_log(c, 2));
    fields_axpby(c, B3, ls, S, sl, 1.0, 0.0);
    base_dtg(c, S, sl, nb, tape_ptr(c, 3, WK_D), tape_ptr(c, 3, WK_T), tape_ptr(c, 3, WK_G), ts);
    base_forcing_sums(c, S, sl);
    forward_assemble(c, S, q, qn, 4, dt, 0, nb, tape_ptr(c, 3, WK_D), tape_ptr(c, 3, WK_T), tape_ptr(c, 3, WK_G), ts);
    run_bc(c, qn, dt, undo_log(c, 3));
    fields_axpby(c, B4, ls, qn, sl, 1.0, 0.0);
    run_filters(c, qn, dt / c->filterCharTime);
  }
  run_sfd(c, qn, dt, true);   // the base mean stays as given
  run_bc(c, qn, dt, undo_log(c, 4));
  // the final boundary conditions (main.F90:405), walked back on qn
  run_bc_adj(c, qn, sl, QB, undo_log(c, 4));
  // the transposed blend: diagonal with a frozen mean, coupled to the mean's adjoint where
  // the mean evolves
  if (c->sfdType > 0) launch_sfd(sfd_lin_args(c, dt, true), c->batch, c->stream);
  // filters in reverse order: qbar <- (1-theta) qbar + theta F^T qbar (theta <= 0: F^T qbar)
  const double theta = dt / c->filterCharTime;
  const int ndir = c->is3d ? 3 : 2;
  for (int d = ndir - 1; d >= 0; --d) {
    if (!c->filterDir[d]) continue;
    // the one-launch transposed solves read their whole input before writing, so the
    // filter can act in place; the two-launch path gets a copy
    const bool in_place = solve_t_one_launch(c, d, GATT_FILTER, (int)state_fields(c).size());
    if (!in_place) fields_axpby(c, KB, ls, QB, ls, 1.0, 0.0);
    double* in[kMaxPtr];
    double* out[kMaxPtr];
    int nf = 0;
    for (int f : state_fields(c)) {
      in[nf] = (in_place ? QB : KB) + (long long)f * c->np;
      out[nf] = QB + (long long)f * c->np;
      ++nf;
    }
    const bool blend = theta > 0.0;
    run_solve_t(c, d, GATT_FILTER, nf, in, out, blend ? out : nullptr, blend ? theta : 1.0, blend ? 1.0 - theta : 0.0,
                0, c->batch, ls, ls);
  }
  // boundary conditions after the last stage
  run_bc_adj(c, B4, ls, QB, undo_log(c, 3));
  // RK4 (the accumulator form of timeIntegration.F90:382-496), backwards, with the
  // transposed boundary conditions after every stage
  // the last stage's RHS adjoint is dt/6 QB: scaled as the point-wise adjoint reads it, unless
  // the forcing's sums need it stored
  if (lin_forcing(c)) {
    fields_axpby(c, KB, ls, QB, ls, dt / 6.0, 0.0);
    adj_rhs(c, B3, ls, KB, SB, tape_ptr(c, 3, WK_D), tape_ptr(c, 3, WK_T), tape_ptr(c, 3, WK_G), ts);
  } else {
    adj_rhs(c, B3, ls, QB, SB, tape_ptr(c, 3, WK_D), tape_ptr(c, 3, WK_T), tape_ptr(c, 3, WK_G), ts, dt / 6.0);
  }
  run_bc_adj(c, B3, ls, SB, undo_log(c, 2));   // B3 is not read again as a state
  // B3 now accumulates the stage adjoints; KB = dt/3 QB + dt SB
  adj_combine(c, QB, SB, B3, KB, dt / 3.0, dt, 1, 0);
  adj_rhs(c, B2, ls, KB, SB, tape_ptr(c, 2, WK_D), tape_ptr(c, 2, WK_T), tape_ptr(c, 2, WK_G), ts);
  run_bc_adj(c, B2, ls, SB, undo_log(c, 1));   // B2 is not read again as a state
  adj_combine(c, QB, SB, B3, KB, dt / 3.0, dt / 2.0, 0, 0);
  adj_rhs(c, B1, ls, KB, SB, tape_ptr(c, 1, WK_D), tape_ptr(c, 1, WK_T), tape_ptr(c, 1, WK_G), ts);
  run_bc_adj(c, B1, ls, SB, undo_log(c, 0));   // B1 is not read again as a state
  adj_combine(c, QB, SB, B3, KB, dt / 6.0, dt / 2.0, 0, 0);
  adj_rhs(c, q, sl, KB, SB, tape_ptr(c, 0, WK_D), tape_ptr(c, 0, WK_T), tape_ptr(c, 0, WK_G), ts);
  // QB = SB + (B3 + QB)
  adj_combine(c, QB, SB, B3, nullptr, 0.0, 0.0, 0, 1);
}

int copy_lin_fields(gatt_ctx* c, int32_t member, const double* const* hq, bool to_dev) {
  if (member < 0 || member >= c->batch) return fail(GATT_E_INVALID, "member out of range");
  int rc = ensure_lin(c);
  if (rc) return rc;
  double* base = lin_ptr(c, LN_Q) + (long long)member * c->lin_stride;
  for (int f = 0; f < 5; ++f) {
    if (!c->is3d && f == GATT_W) continue;
    if (!hq[f]) return fail(GATT_E_INVALID, "null field pointer");
    double* dptr = base + (long long)f * c->np;
    if (to_dev) HIPCHK(hipMemcpyAsync(dptr, hq[f], sizeof(double) * c->np, hipMemcpyHostToDevice, c->stream));
    else HIPCHK(hipMemcpyAsync((void*)hq[f], dptr, sizeof(double) * c->np, hipMemcpyDeviceToHost, c->stream));
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  return GATT_OK;
}

}  // namespace

extern "C" {

int gatt_set_perturbation(gatt_ctx* c, int32_t member, const double* const* q) {
  if (check_ctx(c) || !q) return fail(GATT_E_INVALID, "null argument");
  return copy_lin_fields(c, member, q, true);
}

int gatt_get_perturbation(gatt_ctx* c, int32_t member, double* const* q) {
  if (check_ctx(c) || !q) return fail(GATT_E_INVALID, "null argument");
  int rc = copy_lin_fields(c, member, (const double* const*)q, false);
  if (rc) return rc;
  if (!c->is3d && q[GATT_W]) std::memset(q[GATT_W], 0, sizeof(double) * c->np);
  return GATT_OK;
}

int gatt_copy_fields(gatt_ctx* c, int32_t member, int32_t slot, int32_t to_ctx, double* dev) {
  if (check_ctx(c) || !dev) return fail(GATT_E_INVALID, "null argument");
  if (member < 0 || member >= c->batch) return fail(GATT_E_INVALID, "member out of range");
  double* base;
  if (slot == GATT_SLOT_STATE) {
    base = c->state(c->cur) + (long long)member * c->slab;
  } else if (slot == GATT_SLOT_PERTURBATION || slot == GATT_SLOT_MEAN_PERTURBATION) {
    int rc = ensure_lin(c);
    if (rc) return rc;
    base = lin_ptr(c, slot == GATT_SLOT_PERTURBATION ? LN_Q : LN_M) + (long long)member * c->lin_stride;
  } else if (slot == GATT_SLOT_MEAN) {
    if (c->sfdType <= 0) return fail(GATT_E_INVALID, "the problem has no SFD mean");
    // before the mean is set or first used it is the state (every member's, as at the first step)
    if (to_ctx) ensure_mean(c);
    base = (c->mean_ready ? c->slab_ptr(SL_MEAN) : c->state(c->cur)) + (long long)member * c->slab;
  } else {
    return fail(GATT_E_INVALID, "unknown field slot");
  }
  const size_t bytes = sizeof(double) * 5 * (size_t)c->np;
  if (to_ctx) HIPCHK(hipMemcpyAsync(base, dev, bytes, hipMemcpyDeviceToDevice, c->stream));
  else HIPCHK(hipMemcpyAsync(dev, base, bytes, hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return GATT_OK;
}

int gatt_line_solve_t(gatt_ctx* c, int32_t dir, int32_t kind, int32_t nfields, double* const* in,
                      double* const* out) {
  if (check_ctx(c)) return GATT_E_INVALID;
  if (dir < 0 || dir > 2 || (!c->is3d && dir == 2)) return fail(GATT_E_INVALID, "direction out of range");
  if (kind != GATT_DERIV && kind != GATT_FILTER) return fail(GATT_E_INVALID, "unknown operator kind");
  if (nfields < 1 || nfields > kMaxPtr || !in || !out) return fail(GATT_E_INVALID, "bad field list");
  if (line_solve_t_points_per_lane(c->dir[dir].op[kind].N) == 0)
    return fail(GATT_E_UNSUPPORTED, "transposed line solves are compiled for lines of at most 512 points");
  for (int f = 0; f < nfields; ++f)
    if (!in[f] || !out[f] || in[f] == out[f]) return fail(GATT_E_INVALID, "null or aliasing field pointer");
  run_solve_t(c, dir, kind, nfields, in, out, nullptr, 1.0, 0.0, 0, 1, 0, 0);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  return GATT_OK;
}

int gatt_tlm_step(gatt_ctx* c, double t, double dt, double* maxchange, int32_t* diverged) {
  if (check_ctx(c)) return GATT_E_INVALID;
  int rc = lin_supported(c);
  if (rc) return rc;
  if (c->d_hold && !c->hold_captured)
    return fail(GATT_E_INVALID, "apply the boundary conditions once first (holdInlet captures the inflow)");
  if ((rc = ensure_lin(c))) return rc;
  ensure_mean(c);
  launch_clear_i32(c->d_flag, 1, c->stream);
  enqueue_tlm_step(c, t, dt);
  HIPCHK(hipGetLastError());
  return read_scalars(c, nullptr, maxchange, diverged);
}

int gatt_adj_step(gatt_ctx* c, double t, double dt) {
  if (check_ctx(c)) return GATT_E_INVALID;
  int rc = lin_supported(c);
  if (rc) return rc;
  if (c->d_hold && !c->hold_captured)
    return fail(GATT_E_INVALID, "apply the boundary conditions once first (holdInlet captures the inflow)");
  ensure_mean(c);
  if ((rc = ensure_lin(c))) return rc;
  if ((rc = ensure_tape(c))) return rc;
  // the step's few hundred launches (most of them boundary waves) are captured once
  // per base buffer, time and step size and replayed
  if (c->adj_graph && (c->adj_graph_t != t || c->adj_graph_dt != dt || c->adj_graph_cur != c->cur)) {
    (void)hipGraphExecDestroy(c->adj_graph);
    c->adj_graph = nullptr;
  }
  if (!c->adj_graph) {
    hipGraph_t g;
    HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    enqueue_adj_step(c, t, dt);
    HIPCHK(hipStreamEndCapture(c->stream, &g));
    const hipError_t e = hipGraphInstantiate(&c->adj_graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
      c->adj_graph = nullptr;
      return fail(GATT_E_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    }
    c->adj_graph_t = t;
    c->adj_graph_dt = dt;
    c->adj_graph_cur = c->cur;
  }
  HIPCHK(hipGraphLaunch(c->adj_graph, c->stream));
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  return GATT_OK;
}

}  // extern "C"
