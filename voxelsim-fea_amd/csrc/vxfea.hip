// vxfea.hip -- host side of the MI355X hex8 MG-PCG solver and its C ABI.
//
// One handle = one voxel grid with its multigrid hierarchy (the reference's
// MGSolver, core/multigrid.py:76-583) and the workspace of one solve.  The
// PCG scalars live in device memory; the host reads back one small struct
// per iteration (for the convergence test, the stall guard and the progress
// callback) while the next V-cycle is already queued.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstddef>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "kernels.h"
#include "net_internal.h"
#include "vxfea.h"

using namespace vx;

namespace {

constexpr int EW_BLOCKS = 1024;   // grid of the elementwise / reduction kernels
constexpr int EW_THREADS = 256;

struct Level {
  int nx, ny, nz, Lx, Ly, Lz;
  long long Nn, Ne;
  uint8_t* mask = nullptr;
  double* E64 = nullptr;
  float* E32 = nullptr;
  // V-cycle work vectors (levels >= 1 own b; x/t at every level)
  double *b64 = nullptr, *x64 = nullptr, *t64 = nullptr, *w64 = nullptr;
  float *b32 = nullptr, *x32 = nullptr, *t32 = nullptr, *w32 = nullptr;
  float *v32a = nullptr, *v32b = nullptr;   // blocked-smoother iterates (small levels)
  // blocked smoother plan from vx_autotune: box edge and steps per launch,
  // blk_B = -1: one tile launch per step; 0: not tuned (the size rule)
  int blk_B = 0, blk_S = 0;
  uint8_t* mask_fixed = nullptr;   // injected fixed-DOF mask (set_fixed)
  double* mi64 = nullptr;          // inverse Jacobi diagonal per node
  float* mi32 = nullptr;
  // autotuned node planes per work-group: [precision 0 FP64 / 1 FP32][class
  // 0 operator (matvec forms), 1 smoother / residual forms, 2 smoother forms
  // with the dot-product partials]; 0 = rule
  int zc_tuned[2][5] = {{0, 0, 0, 0, 0}, {0, 0, 0, 0, 0}};
};

template <typename T> struct Sel;
template <> struct Sel<double> {
  static double*& b(Level& l) { return l.b64; }
  static double*& x(Level& l) { return l.x64; }
  static double*& t(Level& l) { return l.t64; }
  static double*& w(Level& l) { return l.w64; }
  static double* E(Level& l) { return l.E64; }
  static double* mi(Level& l) { return l.mi64; }
};
template <> struct Sel<float> {
  static float*& b(Level& l) { return l.b32; }
  static float*& x(Level& l) { return l.x32; }
  static float*& t(Level& l) { return l.t32; }
  static float*& w(Level& l) { return l.w32; }
  static float* E(Level& l) { return l.E32; }
  static float* mi(Level& l) { return l.mi32; }
};

}  // namespace

struct vx_solver {
  int nx, ny, nz;
  double nu;
  vx_options opt;
  std::vector<Level> lv;
  hipStream_t stream = nullptr;
  std::string err;
  ElemConst<double> K64;
  ElemConst<float> K32;
  double kd = 0;
  double lam_max = 0;        // max eigenvalue of D_e^-1 KE: bounds lambda_max(D^-1 A)
  bool fixed_set = false, density_set = false;
  bool active_system = false;
  // level-0 PCG vectors
  double *u64 = nullptr, *r64 = nullptr, *z64 = nullptr, *p64 = nullptr,
         *Ap64 = nullptr, *fm64 = nullptr;
  float *r32 = nullptr, *z32 = nullptr, *p32 = nullptr, *Ap32 = nullptr,
        *d32 = nullptr;
  Level jac;                  // active-DOF system of the Jacobi-PCG fallback
  double* jac_minv = nullptr; // per node
  double* partial = nullptr;
  int n_partial = 0;
  int nb = 1;                 // right-hand sides of the current operation
  Scalars* dscal = nullptr;   // [max_batch]
  Scalars* hscal = nullptr;   // pinned mirror
  int* dflags = nullptr;      // [max_batch] frozen flags staging
  int* hflags = nullptr;      // pinned
  // learned coarse-grid corrector (vx_set_corrector)
  vx_net* corr = nullptr;
  int corr_level = -1;
  float corr_gain = 1.f;
  int corr_budget = 0;        // inner iterations allowed per learned solve (0: no limit)
  bool corr_enabled = false;  // attached and used by solves
  bool corr_on = false;       // active inside the current solve (the fallback clears it)
  float* corr_feat = nullptr; // [Nn][4] of the corrector level
  float* corr_out = nullptr;  // [max_batch][Nn][16]
  float *corr_v = nullptr, *corr_y = nullptr, *corr_Av = nullptr, *corr_Ay = nullptr;
                              // [max_batch][3 Nn] each
  double* corr_part = nullptr;  // [max_batch][5][EW_BLOCKS]
  double* corr_coef = nullptr;  // [max_batch][2]
  float* corr_rms = nullptr;  // [max_batch]
  unsigned int* corr_emax = nullptr;
  long long corr_cap_nn = 0;  // nodes the corrector buffers were sized for
  // coarsest-level direct solve (coarse_mode VX_COARSE_DIRECT): the inverse
  // of the coarsest operator on its free DOFs, FP64, applied in the V-cycle
  bool cd_ok = false;
  int cd_nf = 0;
  double* cd_inv = nullptr;   // [nf][nf], symmetric
  int* cd_pos = nullptr;      // per coarsest DOF: its row in the free list, or -1
  int* cd_idx = nullptr;      // free DOFs of the coarsest level
  long long cd_cap_inv = 0, cd_cap_dof = 0;
  // capture of corrector-level residuals during solves (vx_capture)
  float* cap_buf = nullptr;
  int cap_level = -1, cap_cap = 0, cap_count = 0, cap_every = 1;
  long long cap_calls = 0;
  // V-cycles captured as HIP graphs (one per key; cleared whenever the
  // launches they hold could change) and the stream they are captured on,
  // which never runs work
  struct CycleGraph {
    const void* ptr[4];
    int type, l0, nb, first_done, flexible;
    hipGraphExec_t exec;
  };
  std::vector<CycleGraph> graphs;
  hipStream_t cap_stream = nullptr;
  bool graphs_failed = false;
  int n_cu = 256;             // compute units of the device
  std::vector<void*> allocs;
};

namespace {

int fail(vx_solver* s, int code, const std::string& msg) {
  if (s) s->err = msg;
  return code;
}

// the captured V-cycles hold launch configurations (z-chunks, the coarse
// mode) and buffers that set-up calls may change
void drop_graphs(vx_solver* s) {
  for (auto& g : s->graphs) (void)hipGraphExecDestroy(g.exec);
  s->graphs.clear();
}

#define HIPCHK(call)                                                                   \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess)                                                              \
      return fail(s, VX_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define LAUNCHCHK()                                                                           \
  do {                                                                                        \
    hipError_t e_ = hipGetLastError();                                                        \
    if (e_ != hipSuccess)                                                                     \
      return fail(s, VX_ERR_HIP, std::string("kernel launch: ") + hipGetErrorString(e_)); \
  } while (0)

#define CHK(x)             \
  do {                     \
    int rc_ = (x);         \
    if (rc_ != VX_OK) return rc_; \
  } while (0)

template <typename T>
int dalloc(vx_solver* s, T** p, size_t n) {
  if (*p) return VX_OK;
  void* q = nullptr;
  hipError_t e = hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T));
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return fail(s, VX_ERR_NOMEM, "hipMalloc of " + std::to_string(n * sizeof(T)) + " bytes failed");
  }
  s->allocs.push_back(q);
  *p = static_cast<T*>(q);
  return VX_OK;
}

int ew_grid(long long n) {
  long long g = (n + EW_THREADS - 1) / EW_THREADS;
  return (int)std::max<long long>(1, std::min<long long>(g, EW_BLOCKS));
}

int zchunk(const vx_solver* s, const Level& L, int prec = -1, int cls = 0) {
  if (s->opt.z_chunk > 0) return s->opt.z_chunk;
  if (prec >= 0 && L.zc_tuned[prec][cls] > 0) return L.zc_tuned[prec][cls];
  // enough work-groups to fill 256 CUs several times, long enough columns
  // to amortise the halo layer.
  const long long cols = (long long)((L.Lx + OWN - 1) / OWN) * ((L.Ly + OWN - 1) / OWN);
  int zc = L.Lz;
  while (zc > 16 && cols * ((L.Lz + zc - 1) / zc) < 2048) zc = (zc + 1) / 2;
  return std::max(zc, 1);
}

dim3 tile_grid(const vx_solver* s, const Level& L, int zc) {
  return dim3((L.Lx + OWN - 1) / OWN, (L.Ly + OWN - 1) / OWN, (L.Lz + zc - 1) / zc);
}

template <typename T> ElemConst<T> kconst(const vx_solver* s);
template <> ElemConst<double> kconst<double>(const vx_solver* s) { return s->K64; }
template <> ElemConst<float> kconst<float>(const vx_solver* s) { return s->K32; }

struct SmoothCoef {
  double omega = 0;
  const void* up = nullptr;
  double cu = 0, cr = 0;
};

// Every launcher below covers the s->nb right-hand sides of the current
// operation (vectors [nb][3 Nn] of the level, contiguous).

template <typename T, int MODE, bool DOT>
int launch_tile_L(vx_solver* s, Level& L, const T* u, const T* b, T* out, double* partial,
                  int* n_part, const SmoothCoef* sc = nullptr) {
  TileArgs<T> a;
  a.nx = L.nx; a.ny = L.ny; a.nz = L.nz;
  a.Lx = L.Lx; a.Ly = L.Ly; a.Lz = L.Lz;
  a.zc = zchunk(s, L, std::is_same<T, float>::value ? 1 : 0,
                MODE == M_MATVEC ? 0 : (DOT ? 2 : 1));
  a.nb = s->nb;
  a.vstride = 3 * L.Nn;
  a.u = u; a.E = Sel<T>::E(L); a.mask = L.mask; a.b = b; a.out = out;
  a.omega = (T)(sc ? sc->omega : s->opt.omega);
  a.up = sc ? static_cast<const T*>(sc->up) : nullptr;
  a.cu = (T)(sc ? sc->cu : 0.0);
  a.cr = (T)(sc ? sc->cr : 0.0);
  a.kd = (T)s->kd;
  a.K = kconst<T>(s);
  a.partial = partial;
  dim3 g = tile_grid(s, L, a.zc);
  a.gx = (int)g.x; a.gy = (int)g.y; a.gz = (int)g.z;
  const long long G = (long long)g.x * g.y * g.z;
  if (n_part) *n_part = (int)G;
  if (DOT && G * s->nb > s->n_partial) return fail(s, VX_ERR_ARG, "partial buffer too small");
  const long long launch = ((G * s->nb + 7) / 8) * 8;   // XCD-ordered 1D grid
  hipLaunchKernelGGL((k_tile<T, MODE, DOT>), dim3((unsigned)launch), dim3(NT), 0, s->stream,
                     a);
  LAUNCHCHK();
  return VX_OK;
}

template <typename T, int MODE, bool DOT>
int launch_tile(vx_solver* s, int li, const T* u, const T* b, T* out, double* partial,
                int* n_part, const SmoothCoef* sc = nullptr) {
  return launch_tile_L<T, MODE, DOT>(s, s->lv[li], u, b, out, partial, n_part, sc);
}

// per right-hand side: sum its n partials and apply op to its scalars
int finalize(vx_solver* s, int n, int op) {
  hipLaunchKernelGGL(k_finalize, dim3(s->nb), dim3(FIN_THREADS), 0, s->stream, s->partial, n, op,
                     s->dscal);
  LAUNCHCHK();
  return VX_OK;
}

// partials of a.b per right-hand side over vectors of length n each
template <typename T>
int dot_partials(vx_solver* s, long long n, const T* a, const T* b, int* n_part) {
  hipLaunchKernelGGL(k_dot<T>, dim3(EW_BLOCKS, s->nb), dim3(EW_THREADS), 0, s->stream, n, a, b,
                     s->partial);
  LAUNCHCHK();
  if (n_part) *n_part = EW_BLOCKS;
  return VX_OK;
}

// launch shape of the z-marching transfers: cols threads per plane, planes
// split into chunks so that about TR_TARGET_GROUPS work-groups run
constexpr int PROLONG_TARGET_GROUPS = 8192;
TrGrid tr_grid(long long cols, int planes, bool even_chunks, int target = TR_TARGET_GROUPS);
// the XCD-ordered 1-D grid covers every work-group index only when its size
// is a multiple of 8 (tr_block maps block b to (b mod 8) * per + b / 8)
unsigned tr_blocks(const TrGrid& g, int nb) {
  const long long total = (long long)g.gxy * g.nchunk * nb;
  return (unsigned)(((total + 7) / 8) * 8);
}
TrGrid tr_grid(long long cols, int planes, bool even_chunks, int target) {
  TrGrid g;
  g.gxy = (int)((cols + TR_THREADS - 1) / TR_THREADS);
  int nchunk = (target + g.gxy - 1) / g.gxy;
  nchunk = std::max(1, std::min(nchunk, planes));
  g.per_chunk = (planes + nchunk - 1) / nchunk;
  if (even_chunks) g.per_chunk += g.per_chunk & 1;
  g.nchunk = (planes + g.per_chunk - 1) / g.per_chunk;
  return g;
}

// rc = coarse mask * (R rf) into level C (= level li+1) for nb right-hand sides
template <typename T>
int restrict_to(vx_solver* s, Level& C, const T* rf, T* rc, int nb, T* x1 = nullptr,
                double w0 = 0.0) {
  const TrGrid g = tr_grid((long long)C.Lx * C.Ly, C.Lz, false);
  hipLaunchKernelGGL(k_restrict_z<T>, dim3(tr_blocks(g, nb)), dim3(TR_THREADS), 0, s->stream,
                     C.Lx, C.Ly, C.Lz, g, nb, rf, C.mask, rc, (const T*)Sel<T>::mi(C), (T)w0, x1);
  LAUNCHCHK();
  return VX_OK;
}

// The residual b - K x of level li restricted straight into level li+1
// (M_RESR): rc = the coarse right-hand side, x1 (if set) the coarse level's
// first smoothing step w0 D^-1 rc.  FP32.  The fine residual never goes to
// memory.
int launch_resr(vx_solver* s, int li, const float* x, const float* b, float* rc, float* x1,
                double w0) {
  Level& L = s->lv[li];
  const Level& C = s->lv[li + 1];
  TileArgs<float> a;
  a.nx = L.nx; a.ny = L.ny; a.nz = L.nz;
  a.Lx = L.Lx; a.Ly = L.Ly; a.Lz = L.Lz;
  a.zc = std::max(1, (zchunk(s, L, 1, 4) + 1) / 2);   // coarse planes per work-group
  a.nb = s->nb;
  a.vstride = 3 * L.Nn;
  a.u = x; a.E = L.E32; a.mask = L.mask; a.b = b; a.out = nullptr;
  a.omega = 0.f;
  a.up = nullptr;
  a.kd = (float)s->kd;
  a.K = kconst<float>(s);
  a.partial = nullptr;
  a.Lcx = C.Lx; a.Lcy = C.Ly; a.Lcz = C.Lz;
  a.cvstride = 3 * C.Nn;
  a.rc = rc;
  a.cx1 = x1;
  a.cmask = C.mask;
  a.cminv = C.mi32;
  a.w0 = (float)w0;
  constexpr int RCOL = 2 * (OWN / 2) / 2;
  a.gx = (C.Lx + RCOL - 1) / RCOL;
  a.gy = (C.Ly + RCOL - 1) / RCOL;
  a.gz = (C.Lz + a.zc - 1) / a.zc;
  const long long G = (long long)a.gx * a.gy * a.gz;
  const long long launch = ((G * s->nb + 7) / 8) * 8;   // XCD-ordered 1D grid
  hipLaunchKernelGGL((k_tile<float, M_RESR, false>), dim3((unsigned)launch), dim3(NT), 0,
                     s->stream, a);
  LAUNCHCHK();
  return VX_OK;
}

// x += mask * (P e) from level li+1 into level li for nb right-hand sides
// (default: the s->nb of the current operation)
template <typename T>
int prolong_add(vx_solver* s, int li, const T* ec, T* x, int nb = 0) {
  Level& L = s->lv[li];
  const Level& C = s->lv[li + 1];
  if (nb <= 0) nb = s->nb;
  // one thread per fine value of a plane; fine planes in chunks of an even
  // count (the kernel starts on coarse-aligned planes), enough chunks for
  // several rounds of waves per CU
  const TrGrid g = tr_grid(3LL * L.Lx * L.Ly, L.Lz, true, PROLONG_TARGET_GROUPS);
  hipLaunchKernelGGL(k_prolong_f<T>, dim3(tr_blocks(g, nb)), dim3(TR_THREADS), 0, s->stream,
                     C.Lx, C.Ly, C.Lz, g, nb, ec, L.mask, x);
  LAUNCHCHK();
  return VX_OK;
}

template <typename T>
int copy_vec(vx_solver* s, T* dst, const T* src, long long n) {
  HIPCHK(hipMemcpyAsync(dst, src, sizeof(T) * n, hipMemcpyDeviceToDevice, s->stream));
  return VX_OK;
}

// Where a smoothing pass runs.  LevelTarget: level li of the handle's grid,
// all s->nb right-hand sides.  A target supplies the solver (options and
// work buffers), the tile launch, the first step from zero and the number of
// values of one vector set (the parts group defines its own target).
struct LevelTarget {
  vx_solver* s;
  int li;
};

vx_solver* tsol(const LevelTarget& t) { return t.s; }

// offset of the target's first value in its vectors
long long tfirst(const LevelTarget&) { return 0; }

long long tvals(const LevelTarget& t) { return 3 * t.s->lv[t.li].Nn * t.s->nb; }

template <typename T, int MODE, bool DOT>
int ttile(const LevelTarget& t, const T* u, const T* b, T* out, int* n_part, const SmoothCoef* sc) {
  return launch_tile<T, MODE, DOT>(t.s, t.li, u, b, out, DOT ? t.s->partial : nullptr, n_part,
                                   sc);
}

// the first smoothing step from zero, out = omega mask D^-1 b (A 0 = 0); dot:
// partials of out.b
template <typename T>
int tzero(const LevelTarget& t, const T* b, T* out, double omega, bool dot, int* n_part) {
  vx_solver* s = t.s;
  Level& L = s->lv[t.li];
  // float4 path: every right-hand side's vector and the per-node arrays
  // start 16-byte aligned (allocation alignment; rhs stride 3 Nn floats)
  const bool aligned =
      ((reinterpret_cast<uintptr_t>(b) | reinterpret_cast<uintptr_t>(out) |
        reinterpret_cast<uintptr_t>(Sel<T>::mi(L)) | reinterpret_cast<uintptr_t>(L.mask)) & 15) ==
          0 &&
      (s->nb == 1 || (3 * L.Nn) % 4 == 0);
  hipLaunchKernelGGL(k_zero_step<T>, dim3(ew_grid(L.Nn), s->nb), dim3(EW_THREADS), 0, s->stream,
                     L.Nn, L.mask, Sel<T>::mi(L), b, out, (T)omega, aligned ? 1 : 0);
  LAUNCHCHK();
  if (dot) CHK(dot_partials<T>(s, 3 * L.Nn, out, b, n_part));
  return VX_OK;
}

// one damped-Jacobi sweep with weight omega; u == nullptr means "from zero"
template <typename T, typename Tgt>
int tsweep(const Tgt& t, const T* u, const T* b, T* out, bool dot, int* n_part, double omega) {
  if (!u) return tzero<T>(t, b, out, omega, dot, n_part);
  SmoothCoef sc;
  sc.omega = omega;
  if (dot) return ttile<T, M_JACOBI, true>(t, u, b, out, n_part, &sc);
  return ttile<T, M_JACOBI, false>(t, u, b, out, nullptr, &sc);
}

// weight of step i (0, 1) of a two-step Chebyshev pass over [theta - delta,
// theta + delta] written as two damped-Jacobi steps: 1 / (theta + delta /
// sqrt 2), then 1 / (theta - delta / sqrt 2)
static double rich2_weight(double theta, double delta, int i) {
  const double r = delta / std::sqrt(2.0);
  return 1.0 / (i == 0 ? theta + r : theta - r);
}

// the FP32 smoother runs a two-step pass from a nonzero start as two
// damped-Jacobi steps (rich2_weight): its last step is M_JACOBI
static bool post_pass_jacobi(const vx_solver* s, int k) {
  return s->opt.smoother == VX_SMOOTH_CHEBYSHEV && k == 2;
}

// lower end of the Chebyshev interval (over lambda_max) of a k-step pass
static double cheb_ratio_for(const vx_solver* s, int k) {
  const int p = s->opt.cheb_ratio_power;
  if (p <= 0 || k <= 0) return s->opt.cheb_ratio;
  return s->opt.cheb_ratio * std::pow((double)s->opt.n_smooth / (double)k, (double)p);
}

// Chebyshev smoothing (degree k) on D^-1 A over [ratio*lmax, lmax]:
// x1 = x0 + D^-1 r0 / theta, x_{j+1} = x_j + rho_j rho_{j-1} (x_j - x_{j-1})
// + (2 rho_j / delta) D^-1 r_j.  u_in == nullptr starts from zero.  The
// iterates alternate between tmp and tmp2 (a step may overwrite the iterate
// before the one it reads) and the k-th lands in dst; the first is in tmp
// for every k >= 2.  FP32 with k = 2 from u_in: the same polynomial as two
// damped-Jacobi steps (rich2_weight), the first in tmp.
template <typename T, typename Tgt>
int tcheb_n(const Tgt& tg, const T* u_in, const T* b, int k, T* dst, T* tmp, T* tmp2, bool dot,
            int* n_part, bool first_done = false) {
  vx_solver* s = tsol(tg);
  const double beta = s->lam_max, alpha = cheb_ratio_for(s, k) * beta;
  const double theta = 0.5 * (beta + alpha), delta = 0.5 * (beta - alpha);
  const double sigma = theta / delta;
  if (std::is_same<T, float>::value && k == 2 && u_in && !first_done) {
    // FP32, two steps from a nonzero start: the same degree-2 polynomial,
    // T_2((theta - lambda) / delta) / T_2(theta / delta), as two damped-Jacobi
    // steps weighted by the reciprocals of its roots theta +- delta / sqrt 2
    // (the smaller weight first; the other order measured the same), so the
    // second step reads no previous iterate: one vector less per pass
    if (tmp == u_in || tmp == dst) return fail(s, VX_ERR_ARG, "chebyshev aliasing");
    CHK(tsweep<T>(tg, u_in, b, tmp, false, nullptr, rich2_weight(theta, delta, 0)));
    return tsweep<T>(tg, tmp, b, dst, dot, n_part, rich2_weight(theta, delta, 1));
  }
  auto target = [&](int j) { return j == k - 1 ? dst : (j % 2 == 0 ? tmp : tmp2); };
  const T* xprev = u_in;
  const T* xcur = u_in;
  double rho_prev = 1.0 / sigma, rho = rho_prev;
  for (int j = 0; j < k; ++j) {
    T* o = target(j);
    const bool last = (j == k - 1);
    if (j == 0) {
      if (o == xcur) return fail(s, VX_ERR_ARG, "chebyshev aliasing");
      if (!first_done) CHK(tsweep<T>(tg, xcur, b, o, dot && last, n_part, 1.0 / theta));
    } else {
      rho = 1.0 / (2.0 * sigma - rho_prev);
      SmoothCoef sc;
      sc.up = (j == 1 && !u_in) ? nullptr : (const void*)xprev;
      sc.cu = rho * rho_prev;
      sc.cr = 2.0 * rho / delta;
      // out may overwrite up (read only at the owner's own node, before its
      // write) but never u, whose halo other work-groups read
      if (o == xcur) return fail(s, VX_ERR_ARG, "chebyshev aliasing");
      if (dot && last)
        CHK((ttile<T, M_CHEB, true>(tg, xcur, b, o, n_part, &sc)));
      else
        CHK((ttile<T, M_CHEB, false>(tg, xcur, b, o, nullptr, &sc)));
      rho_prev = rho;
    }
    xprev = xcur;
    xcur = o;
  }
  return VX_OK;
}

// k damped-Jacobi sweeps starting at u_in (nullptr = zero) writing the result
// to `dst`, using `tmp` as the other ping-pong buffer.  dot: last sweep adds
// partials of out.b.
template <typename T, typename Tgt>
int tsmooth_n(const Tgt& tg, const T* u_in, const T* b, int k, T* dst, T* tmp, bool dot,
              int* n_part, bool first_done = false) {
  vx_solver* s = tsol(tg);
  T* bufs[2] = {dst, tmp};
  int cur = (k % 2 == 1) ? 0 : 1;   // first write goes to bufs[cur]
  const T* src = u_in;
  for (int i = 0; i < k; ++i) {
    T* o = bufs[cur];
    if (src == o) return fail(s, VX_ERR_ARG, "smoother aliasing");
    if (!(first_done && i == 0))
      CHK(tsweep<T>(tg, src, b, o, dot && i == k - 1, n_part, s->opt.omega));
    src = o;
    cur ^= 1;
  }
  return VX_OK;
}

// FP32 smoothing passes of levels li > 0 with at most BLK_LEVEL_MAX_NODES
// nodes may run blocked (k_smooth_block): a tile launch there costs its
// latency, not its bytes.  Untuned, the blocked form is used up to
// BLK_RULE_MAX_NODES (measured on MI355X: a step of a 17^3-node level takes
// 3.6 us blocked against 4.9 us per tile launch; at 33^3 both take 5.6 us);
// vx_autotune times both forms on every level up to the larger bound.
constexpr long long BLK_LEVEL_MAX_NODES = 40000;
constexpr long long BLK_RULE_MAX_NODES = 12000;

// level work vectors for max_batch right-hand sides
template <typename T>
int alloc_level_vectors(vx_solver* s) {
  for (size_t li = 0; li < s->lv.size(); ++li) {
    Level& L = s->lv[li];
    const size_t n = 3 * (size_t)L.Nn * s->opt.max_batch;
    if (li > 0) CHK(dalloc(s, &Sel<T>::b(L), n));
    CHK(dalloc(s, &Sel<T>::x(L), n));
    CHK(dalloc(s, &Sel<T>::t(L), n));
    CHK(dalloc(s, &Sel<T>::w(L), n));
    if (std::is_same<T, float>::value && li > 0 && L.Nn <= BLK_LEVEL_MAX_NODES) {
      CHK(dalloc(s, &L.v32a, n));
      CHK(dalloc(s, &L.v32b, n));
    }
  }
  return VX_OK;
}

// smoothing steps (pre and post) of level li
static int level_ns(const vx_solver* s, int li) {
  return li < 16 && s->opt.level_smooth[li] ? (int)s->opt.level_smooth[li] : s->opt.n_smooth;
}

// k smoothing steps from zero leave their first step in tmp (cheb_n for any
// k >= 2, smooth_n for even k): the step another kernel may write ahead
static bool first_in_tmp(const vx_solver* s, int k) {
  return s->opt.smoother == VX_SMOOTH_CHEBYSHEV ? k >= 2 : k >= 2 && k % 2 == 0;
}

// weight of the first smoothing step from zero of a k-step pass, x1 = c D^-1 b:
// 1/theta for Chebyshev, omega for damped Jacobi
static double first_step_coef(const vx_solver* s, int k) {
  if (s->opt.smoother == VX_SMOOTH_CHEBYSHEV) {
    const double beta = s->lam_max, alpha = cheb_ratio_for(s, k) * beta;
    return 1.0 / (0.5 * (beta + alpha));
  }
  return s->opt.omega;
}

// coefficients (cu, cr) of step j >= 1 of a k-step Chebyshev pass, the
// recurrence of tcheb_n
static SmoothCoef cheb_step_coef(const vx_solver* s, int k, int j) {
  const double beta = s->lam_max, alpha = cheb_ratio_for(s, k) * beta;
  const double theta = 0.5 * (beta + alpha), delta = 0.5 * (beta - alpha);
  const double sigma = theta / delta;
  double rho_prev = 1.0 / sigma, rho = rho_prev;
  SmoothCoef sc;
  for (int i = 1; i <= j; ++i) {
    rho = 1.0 / (2.0 * sigma - rho_prev);
    sc.cu = rho * rho_prev;
    sc.cr = 2.0 * rho / delta;
    rho_prev = rho;
  }
  return sc;
}

// The first two steps of a Chebyshev pass from zero on level li in one launch
// (M_CHEB_Z2): x1 = c1 D^-1 b formed pointwise while b is staged, then the
// second step into x; the pass's own coefficients (the recurrence of tcheb_n)
int launch_cheb_z2(vx_solver* s, int li, int k, const float* b, float* x) {
  Level& L = s->lv[li];
  const SmoothCoef sc = cheb_step_coef(s, k, 1);
  TileArgs<float> a;
  a.nx = L.nx; a.ny = L.ny; a.nz = L.nz;
  a.Lx = L.Lx; a.Ly = L.Ly; a.Lz = L.Lz;
  a.zc = zchunk(s, L, 1, 3);
  a.nb = s->nb;
  a.vstride = 3 * L.Nn;
  a.u = b; a.E = L.E32; a.mask = L.mask; a.b = nullptr; a.out = x;
  a.omega = 0.f;
  a.up = nullptr;
  a.cu = (float)sc.cu;
  a.cr = (float)sc.cr;
  a.kd = (float)s->kd;
  a.K = kconst<float>(s);
  a.partial = nullptr;
  a.minv = L.mi32;
  a.c1 = (float)first_step_coef(s, k);
  dim3 g = tile_grid(s, L, a.zc);
  a.gx = (int)g.x; a.gy = (int)g.y; a.gz = (int)g.z;
  const long long G = (long long)g.x * g.y * g.z;
  const long long launch = ((G * s->nb + 7) / 8) * 8;   // XCD-ordered 1D grid
  hipLaunchKernelGGL((k_tile<float, M_CHEB_Z2, false>), dim3((unsigned)launch), dim3(NT), 0,
                     s->stream, a);
  LAUNCHCHK();
  return VX_OK;
}

// smoothing step of the configured kind; result in dst (tmp, tmp2 scratch)
template <typename T, typename Tgt>
int tsmooth_step(const Tgt& tg, const T* u_in, const T* b, int k, T* dst, T* tmp, T* tmp2,
                 bool dot, int* n_part, bool first_done = false) {
  vx_solver* s = tsol(tg);
  const long long n = tvals(tg), off = tfirst(tg);
  if (first_done) {
    // the first step from zero (k_cg_update_x1 or the restriction) is
    // already in tmp, where the smoother writes it
    if (u_in || !first_in_tmp(s, k))
      return fail(s, VX_ERR_ARG, "a precomputed first step needs a start from zero and its step in tmp");
    if (s->opt.smoother == VX_SMOOTH_CHEBYSHEV)
      return tcheb_n<T>(tg, u_in, b, k, dst, tmp, tmp2, dot, n_part, true);
    return tsmooth_n<T>(tg, u_in, b, k, dst, tmp, dot, n_part, true);
  }
  if (s->opt.smoother == VX_SMOOTH_CHEBYSHEV) {
    if (u_in == dst && k == 1) {
      // the only write would land on the start vector: go through tmp
      CHK(tcheb_n<T>(tg, u_in, b, k, tmp, dst, tmp2, dot, n_part));
      return copy_vec<T>(s, dst + off, tmp + off, n);
    }
    return tcheb_n<T>(tg, u_in, b, k, dst, tmp, tmp2, dot, n_part);
  }
  if (u_in == dst) {
    // Jacobi from dst: ping-pong through tmp, copy back for odd counts
    if (k % 2 == 0) return tsmooth_n<T>(tg, u_in, b, k, dst, tmp, dot, n_part);
    CHK(tsmooth_n<T>(tg, u_in, b, k, tmp, dst, dot, n_part));
    return copy_vec<T>(s, dst + off, tmp + off, n);
  }
  return tsmooth_n<T>(tg, u_in, b, k, dst, tmp, dot, n_part);
}

static bool block_eligible(const vx_solver* s, int li, int k, bool dot);
static int smooth_blocked(vx_solver* s, int li, const float* u_in, const float* b, int k,
                          float* dst, float* tmp, float* tmp2, bool first_done);

template <typename T>
int smooth_step(vx_solver* s, int li, const T* u_in, const T* b, int k, T* dst, T* tmp, T* tmp2,
                bool dot, int* n_part, bool first_done = false) {
  if (std::is_same<T, float>::value && block_eligible(s, li, k, dot))
    return smooth_blocked(s, li, reinterpret_cast<const float*>(u_in),
                          reinterpret_cast<const float*>(b), k, reinterpret_cast<float*>(dst),
                          reinterpret_cast<float*>(tmp), reinterpret_cast<float*>(tmp2),
                          first_done);
  return tsmooth_step<T>(LevelTarget{s, li}, u_in, b, k, dst, tmp, tmp2, dot, n_part, first_done);
}

// the step coefficients of k smoothing steps (k <= COARSE_MAX_STEPS), the
// sequence cheb_n / smooth_n use on every level
static CoarseCoef smooth_coef(const vx_solver* s, int k) {
  CoarseCoef cc{};
  cc.steps = k;
  if (s->opt.smoother == VX_SMOOTH_CHEBYSHEV) {
    const double beta = s->lam_max, alpha = cheb_ratio_for(s, k) * beta;
    const double theta = 0.5 * (beta + alpha), delta = 0.5 * (beta - alpha);
    const double sigma = theta / delta;
    cc.cheb = 1;
    cc.w0 = 1.0 / theta;
    double rho_prev = 1.0 / sigma;
    for (int j = 1; j < k; ++j) {
      const double rho = 1.0 / (2.0 * sigma - rho_prev);
      cc.cu[j] = rho * rho_prev;
      cc.cr[j] = 2.0 * rho / delta;
      rho_prev = rho;
    }
  } else {
    cc.cheb = 0;
    cc.w0 = s->opt.omega;
  }
  return cc;
}

// ---- blocked smoothing passes (k_smooth_block) ----------------------------

static bool block_eligible(const vx_solver* s, int li, int k, bool dot) {
  if (dot || li <= 0 || li >= (int)s->lv.size() || s->opt.block_steps == 1 || k < 2 ||
      k > COARSE_MAX_STEPS)
    return false;
  const Level& L = s->lv[li];
  if (L.Nn > BLK_LEVEL_MAX_NODES || !L.v32a || !L.v32b) return false;
  if (s->opt.block_steps > 0 || s->opt.block_edge > 0) return true;   // requested
  if (L.blk_B != 0) return L.blk_B > 0;                               // tuned
  return L.Nn <= BLK_RULE_MAX_NODES;
}

// region extent per axis of a box of edge B grown by S (its largest, interior form)
static void block_region(const Level& L, int B, int S, long long* nr, long long* ne) {
  const long long rx = std::min<long long>(L.Lx, B + 2 * S);
  const long long ry = std::min<long long>(L.Ly, B + 2 * S);
  const long long rz = std::min<long long>(L.Lz, B + 2 * S);
  *nr = rx * ry * rz;
  *ne = (rx - 1) * (ry - 1) * (rz - 1);
}

static size_t block_lds_bytes(const Level& L, int B, int S) {
  long long nr, ne;
  block_region(L, B, S, &nr, &ne);
  return sizeof(float) * (size_t)(3 * nr + 25 * ne);
}

static bool block_fits(const Level& L, int B, int S) {
  long long nr, ne;
  block_region(L, B, S, &nr, &ne);
  return nr <= (long long)BLK_NPT * BLK_THREADS && block_lds_bytes(L, B, S) <= BLK_LDS_BYTES;
}

static long long block_count(const Level& L, int B) {
  return (long long)((L.Lx + B - 1) / B) * ((L.Ly + B - 1) / B) * ((L.Lz + B - 1) / B);
}

// threads of a blocked work-group: one element per thread in the first step
// (whole waves), at least half the region's nodes (BLK_NPT per thread)
static int block_threads(const Level& L, int B, int S) {
  long long nr, ne;
  block_region(L, B, S, &nr, &ne);
  const long long t = std::max(ne, (nr + BLK_NPT - 1) / BLK_NPT);
  return (int)std::min<long long>(BLK_THREADS, std::max<long long>(64, (t + 63) / 64 * 64));
}

// Box edge B and steps per launch S for a pass on level L: the options, the
// autotuned pair, or the pair of least estimated time per step.  Fitted to
// MI355X timings of 17^3- and 33^3-node levels: a launch costs about 3 us
// (dispatch, loads, stores) per round of co-resident work-groups, and each
// step about 1.3 us plus 4.5 ns per element application of the CU's
// work-groups.
struct BlockPlan {
  int B = 0, S = 0, nt = 0;
};

static BlockPlan block_plan(const vx_solver* s, const Level& L, int k) {
  BlockPlan best;
  double best_cost = 0;
  if (s->opt.block_steps <= 0 && s->opt.block_edge <= 0 && L.blk_B > 0 &&
      block_fits(L, L.blk_B, std::min(L.blk_S, k))) {
    best.B = L.blk_B;
    best.S = std::min(L.blk_S, k);
    best.nt = block_threads(L, best.B, best.S);
    return best;
  }
  // requested steps per launch: the most of at most that many that fit
  int s_lo = 2, s_hi = std::min(4, k);
  if (s->opt.block_steps > 0) {
    s_hi = std::min<int>(s->opt.block_steps, k);
    while (s_hi > 1) {
      bool any = false;
      for (int B = 2; B <= 16 && !any; ++B)
        any = (s->opt.block_edge <= 0 || B == s->opt.block_edge) && block_fits(L, B, s_hi);
      if (any) break;
      --s_hi;
    }
    s_lo = s_hi;
  }
  for (int S = std::min(s_lo, s_hi); S <= s_hi; ++S) {
    for (int B = 2; B <= 16; ++B) {
      if (s->opt.block_edge > 0 && B != s->opt.block_edge) continue;
      if (!block_fits(L, B, S)) continue;
      const long long nblk = block_count(L, B) * s->nb;
      const long long w = (nblk + s->n_cu - 1) / s->n_cu;   // work-groups per CU
      const long long conc = std::max<long long>(
          1, std::min<long long>(2048 / block_threads(L, B, S),
                                 BLK_LDS_BYTES / (long long)block_lds_bytes(L, B, S)));
      double t = 3.0 * (double)((w + conc - 1) / conc);
      for (int j = 0; j < S; ++j) {
        const int e = B + 2 * (S - 1 - j) + 1;
        const double ne = (double)std::min(L.nx, e) * std::min(L.ny, e) * std::min(L.nz, e);
        t += 1.3 + (double)w * ne * 4.5e-3;
      }
      const double cost = t / S;
      if (!best.B || cost < best_cost) {
        best.B = B;
        best.S = S;
        best_cost = cost;
      }
    }
  }
  if (best.B) best.nt = block_threads(L, best.B, best.S);
  return best;
}

// k smoothing steps of level li (the step sequence of smooth_step: from zero,
// from u_in with a first Jacobi-form step, or after a first step from zero
// already in tmp), `block_steps` per k_smooth_block launch.  A launch reads
// the iterate and the one before it and writes both anew: they rotate over
// tmp, tmp2 and the level's two blocked vectors, and the last launch writes
// dst, which no launch before it reads or writes unless it is u_in.
static int smooth_blocked(vx_solver* s, int li, const float* u_in, const float* b, int k,
                          float* dst, float* tmp, float* tmp2, bool first_done) {
  Level& L = s->lv[li];
  const CoarseCoef cc = smooth_coef(s, k);
  const BlockPlan plan = block_plan(s, L, k);
  if (!plan.B)   // no box of the requested edge fits the LDS: one launch per step
    return tsmooth_step<float>(LevelTarget{s, li}, u_in, b, k, dst, tmp, tmp2, false, nullptr,
                               first_done);
  const int S = plan.S, B = plan.B;
  int j = 0;
  const float* cur = u_in;
  const float* prev = nullptr;
  if (first_done) {
    if (u_in) return fail(s, VX_ERR_ARG, "a precomputed first step needs a start from zero");
    cur = tmp;   // x1 from zero; x0 = 0
    j = 1;
  }
  float* pool[4] = {tmp, tmp2, L.v32a, L.v32b};
  BlockArgs a;
  a.nx = L.nx; a.ny = L.ny; a.nz = L.nz;
  a.Lx = L.Lx; a.Ly = L.Ly; a.Lz = L.Lz;
  a.B = B;
  a.gx = (L.Lx + B - 1) / B; a.gy = (L.Ly + B - 1) / B; a.gz = (L.Lz + B - 1) / B;
  a.vstride = 3 * L.Nn;
  a.E = L.E32; a.mask = L.mask; a.minv = L.mi32; a.b = b;
  a.K = kconst<float>(s);
  const unsigned nblk = (unsigned)(a.gx * a.gy * a.gz);
  while (j < k) {
    int n = std::min(S, k - j);
    bool last = j + n == k;
    if (last && cur == dst) {   // the first launch reads dst (u_in): finish in a second one
      n = (k - j + 1) / 2;
      last = false;
    }
    float* o = dst;
    float* op = nullptr;
    if (!last) {
      o = nullptr;
      for (float* p : pool) {
        if (p == cur || p == prev || p == dst || p == b) continue;
        if (!o) {
          o = p;
        } else {
          op = p;
          break;
        }
      }
      if (!o || !op) return fail(s, VX_ERR_ARG, "blocked smoother: aliased buffers");
    }
    a.steps = n;
    for (int t = 0; t < n; ++t) {
      const int st = j + t;
      a.c1[t] = 0.f;
      if (st == 0) {
        a.form[t] = u_in ? BLK_JAC : BLK_ZERO;
        a.c0[t] = (float)cc.w0;
      } else if (cc.cheb) {
        a.form[t] = BLK_CHEB;
        a.c0[t] = (float)cc.cu[st];
        a.c1[t] = (float)cc.cr[st];
      } else {
        a.form[t] = BLK_JAC;
        a.c0[t] = (float)cc.w0;
      }
    }
    a.x = cur;
    a.xp = prev;
    a.xo = o;
    a.xpo = op;
    hipLaunchKernelGGL(k_smooth_block, dim3(nblk, s->nb), dim3(plan.nt),
                       block_lds_bytes(L, B, n), s->stream, a);
    LAUNCHCHK();
    prev = op;
    cur = o;
    j += n;
  }
  return VX_OK;
}

// The two bottom levels of the cycle in one launch (k_bottom): FP32, the
// second-to-last level li below the cycle's first level, within the kernel's
// LDS budget, and neither the exact coarse solve nor a residual capture on li
template <typename T>
bool bottom_fusable(const vx_solver* s, int li, int l0) {
  if (!std::is_same<T, float>::value) return false;
  const int nl = (int)s->lv.size();
  if (li != nl - 2 || li <= l0 || s->cd_ok) return false;
  if (s->cap_buf && s->cap_level == li) return false;
  const Level& F = s->lv[li];
  const Level& C = s->lv[li + 1];
  return F.Ne <= BOT_MAXE && F.Nn <= BOT_MAXN && C.Ne <= BOT_MAXEC && C.Nn <= BOT_MAXNC &&
         2 * C.nx == F.nx && 2 * C.ny == F.ny && 2 * C.nz == F.nz &&
         level_ns(s, li) <= COARSE_MAX_STEPS && s->opt.coarse_sweeps <= COARSE_MAX_STEPS;
}

static int bottom_fused(vx_solver* s, int li, const float* b, float* x) {
  Level& F = s->lv[li];
  Level& C = s->lv[li + 1];
  BottomArgs a;
  a.nx = F.nx;
  a.ny = F.ny;
  a.nz = F.nz;
  a.E = F.E32;
  a.mask = F.mask;
  a.minv = F.mi32;
  a.b = b;
  a.x = x;
  a.Ec = C.E32;
  a.maskc = C.mask;
  a.minvc = C.mi32;
  a.kf = level_ns(s, li);
  a.kc = s->opt.coarse_sweeps;
  a.K = kconst<float>(s);
  const CoarseCoef cf = smooth_coef(s, a.kf), cc = smooth_coef(s, a.kc);
  hipLaunchKernelGGL(k_bottom, dim3(s->nb), dim3(BOT_THREADS), 0, s->stream, a, cf, cc);
  LAUNCHCHK();
  return VX_OK;
}

// All coarsest-level smoothing steps in one single-work-group launch per
// right-hand side (same step sequence as smooth_step from zero).
template <typename T>
int coarse_fused(vx_solver* s, Level& L, const T* b, T* x) {
  const int k = s->opt.coarse_sweeps;
  const CoarseCoef cc = smooth_coef(s, k);
  // one element and one node per thread where 1024 threads allow (whole
  // waves; at most cs_slots<T>() of each per thread at the size limits)
  const long long work = std::max<long long>(L.Ne, L.Nn);
  const int threads = (int)std::min<long long>(1024, std::max<long long>(64, (work + 63) / 64 * 64));
  hipLaunchKernelGGL(k_coarse_smooth<T>, dim3(s->nb), dim3(threads), 0, s->stream, L.nx, L.ny, L.nz,
                     Sel<T>::E(L), L.mask, Sel<T>::mi(L), b, x, kconst<T>(s), cc);
  LAUNCHCHK();
  return VX_OK;
}

// x = K_ff^-1 b on the free DOFs of the coarsest level (0 elsewhere), FP64
// accumulation; the free part of b is staged in LDS.  blockIdx.y = rhs.
template <typename T>
__global__ __launch_bounds__(256) void k_coarse_direct(long long ndof, int nf,
                                                       const int* __restrict__ pos,
                                                       const int* __restrict__ idx,
                                                       const double* __restrict__ inv,
                                                       const T* __restrict__ b, T* __restrict__ x) {
  extern __shared__ double sb[];
  b += (long long)blockIdx.y * ndof;
  x += (long long)blockIdx.y * ndof;
  for (int j = threadIdx.x; j < nf; j += blockDim.x) sb[j] = (double)b[idx[j]];
  __syncthreads();
  const long long d = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (d >= ndof) return;
  const int p = pos[d];
  double acc = 0.0;
  if (p >= 0)
    for (int j = 0; j < nf; ++j) acc += inv[(long long)j * nf + p] * sb[j];
  x[d] = (T)acc;
}

// the coarsest level's correction by the direct solve
template <typename T>
int coarse_direct(vx_solver* s, const T* b, T* x) {
  Level& C = s->lv.back();
  const long long ndof = 3 * C.Nn;
  hipLaunchKernelGGL(k_coarse_direct<T>, dim3((unsigned)((ndof + 255) / 256), s->nb), dim3(256),
                     sizeof(double) * s->cd_nf, s->stream, ndof, s->cd_nf, s->cd_pos, s->cd_idx,
                     s->cd_inv, b, x);
  LAUNCHCHK();
  return VX_OK;
}

// first_done: 0 none, 1 the first level-l0 smoothing step from zero is in t0,
// 2 the level-l0 pre-smoothing pass (two Chebyshev steps from zero, FP32) runs
// as one M_CHEB_Z2 launch
template <typename T>
int vcycle(vx_solver* s, const T* b0, T* z0, T* t0, T* w0, bool dot, int* n_part, int l0,
           int first_done = 0);

// The plain V-cycle's coarse-grid correction of level li for residual t:
// restrict, V-cycle from level li+1, prolong (v = 0 + P x); at the coarsest
// level, its smoothing from zero.  tmp, tmp2: level-li scratch.
template <typename T>
int plain_coarse(vx_solver* s, int li, const T* t, T* v, T* tmp, T* tmp2) {
  const int nl = (int)s->lv.size();
  Level& L = s->lv[li];
  const int nb = s->nb;
  if (li == nl - 1) {
    const int k = s->opt.coarse_sweeps;
    if (s->cd_ok) return coarse_direct<T>(s, t, v);
    if (li > 0 && L.Nn <= coarse_max_nodes<T>() && L.Ne <= coarse_max_elems<T>() &&
        k <= COARSE_MAX_STEPS)
      return coarse_fused<T>(s, L, t, v);
    return smooth_step<T>(s, li, nullptr, t, k, v, tmp, tmp2, false, nullptr);
  }
  Level& C = s->lv[li + 1];
  CHK(restrict_to<T>(s, C, t, Sel<T>::b(C), nb));
  CHK(vcycle<T>(s, Sel<T>::b(C), Sel<T>::x(C), Sel<T>::t(C), Sel<T>::w(C), false, nullptr,
                li + 1));
  HIPCHK(hipMemsetAsync(v, 0, sizeof(T) * 3 * L.Nn * nb, s->stream));
  CHK(prolong_add<T>(s, li, Sel<T>::x(C), v));
  return VX_OK;
}

// network correction y [nb][3 Nn] at level li from t and v (FP32)
int corrector_net(vx_solver* s, int li, const float* t, const float* v, float* y, float* rms) {
  Level& L = s->lv[li];
  int rc = vxnet_api::prepare(s->corr, s->stream, s->nb, L.Lx, L.Ly, L.Lz);
  if (rc != VX_OK) return fail(s, rc, std::string("corrector: ") + vxnet_api::error(s->corr));
  hipLaunchKernelGGL(k_net_input, dim3(ew_grid(L.Nn), s->nb), dim3(EW_THREADS), 0, s->stream,
                     L.Nn, t, v, s->corr_feat, L.mi32, rms, s->corr_gain,
                     vxnet_api::input(s->corr));
  LAUNCHCHK();
  rc = vxnet_api::forward_bf16(s->corr, s->corr_out);
  if (rc != VX_OK) return fail(s, rc, std::string("corrector: ") + vxnet_api::error(s->corr));
  hipLaunchKernelGGL(k_net_apply, dim3(ew_grid(L.Nn), s->nb), dim3(EW_THREADS), 0, s->stream,
                     L.Nn, s->corr_out, L.mi32, L.mask, rms, s->corr_gain, y, 0);
  LAUNCHCHK();
  return VX_OK;
}

// Learned coarse-grid correction at level li for the s->nb right-hand sides:
// x += alpha v + beta y, the energy-optimal combination of the plain
// correction v and the network's y.  FP32 work only.
int corrector_apply(vx_solver* s, int li, const float* t, float* x) {
  Level& L = s->lv[li];
  const long long n3 = 3 * L.Nn;
  const int nb = s->nb;
  // (A v and A y are formed later: their buffers are the scratch here)
  CHK(plain_coarse<float>(s, li, t, s->corr_v, s->corr_Av, s->corr_Ay));
  int np = 0;
  CHK(dot_partials<float>(s, n3, t, t, &np));
  CHK(finalize(s, np, F_SET_CS));
  hipLaunchKernelGGL(k_net_rms, dim3(1), dim3(64), 0, s->stream, s->dscal, nb, (double)n3,
                     s->corr_rms);
  LAUNCHCHK();
  CHK(corrector_net(s, li, t, s->corr_v, s->corr_y, s->corr_rms));
  CHK((launch_tile_L<float, M_MATVEC, false>(s, L, s->corr_v, nullptr, s->corr_Av, nullptr,
                                             nullptr)));
  CHK((launch_tile_L<float, M_MATVEC, false>(s, L, s->corr_y, nullptr, s->corr_Ay, nullptr,
                                             nullptr)));
  hipLaunchKernelGGL(k_dot5, dim3(EW_BLOCKS, nb), dim3(EW_THREADS), 0, s->stream, n3, s->corr_v,
                     s->corr_y, s->corr_Av, s->corr_Ay, t, s->corr_part);
  hipLaunchKernelGGL(k_combine_coef, dim3(nb), dim3(256), 0, s->stream, s->corr_part, EW_BLOCKS,
                     s->corr_coef);
  hipLaunchKernelGGL(k_axpy2, dim3(ew_grid(n3), nb), dim3(EW_THREADS), 0, s->stream, n3,
                     s->corr_coef, s->corr_v, s->corr_y, x);
  LAUNCHCHK();
  return VX_OK;
}

// store the level residual of the first right-hand side (vx_capture)
template <typename T>
int capture_residual(vx_solver* s, int li, const T* t) {
  if (!s->cap_buf || li != s->cap_level || s->cap_count >= s->cap_cap) return VX_OK;
  if ((s->cap_calls++ % s->cap_every) != 0) return VX_OK;
  const long long n3 = 3 * s->lv[li].Nn;
  float* dst = s->cap_buf + (long long)s->cap_count * n3;
  if (std::is_same<T, float>::value) {
    HIPCHK(hipMemcpyAsync(dst, t, sizeof(float) * n3, hipMemcpyDeviceToDevice, s->stream));
  } else {
    hipLaunchKernelGGL((k_convert<T, float>), dim3(ew_grid(n3)), dim3(EW_THREADS), 0, s->stream,
                       n3, t, dst);
    LAUNCHCHK();
  }
  ++s->cap_count;
  return VX_OK;
}

// V-cycle (core/multigrid.py:426-436) from level l0: b at level l0 -> result
// in z0.  t0 and w0 are level-l0 scratch vectors (distinct from b0, z0).
// dot: partials of z0.b0 from the last level-l0 write.  With an attached
// corrector (FP32 work) the cycle bottoms out at the corrector level, whose
// coarse-grid correction (restrict, coarser levels, prolong) is replaced by
// the network's correction between its pre- and post-smoothing.
template <typename T>
int vcycle(vx_solver* s, const T* b0, T* z0, T* t0, T* w0, bool dot, int* n_part, int l0,
           int first_done) {
  const int nl = (int)s->lv.size();
  const int nb = s->nb;
  const bool learned = std::is_same<T, float>::value && s->corr_on && s->corr_level >= l0 &&
                       s->corr_level < nl;
  const int bottom = learned ? s->corr_level : nl - 1;
  // the two bottom levels in one launch (k_bottom) from level fuse_li
  const int fuse_li = !learned && bottom_fusable<T>(s, nl - 2, l0) ? nl - 2 : -1;
  // descend.  The restriction into a level that smooths from zero (not the
  // coarsest) also writes that level's first step into its t vector, where
  // the smoother expects it (first_in_tmp).
  bool restricted_x1 = false, z2_next = false;
  for (int li = l0; li <= bottom; ++li) {
    Level& L = s->lv[li];
    const T* b = li == l0 ? b0 : Sel<T>::b(L);
    T* x = li == l0 ? z0 : Sel<T>::x(L);
    T* t = li == l0 ? t0 : Sel<T>::t(L);
    T* w = li == l0 ? w0 : Sel<T>::w(L);
    const bool pre_all = std::is_same<T, float>::value && (li == l0 ? first_done == 2 : z2_next);
    const bool pre_done = li == l0 ? first_done == 1 : restricted_x1;
    const int ns = level_ns(s, li);
    if (li == bottom && learned) {
      if (pre_all)
        CHK(launch_cheb_z2(s, li, ns, reinterpret_cast<const float*>(b), reinterpret_cast<float*>(x)));
      else
        CHK(smooth_step<T>(s, li, nullptr, b, ns, x, t, w, false, nullptr, pre_done));
      CHK((launch_tile<T, M_RESID, false>(s, li, x, b, t, nullptr, nullptr)));
      if (l0 == 0) CHK(capture_residual<T>(s, li, t));
      CHK(corrector_apply(s, li, reinterpret_cast<const float*>(t), reinterpret_cast<float*>(x)));
      CHK(smooth_step<T>(s, li, x, b, ns, x, t, w, dot && li == l0, n_part));
      break;
    }
    if (li == fuse_li) {
      CHK(bottom_fused(s, li, reinterpret_cast<const float*>(b), reinterpret_cast<float*>(x)));
      break;
    }
    if (li == nl - 1) {
      const int k = s->opt.coarse_sweeps;
      if (s->cd_ok && li > 0 && !(dot && li == l0)) {
        CHK(coarse_direct<T>(s, b, x));
      } else if (li > 0 && L.Nn <= coarse_max_nodes<T>() && L.Ne <= coarse_max_elems<T>() &&
          k <= COARSE_MAX_STEPS) {
        CHK(coarse_fused<T>(s, L, b, x));
      } else {
        CHK(smooth_step<T>(s, li, nullptr, b, k, x, t, w, dot && li == l0, n_part));
      }
      break;
    }
    if (pre_all)
      CHK(launch_cheb_z2(s, li, ns, reinterpret_cast<const float*>(b), reinterpret_cast<float*>(x)));
    else
      CHK(smooth_step<T>(s, li, nullptr, b, ns, x, t, w, false, nullptr, pre_done));
    // residual into t, then restrict into b of level li+1 (with, for a level
    // that smooths from zero, its first step into its t vector)
    Level& C = s->lv[li + 1];
    restricted_x1 = li + 1 < nl - 1 && li + 1 != fuse_li && first_in_tmp(s, level_ns(s, li + 1));
    // a two-step Chebyshev pass from zero on a tile-smoothed level runs as one
    // M_CHEB_Z2 launch, which forms the first step itself
    z2_next = restricted_x1 && std::is_same<T, float>::value && li + 1 != bottom &&
              s->opt.smoother == VX_SMOOTH_CHEBYSHEV && level_ns(s, li + 1) == 2 &&
              !block_eligible(s, li + 1, 2, false);
    if (z2_next) restricted_x1 = false;
    const double x1w = restricted_x1 ? first_step_coef(s, level_ns(s, li + 1)) : 0.0;
    T* const cx1 = restricted_x1 ? Sel<T>::t(C) : nullptr;
    if (std::is_same<T, float>::value && !(l0 == 0 && s->cap_buf && s->cap_level == li)) {
      // the residual goes from the tiles straight into the restriction
      CHK(launch_resr(s, li, reinterpret_cast<const float*>(x), reinterpret_cast<const float*>(b),
                      reinterpret_cast<float*>(Sel<T>::b(C)), reinterpret_cast<float*>(cx1), x1w));
    } else {
      CHK((launch_tile<T, M_RESID, false>(s, li, x, b, t, nullptr, nullptr)));
      if (l0 == 0) CHK(capture_residual<T>(s, li, t));
      if (restricted_x1)
        CHK(restrict_to<T>(s, C, t, Sel<T>::b(C), nb, cx1, x1w));
      else
        CHK(restrict_to<T>(s, C, t, Sel<T>::b(C), nb));
    }
  }
  // ascend (from below the fused levels' result when k_bottom ran)
  for (int li = (fuse_li >= 0 ? fuse_li : bottom) - 1; li >= l0; --li) {
    Level& L = s->lv[li];
    Level& C = s->lv[li + 1];
    const T* b = li == l0 ? b0 : Sel<T>::b(L);
    T* x = li == l0 ? z0 : Sel<T>::x(L);
    T* t = li == l0 ? t0 : Sel<T>::t(L);
    T* w = li == l0 ? w0 : Sel<T>::w(L);
    CHK(prolong_add<T>(s, li, Sel<T>::x(C), x));
    CHK(smooth_step<T>(s, li, x, b, level_ns(s, li), x, t, w, dot && li == l0, n_part));
  }
  return VX_OK;
}

int require_ready(vx_solver* s) {
  if (!s->fixed_set) return fail(s, VX_ERR_ARG, "call vx_set_fixed() before solving");
  if (!s->density_set) return fail(s, VX_ERR_ARG, "call vx_set_density() before solving");
  return VX_OK;
}

// copy the nb scalar structs to the pinned host mirror; event marks arrival
int read_scalars(vx_solver* s, hipEvent_t ev) {
  HIPCHK(hipMemcpyAsync(s->hscal, s->dscal, sizeof(Scalars) * s->nb, hipMemcpyDeviceToHost,
                        s->stream));
  HIPCHK(hipEventRecord(ev, s->stream));
  return VX_OK;
}

int sync_scalars(vx_solver* s) {
  HIPCHK(hipMemcpyAsync(s->hscal, s->dscal, sizeof(Scalars) * s->nb, hipMemcpyDeviceToHost,
                        s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  return VX_OK;
}

// push the host's frozen flags of all right-hand sides to the device
int push_frozen(vx_solver* s, const std::vector<int>& fr) {
  for (int r = 0; r < s->nb; ++r) s->hflags[r] = fr[r];
  HIPCHK(hipMemcpyAsync(s->dflags, s->hflags, sizeof(int) * s->nb, hipMemcpyHostToDevice,
                        s->stream));
  hipLaunchKernelGGL(k_set_frozen, dim3(1), dim3(256), 0, s->stream, s->dscal, s->dflags, s->nb);
  LAUNCHCHK();
  return VX_OK;
}

// A V-cycle holds only kernel launches unless a copy, a corrector or a
// residual capture is in it: then it cannot be replayed from a graph
bool cycle_graphable(const vx_solver* s) {
  if (s->opt.eager || s->graphs_failed || s->corr_on || s->cap_buf) return false;
  for (int li = 0; li < (int)s->lv.size(); ++li) {
    const int k = level_ns(s, li);
    if (s->opt.smoother == VX_SMOOTH_CHEBYSHEV ? k == 1 : k % 2 == 1) return false;   // copy_vec
  }
  return true;
}

// Run `enqueue` (kernel launches on s->stream) from a HIP graph captured
// once per key on a stream that only records, and replayed on s->stream (the
// null stream included): one submission instead of one per kernel.  If
// capture or instantiation fails the launches run directly, from then on.
template <typename F>
int run_graphed(vx_solver* s, const void* const (&ptr)[4], int type, int l0, int first_done,
                int flexible, F&& enqueue) {
  if (!cycle_graphable(s)) return enqueue();
  for (const auto& g : s->graphs)
    if (g.type == type && g.l0 == l0 && g.nb == s->nb && g.first_done == first_done &&
        g.flexible == flexible && std::equal(ptr, ptr + 4, g.ptr)) {
      HIPCHK(hipGraphLaunch(g.exec, s->stream));
      return VX_OK;
    }
  if (!s->cap_stream && hipStreamCreateWithFlags(&s->cap_stream, hipStreamNonBlocking) != hipSuccess) {
    (void)hipGetLastError();
    s->graphs_failed = true;
    return enqueue();
  }
  hipStream_t run = s->stream;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  bool ok = hipStreamBeginCapture(s->cap_stream, hipStreamCaptureModeThreadLocal) == hipSuccess;
  int rc = VX_OK;
  if (ok) {
    s->stream = s->cap_stream;
    rc = enqueue();
    s->stream = run;
    ok = hipStreamEndCapture(s->cap_stream, &graph) == hipSuccess && rc == VX_OK && graph;
  }
  if (ok) ok = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) == hipSuccess;
  if (graph) (void)hipGraphDestroy(graph);
  if (!ok) {
    (void)hipGetLastError();
    s->graphs_failed = true;
    return enqueue();
  }
  vx_solver::CycleGraph g{};
  std::copy(ptr, ptr + 4, g.ptr);
  g.type = type;
  g.l0 = l0;
  g.nb = s->nb;
  g.first_done = first_done;
  g.flexible = flexible;
  g.exec = exec;
  s->graphs.push_back(g);
  HIPCHK(hipGraphLaunch(exec, s->stream));
  return VX_OK;
}

struct PcgResult {
  int iters = 0;
  bool converged = false;
  bool stalled = false;
  double ratio = NAN;
};

// Preconditioned CG of core/multigrid.py:474-546 on level-0 vectors of type T,
// for the s->nb right-hand sides at once.  On entry r holds the masked
// initial residuals and u the initial iterates; the device scalar fnorm of
// each right-hand side is its ratio's denominator.  A right-hand side stops
// (its updates frozen) when ||r||/fnorm < tol_rel, or on its stall guard; the
// loop ends when all have stopped or max_iter is reached.  The history and
// the callback see ratios times ratio_scale[r] (1 for a plain solve; the
// outer residual ratio for an inner solve of iterative refinement).
//
// l0 > 0 runs the same iteration on level l0's system (its V-cycle starts
// there).  With the learned corrector active (FP32) the preconditioner is
// not a fixed linear operator, so the direction update is the flexible one,
// beta = -z_new.A p / p.A p (A-orthogonal to the previous direction), which
// equals Fletcher-Reeves for a fixed symmetric preconditioner.
template <typename T>
int pcg(vx_solver* s, T* u, T* r, T* z, T* p, T* Ap, T* t0, const std::vector<double>& tol_rel, int max_iter,
        bool stall_guard, const std::vector<double>& ratio_scale, double cb_tol,
        const std::vector<int>& iter_offset, vx_progress_cb cb, void* user, double* hist,
        int hist_cap, std::vector<PcgResult>& res, const std::vector<int>& skip, int l0 = 0,
        bool keep_p = false) {
  Level& L0 = s->lv[l0];
  const bool flexible = (std::is_same<T, float>::value && s->corr_on) || s->opt.flexible;
  const int nb = s->nb;
  const long long n = 3 * L0.Nn;
  int np = 0;
  hipEvent_t ev;
  HIPCHK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
  auto cleanup = [&]() { (void)hipEventDestroy(ev); };

  res.assign(nb, PcgResult());
  std::vector<int> frozen(skip.begin(), skip.end());
  int rc = vcycle<T>(s, r, z, t0, Sel<T>::w(L0), true, &np, l0);
  // also alpha = 0, frozen = 0; beta = 0, or with keep_p the previous
  // solve's direction continued (p is kept, scaled through beta)
  if (rc == VX_OK) rc = finalize(s, np, keep_p ? F_RZ0_KEEP : F_RZ0);
  if (rc == VX_OK && std::any_of(frozen.begin(), frozen.end(), [](int f) { return f; }))
    rc = push_frozen(s, frozen);
  if (rc != VX_OK) { cleanup(); return rc; }
  // FP32 with a smoothing pass from zero below the coarsest level whose
  // first step lands in t0 (first_in_tmp): the update kernel also writes the
  // next V-cycle's first step there
  const auto al16 = [](const void* q) { return (reinterpret_cast<uintptr_t>(q) & 15) == 0; };
  const bool fuse_x1 = std::is_same<T, float>::value && first_in_tmp(s, level_ns(s, l0)) &&
                       l0 < (int)s->lv.size() - 1 && (nb == 1 || (3 * L0.Nn) % 4 == 0) &&
                       al16(u) && al16(r) && al16(p) && al16(Ap) && al16(t0) && al16(L0.mask) &&
                       al16(L0.mi32);
  // with that fusion the solution update u += alpha p is deferred to the next
  // direction update (which streams p anyway) and a flush at the end
  const bool lazy_u = fuse_x1;
  const double x1_coef = first_step_coef(s, level_ns(s, l0));
  // a two-step Chebyshev pre-smoothing pass on level l0 runs as one launch
  // that forms its first step from r itself (M_CHEB_Z2): the update then
  // writes only r (k_cg_update_r), not the first step
  const bool fuse_z2 = fuse_x1 && s->opt.smoother == VX_SMOOTH_CHEBYSHEV &&
                       level_ns(s, l0) == 2 && !(s->corr_on && s->corr_level == l0);
  // p starts at 0 so that the first p = z + 0 * p is exact
  if (!keep_p && hipMemsetAsync(p, 0, sizeof(T) * n * nb, s->stream) != hipSuccess) {
    cleanup();
    return fail(s, VX_ERR_HIP, "memset p");
  }
  std::vector<double> best(nb, INFINITY), prev(nb, NAN), prev2(nb, NAN);
  for (int it = 0; it < max_iter; ++it) {
    // p = z + beta p, then Ap with the p.Ap partials.  (Forming p inside the
    // operator kernel instead costs more than this extra streaming pass: that
    // kernel then stages two halo streams and writes two outputs.)
    if (lazy_u)
      hipLaunchKernelGGL(k_p_update_u<T>, dim3(EW_BLOCKS, nb), dim3(EW_THREADS), 0, s->stream, n,
                         s->dscal, z, p, u);
    else
      hipLaunchKernelGGL(k_p_update_b<T>, dim3(EW_BLOCKS, nb), dim3(EW_THREADS), 0, s->stream, n,
                         s->dscal, z, p);
    if (hipGetLastError() != hipSuccess) { rc = fail(s, VX_ERR_HIP, "p update launch"); break; }
    rc = launch_tile_L<T, M_MATVEC, true>(s, L0, p, nullptr, Ap, s->partial, &np);
    if (rc == VX_OK) rc = finalize(s, np, F_ALPHA);
    if (rc != VX_OK) break;
    if (fuse_z2) {
      // r -= alpha Ap (u += alpha p waits for the next direction update or the flush)
      hipLaunchKernelGGL(k_cg_update_r, dim3(EW_BLOCKS, nb), dim3(EW_THREADS), 0, s->stream, n,
                         s->dscal, reinterpret_cast<float*>(r),
                         reinterpret_cast<const float*>(Ap), s->partial);
    } else if (fuse_x1) {
      // also the next V-cycle's first smoothing step from zero, into t0
      // (u += alpha p waits for the next direction update or the flush)
      hipLaunchKernelGGL(k_cg_update_x1, dim3(EW_BLOCKS, nb), dim3(EW_THREADS), 0, s->stream,
                         L0.Nn, s->dscal, reinterpret_cast<float*>(r),
                         reinterpret_cast<const float*>(Ap), s->partial, L0.mask, L0.mi32,
                         (float)x1_coef, reinterpret_cast<float*>(t0));
    } else {
      hipLaunchKernelGGL(k_cg_update<T>, dim3(EW_BLOCKS, nb), dim3(EW_THREADS), 0, s->stream, n,
                         s->dscal, u, r, p, Ap, s->partial);
    }
    if (hipGetLastError() != hipSuccess) { rc = fail(s, VX_ERR_HIP, "cg_update launch"); break; }
    rc = finalize(s, EW_BLOCKS, F_RATIO);
    if (rc == VX_OK) rc = read_scalars(s, ev);
    if (rc != VX_OK) break;
    // Queue the next preconditioner application while the host waits for the
    // ratios, unless every open right-hand side looks about to converge
    // (then waiting is cheaper than a wasted V-cycle).
    bool predicted_done = true;
    for (int q = 0; q < nb; ++q) {
      if (frozen[q]) continue;
      bool d = false;
      if (!std::isnan(prev[q]) && !std::isnan(prev2[q]) && prev2[q] > 0)
        d = prev[q] * (prev[q] / prev2[q]) < 3.0 * tol_rel[q];
      predicted_done = predicted_done && d;
    }
    // the V-cycle and its beta: the same launches every step, replayed from
    // a captured graph where the cycle allows it
    auto cycle_and_beta = [&]() -> int {
      int np2 = 0;
      CHK(vcycle<T>(s, r, z, t0, Sel<T>::w(L0), true, &np2, l0, fuse_z2 ? 2 : (fuse_x1 ? 1 : 0)));
      if (!flexible) return finalize(s, np2, F_BETA);
      CHK(finalize(s, np2, F_RZNEW));
      CHK(dot_partials<T>(s, n, z, Ap, &np2));
      return finalize(s, np2, F_BETA_FLEX);
    };
    auto queue_next = [&]() -> int {
      const void* const key[4] = {r, z, t0, Ap};
      return run_graphed(s, key, std::is_same<T, float>::value ? 1 : 0, l0,
                         fuse_z2 ? 2 : (fuse_x1 ? 1 : 0), flexible ? 1 : 0, cycle_and_beta);
    };
    bool queued = false;
    if (!predicted_done) {
      rc = queue_next();
      if (rc != VX_OK) break;
      queued = true;
    }
    if (hipEventSynchronize(ev) != hipSuccess) { rc = fail(s, VX_ERR_HIP, "event sync"); break; }
    bool changed = false;
    double worst = 0.0;
    for (int q = 0; q < nb; ++q) {
      if (frozen[q]) continue;
      const Scalars sc = s->hscal[q];
      PcgResult& R = res[q];
      if (sc.pAp == 0.0) {   // nothing left along p (core/multigrid.py:500-511)
        R.converged = R.ratio < tol_rel[q];
        frozen[q] = 1;
        changed = true;
        continue;
      }
      const double ratio = sc.ratio;
      R.iters = it + 1;
      R.ratio = ratio;
      const int gi = iter_offset[q] + it;
      if (hist && gi < hist_cap) hist[(long long)q * hist_cap + gi] = ratio * ratio_scale[q];
      worst = std::max(worst, ratio * ratio_scale[q]);
      if (ratio < tol_rel[q]) {
        R.converged = true;
        frozen[q] = 1;
        changed = true;
        continue;
      }
      best[q] = std::min(best[q], ratio);
      if (stall_guard && it + 1 >= 10 && ratio > 5.0 * best[q]) {
        R.stalled = true;
        frozen[q] = 1;
        changed = true;
        char buf[200];
        snprintf(buf, sizeof buf,
                 "MG-PCG stalled/diverged at iter %d (resid ratio %.3e, best seen %.3e)", it + 1,
                 ratio, best[q]);
        s->err = buf;
        continue;
      }
      prev2[q] = prev[q];
      prev[q] = ratio;
    }
    if (cb) cb(iter_offset[0] + it + 1, worst, cb_tol, user);
    const bool all_done = std::all_of(frozen.begin(), frozen.end(), [](int f) { return f; });
    if (all_done) break;
    if (changed) {
      // stopped right-hand sides get alpha = 0 from the next finalize on;
      // the flags land before the queued V-cycle's beta, which is harmless
      rc = push_frozen(s, frozen);
      if (rc != VX_OK) break;
    }
    if (!queued) {
      rc = queue_next();
      if (rc != VX_OK) break;
    }
  }
  if (lazy_u && rc == VX_OK) {
    // the last step's u += alpha p (alpha is 0 for a right-hand side that
    // stopped earlier: its pending update was applied by the step after)
    hipLaunchKernelGGL(k_u_flush<T>, dim3(EW_BLOCKS, nb), dim3(EW_THREADS), 0, s->stream, n,
                       s->dscal, p, u);
    if (hipGetLastError() != hipSuccess) rc = fail(s, VX_ERR_HIP, "u flush launch");
  }
  cleanup();
  return rc;
}

// corrector buffers sized for level `level` (and max_batch right-hand sides)
int corr_buffers(vx_solver* s, int level) {
  const Level& L = s->lv[level];
  const size_t mb = (size_t)s->opt.max_batch;
  if (L.Nn > s->corr_cap_nn) {
    // older (smaller) buffers stay owned by s->allocs
    s->corr_feat = s->corr_out = s->corr_v = s->corr_y = s->corr_Av = s->corr_Ay = nullptr;
    CHK(dalloc(s, &s->corr_feat, 4 * (size_t)L.Nn));
    CHK(dalloc(s, &s->corr_out, 16 * (size_t)L.Nn * mb));
    CHK(dalloc(s, &s->corr_v, 3 * (size_t)L.Nn * mb));
    CHK(dalloc(s, &s->corr_y, 3 * (size_t)L.Nn * mb));
    CHK(dalloc(s, &s->corr_Av, 3 * (size_t)L.Nn * mb));
    CHK(dalloc(s, &s->corr_Ay, 3 * (size_t)L.Nn * mb));
    s->corr_cap_nn = L.Nn;
  }
  CHK(dalloc(s, &s->corr_part, 5 * (size_t)EW_BLOCKS * mb));
  CHK(dalloc(s, &s->corr_coef, 2 * mb));
  CHK(dalloc(s, &s->corr_rms, mb));
  CHK(dalloc(s, &s->corr_emax, 1));
  return VX_OK;
}

// node features of the corrector level from its E and mask
int corr_features(vx_solver* s, int level) {
  Level& L = s->lv[level];
  HIPCHK(hipMemsetAsync(s->corr_emax, 0, sizeof(unsigned int), s->stream));
  hipLaunchKernelGGL(k_max_nonneg, dim3(ew_grid(L.Ne)), dim3(EW_THREADS), 0, s->stream, L.Ne,
                     L.E32, s->corr_emax);
  hipLaunchKernelGGL(k_net_feat, dim3(ew_grid(L.Nn)), dim3(EW_THREADS), 0, s->stream, L.nx, L.ny,
                     L.nz, L.E32, L.mask, s->corr_emax, s->corr_feat);
  LAUNCHCHK();
  return VX_OK;
}

}  // namespace

static int alloc_pcg64(vx_solver* s);

// Direct solve of the coarsest level: assemble its operator densely on the
// host from the level's E (sum of E_e KE over elements, the same operator
// the tile kernel applies), keep the free DOFs, factor (Cholesky, FP64) and
// invert.  Leaves cd_ok false (the V-cycle then smooths as configured) when
// the level has more than CD_MAX_FREE free DOFs or the free block is not
// positive definite (a support that injection lost on the coarse grid).
constexpr int CD_MAX_FREE = 1500;
static int build_coarse_direct(vx_solver* s) {
  s->cd_ok = false;
  Level& C = s->lv.back();
  if (s->lv.size() < 2) return VX_OK;
  const long long ndof = 3 * C.Nn;
  std::vector<double> E(C.Ne);
  std::vector<uint8_t> m(C.Nn);
  HIPCHK(hipMemcpyAsync(E.data(), C.E64, sizeof(double) * C.Ne, hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipMemcpyAsync(m.data(), C.mask, C.Nn, hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  std::vector<int> pos(ndof, -1), idx;
  for (long long d = 0; d < ndof; ++d)
    if ((m[d / 3] >> (d % 3)) & 1u) {
      pos[d] = (int)idx.size();
      idx.push_back((int)d);
    }
  const int nf = (int)idx.size();
  if (nf == 0 || nf > CD_MAX_FREE) return VX_OK;
  // unit-E element matrix in binary corner order
  double KE[24][24];
  for (int j = 0; j < 24; ++j) {
    double u[8][3] = {}, y[8][3];
    u[j / 3][j % 3] = 1.0;
    elem_apply<double>(u, 1.0, s->K64, y);
    for (int i = 0; i < 24; ++i) KE[i][j] = y[i / 3][i % 3];
  }
  std::vector<double> A((size_t)nf * nf, 0.0);
  for (int ez = 0; ez < C.nz; ++ez)
    for (int ey = 0; ey < C.ny; ++ey)
      for (int ex = 0; ex < C.nx; ++ex) {
        const double Ee = E[((long long)ez * C.ny + ey) * C.nx + ex];
        if (Ee == 0.0) continue;
        int row[24];
        for (int q = 0; q < 8; ++q) {
          const long long n = (ex + (q & 1)) + (long long)C.Lx * ((ey + ((q >> 1) & 1)) +
                                                                 (long long)C.Ly * (ez + (q >> 2)));
          for (int c = 0; c < 3; ++c) row[3 * q + c] = pos[3 * n + c];
        }
        for (int i = 0; i < 24; ++i) {
          if (row[i] < 0) continue;
          for (int j = 0; j < 24; ++j)
            if (row[j] >= 0) A[(size_t)row[i] * nf + row[j]] += Ee * KE[i][j];
        }
      }
  // Cholesky A = L L^T (lower triangle in place)
  double dmax = 0.0;
  for (int i = 0; i < nf; ++i) dmax = std::max(dmax, A[(size_t)i * nf + i]);
  for (int j = 0; j < nf; ++j) {
    double d = A[(size_t)j * nf + j];
    for (int k = 0; k < j; ++k) d -= A[(size_t)j * nf + k] * A[(size_t)j * nf + k];
    if (!(d > 1e-13 * dmax)) return VX_OK;   // not positive definite
    const double ljj = std::sqrt(d);
    A[(size_t)j * nf + j] = ljj;
    for (int i = j + 1; i < nf; ++i) {
      double v = A[(size_t)i * nf + j];
      const double* ri = &A[(size_t)i * nf];
      const double* rj = &A[(size_t)j * nf];
      for (int k = 0; k < j; ++k) v -= ri[k] * rj[k];
      A[(size_t)i * nf + j] = v / ljj;
    }
  }
  // inverse: solve L L^T X = I column by column (X symmetric)
  std::vector<double> inv((size_t)nf * nf), col(nf);
  for (int c = 0; c < nf; ++c) {
    for (int i = 0; i < nf; ++i) {
      double v = (i == c) ? 1.0 : 0.0;
      const double* ri = &A[(size_t)i * nf];
      for (int k = c; k < i; ++k) v -= ri[k] * col[k];
      col[i] = (i < c) ? 0.0 : v / ri[i];
    }
    for (int i = nf - 1; i >= 0; --i) {
      double v = col[i];
      for (int k = i + 1; k < nf; ++k) v -= A[(size_t)k * nf + i] * col[k];
      col[i] = v / A[(size_t)i * nf + i];
    }
    for (int i = 0; i < nf; ++i) inv[(size_t)c * nf + i] = col[i];
  }
  if ((long long)nf * nf > s->cd_cap_inv) {
    s->cd_inv = nullptr;   // smaller buffers stay owned by s->allocs
    CHK(dalloc(s, &s->cd_inv, (size_t)nf * nf));
    s->cd_cap_inv = (long long)nf * nf;
  }
  if (ndof > s->cd_cap_dof) {
    s->cd_pos = nullptr;
    s->cd_idx = nullptr;
    CHK(dalloc(s, &s->cd_pos, (size_t)ndof));
    CHK(dalloc(s, &s->cd_idx, (size_t)ndof));
    s->cd_cap_dof = ndof;
  }
  HIPCHK(hipMemcpyAsync(s->cd_inv, inv.data(), sizeof(double) * nf * nf, hipMemcpyHostToDevice,
                        s->stream));
  HIPCHK(hipMemcpyAsync(s->cd_pos, pos.data(), sizeof(int) * ndof, hipMemcpyHostToDevice,
                        s->stream));
  HIPCHK(hipMemcpyAsync(s->cd_idx, idx.data(), sizeof(int) * nf, hipMemcpyHostToDevice,
                        s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  s->cd_nf = nf;
  s->cd_ok = true;
  return VX_OK;
}

// PCG on level l's own system (V-cycle from l down): x = A_l^-1 b to tol.
static int level_solve(vx_solver* s, int l, const double* b, double* x, double tol, int max_iter,
                       PcgResult* out) {
  CHK(alloc_pcg64(s));
  s->nb = 1;
  Level& L = s->lv[l];
  const long long n3 = 3 * L.Nn;
  hipLaunchKernelGGL((k_mask_copy<double, double>), dim3(EW_BLOCKS, 1), dim3(EW_THREADS), 0,
                     s->stream, L.Nn, L.mask, b, s->r64, s->partial);
  LAUNCHCHK();
  CHK(finalize(s, EW_BLOCKS, F_FNORM));
  HIPCHK(hipMemsetAsync(x, 0, sizeof(double) * n3, s->stream));
  CHK(sync_scalars(s));
  *out = PcgResult();
  if (s->hscal[0].fnorm == 0.0) {
    out->converged = true;
    out->ratio = 0.0;
    return VX_OK;
  }
  std::vector<double> one(1, 1.0);
  std::vector<int> zero(1, 0);
  std::vector<PcgResult> pr;
  CHK(pcg<double>(s, x, s->r64, s->z64, s->p64, s->Ap64, L.t64, std::vector<double>(1, tol),
                  max_iter, false, one, tol, zero, nullptr, nullptr, nullptr, 0, pr, zero, l));
  *out = pr[0];
  return VX_OK;
}

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

void vx_default_options(vx_options* o) {
  std::memset(o, 0, sizeof(*o));
  o->struct_size = sizeof(vx_options);
  o->max_levels = 6;
  o->min_elems = 4;
  o->n_smooth = 2;
  o->omega = 0.6;
  o->coarse_sweeps = 30;
  o->precision = VX_FP64;
  o->smoother = VX_SMOOTH_JACOBI;
  o->coarse_mode = VX_COARSE_SWEEPS;
  o->cheb_ratio_power = 0;
  o->cheb_ratio = 0.3;
  o->inner_tol = 1e-5;
  o->max_outer = 20;
  o->z_chunk = 0;
  o->e_coarsen = 0;
  o->max_batch = 1;
}

int vx_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return n;
}

const char* vx_last_error(vx_handle h) { return h ? h->err.c_str() : "null handle"; }

void vx_element_apply_host(double nu, const double* u24, double* y24) {
  ElemConst<double> K = elem_const(nu);
  double u[8][3], y[8][3];
  for (int k = 0; k < 8; ++k)
    for (int c = 0; c < 3; ++c) u[ref_to_bin(k)][c] = u24[3 * k + c];
  elem_apply<double>(u, 1.0, K, y);
  for (int k = 0; k < 8; ++k)
    for (int c = 0; c < 3; ++c) y24[3 * k + c] = y[ref_to_bin(k)][c];
}

}  // extern "C"

namespace {

// hierarchy rule of core/multigrid.py:141-148: halve while every dim is even,
// the smallest stays >= min_elems / 2 and max_levels is not reached
std::vector<std::array<int, 3>> level_dims_rule(int nx, int ny, int nz, const vx_options& o) {
  std::vector<std::array<int, 3>> dims;
  dims.push_back({nx, ny, nz});
  while ((int)dims.size() < o.max_levels) {
    auto d = dims.back();
    const int mn = std::min(d[0], std::min(d[1], d[2]));
    if (d[0] % 2 || d[1] % 2 || d[2] % 2) break;
    if (!(mn / 2 >= o.min_elems / 2 && mn >= 2)) break;
    dims.push_back({d[0] / 2, d[1] / 2, d[2] / 2});
  }
  return dims;
}

// Copy and check the options (nullptr: defaults).  Every Chebyshev pass needs
// an interval [r lambda, lambda] with 0 < r < 1 (r = 1 makes its width 0 and
// the step coefficients divide by it): the ratio of each level's step count
// and of the coarsest sweeps of a hierarchy with `nlev` levels is checked.
int check_options(const vx_options* opt, int nlev, vx_options* o) {
  if (opt) {
    if (opt->struct_size != (int32_t)sizeof(vx_options)) return VX_ERR_ARG;
    *o = *opt;
  } else {
    vx_default_options(o);
  }
  if (o->max_batch < 1) o->max_batch = 1;
  if (o->n_smooth < 1 || o->coarse_sweeps < 1 || o->max_levels < 1 ||
      (o->coarse_mode != VX_COARSE_SWEEPS && o->coarse_mode != VX_COARSE_DIRECT))
    return VX_ERR_ARG;
  if (o->block_steps < 0 || o->block_steps > BLK_MAX_STEPS || o->block_edge < 0 ||
      o->block_edge == 1 || o->block_edge > 16 || o->keep_direction < 0 || o->keep_direction > 1 ||
      !(o->keep_drift >= 0.0))
    return VX_ERR_ARG;
  if (o->smoother == VX_SMOOTH_CHEBYSHEV) {
    vx_solver probe;
    probe.opt = *o;
    for (int li = 0; li < nlev; ++li) {
      const int k = li + 1 == nlev ? o->coarse_sweeps : level_ns(&probe, li);
      const double r = cheb_ratio_for(&probe, k);
      if (!(r > 0.0 && r < 1.0)) return VX_ERR_ARG;
    }
  }
  return VX_OK;
}

}  // namespace

extern "C" {

int vx_create(int32_t nx, int32_t ny, int32_t nz, double nu, const vx_options* opt,
              vx_handle* out) {
  if (!out) return VX_ERR_ARG;
  *out = nullptr;
  if (nx < 1 || ny < 1 || nz < 1 || !(nu > -1.0 && nu < 0.5)) return VX_ERR_ARG;
  vx_options o;
  if (check_options(opt, 0, &o) != VX_OK) return VX_ERR_ARG;
  const std::vector<std::array<int, 3>> dims = level_dims_rule(nx, ny, nz, o);
  if (check_options(opt, (int)dims.size(), &o) != VX_OK) return VX_ERR_ARG;
  if (vx_device_count() < 1) return VX_ERR_NO_DEVICE;
  vx_solver* s = new vx_solver();
  s->nx = nx; s->ny = ny; s->nz = nz; s->nu = nu;
  s->opt = o;
  s->K64 = elem_const(nu);
  s->K32 = {(float)s->K64.k1, (float)s->K64.k2, (float)s->K64.b1, (float)s->K64.b2,
            (float)s->K64.t1};
  s->kd = elem_diag(nu);
  s->lam_max = elem_lambda_max(nu) / s->kd;
  long long maxpart = 0;
  for (auto& d : dims) {
    Level L;
    L.nx = d[0]; L.ny = d[1]; L.nz = d[2];
    L.Lx = d[0] + 1; L.Ly = d[1] + 1; L.Lz = d[2] + 1;
    L.Nn = (long long)L.Lx * L.Ly * L.Lz;
    L.Ne = (long long)L.nx * L.ny * L.nz;
    s->lv.push_back(L);
  }
  for (auto& L : s->lv) {
    const int zc = zchunk(s, L);
    dim3 g = tile_grid(s, L, zc);
    maxpart = std::max<long long>(maxpart, (long long)g.x * g.y * g.z);
  }
  s->n_partial = (int)(std::max<long long>(maxpart, EW_BLOCKS) * s->opt.max_batch);
  int rc = VX_OK;
  for (auto& L : s->lv) {
    if ((rc = dalloc(s, &L.mask, L.Nn)) != VX_OK) break;
    if ((rc = dalloc(s, &L.mask_fixed, L.Nn)) != VX_OK) break;
    if ((rc = dalloc(s, &L.mi64, L.Nn)) != VX_OK) break;
    if ((rc = dalloc(s, &L.mi32, L.Nn)) != VX_OK) break;
    if ((rc = dalloc(s, &L.E64, L.Ne)) != VX_OK) break;
    if ((rc = dalloc(s, &L.E32, L.Ne)) != VX_OK) break;
  }
  if (rc == VX_OK) rc = dalloc(s, &s->partial, (size_t)s->n_partial);
  if (rc == VX_OK) rc = dalloc(s, &s->dscal, (size_t)s->opt.max_batch);
  if (rc == VX_OK) rc = dalloc(s, &s->dflags, (size_t)s->opt.max_batch);
  if (rc == VX_OK &&
      hipHostMalloc((void**)&s->hscal, sizeof(Scalars) * s->opt.max_batch) != hipSuccess)
    rc = VX_ERR_NOMEM;
  if (rc == VX_OK &&
      hipHostMalloc((void**)&s->hflags, sizeof(int) * s->opt.max_batch) != hipSuccess)
    rc = VX_ERR_NOMEM;
  if (rc == VX_OK) {
    int dev = 0, ncu = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
        ncu > 0)
      s->n_cu = ncu;
    (void)hipGetLastError();
    // the blocked smoother's LDS exceeds the default dynamic limit
    if (hipFuncSetAttribute(reinterpret_cast<const void*>(&k_smooth_block),
                            hipFuncAttributeMaxDynamicSharedMemorySize,
                            BLK_LDS_BYTES) != hipSuccess) {
      (void)hipGetLastError();
      rc = VX_ERR_HIP;
    }
  }
  if (rc != VX_OK) {
    vx_destroy(s);
    return rc;
  }
  *out = s;
  return VX_OK;
}

void vx_destroy(vx_handle s) {
  if (!s) return;
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  else (void)hipDeviceSynchronize();
  drop_graphs(s);
  if (s->cap_stream) (void)hipStreamDestroy(s->cap_stream);
  for (void* p : s->allocs) (void)hipFree(p);
  if (s->hscal) (void)hipHostFree(s->hscal);
  if (s->hflags) (void)hipHostFree(s->hflags);
  delete s;
}

int vx_set_stream(vx_handle s, void* stream) {
  if (!s) return VX_ERR_ARG;
  s->stream = (hipStream_t)stream;
  return VX_OK;
}

int vx_num_levels(vx_handle s) { return s ? (int)s->lv.size() : VX_ERR_ARG; }

int vx_level_dims(vx_handle s, int32_t l, int32_t* nx, int32_t* ny, int32_t* nz) {
  if (!s || l < 0 || l >= (int)s->lv.size()) return VX_ERR_ARG;
  *nx = s->lv[l].nx; *ny = s->lv[l].ny; *nz = s->lv[l].nz;
  return VX_OK;
}

int vx_set_fixed(vx_handle s, const int64_t* fixed, int64_t n_fixed) {
  if (!s || (n_fixed > 0 && !fixed)) return fail(s, VX_ERR_ARG, "bad fixed-DOF arguments");
  Level& L0 = s->lv[0];
  std::vector<uint8_t> m(L0.Nn, 7);
  for (int64_t i = 0; i < n_fixed; ++i) {
    const int64_t d = fixed[i];
    if (d < 0 || d >= 3 * L0.Nn) return fail(s, VX_ERR_ARG, "fixed DOF id out of range");
    m[d / 3] &= (uint8_t)~(1u << (d % 3));
  }
  HIPCHK(hipMemcpyAsync(L0.mask_fixed, m.data(), L0.Nn, hipMemcpyHostToDevice, s->stream));
  for (size_t li = 1; li < s->lv.size(); ++li) {
    Level& C = s->lv[li];
    hipLaunchKernelGGL(k_inject_mask, dim3(ew_grid(C.Nn)), dim3(EW_THREADS), 0, s->stream, C.Lx,
                       C.Ly, C.Lz, s->lv[li - 1].mask_fixed, C.mask_fixed);
    LAUNCHCHK();
  }
  for (auto& L : s->lv)
    HIPCHK(hipMemcpyAsync(L.mask, L.mask_fixed, L.Nn, hipMemcpyDeviceToDevice, s->stream));
  s->density_set = false;   // the effective masks are rebuilt by set_density
  s->cd_ok = false;
  drop_graphs(s);           // captured with the previous coarsest-level treatment
  HIPCHK(hipStreamSynchronize(s->stream));
  s->fixed_set = true;
  return VX_OK;
}

int vx_set_density(vx_handle s, const double* E, const uint8_t* active, double e_min) {
  if (!s || !E) return fail(s, VX_ERR_ARG, "null density");
  if (!s->fixed_set) return fail(s, VX_ERR_ARG, "call vx_set_fixed() before vx_set_density()");
  Level& L0 = s->lv[0];
  // e_min < 0 selects the active-DOF system: voids carry no stiffness and
  // DOFs that no element with E > 0 touches leave the system.
  const bool active_sys = e_min < 0.0;
  s->active_system = active_sys;
  hipLaunchKernelGGL(k_import_E, dim3(ew_grid(L0.Ne)), dim3(EW_THREADS), 0, s->stream, L0.nx,
                     L0.ny, L0.nz, E, active, active_sys ? 0.0 : e_min, L0.E64, L0.E32);
  LAUNCHCHK();
  for (size_t li = 1; li < s->lv.size(); ++li) {
    Level& F = s->lv[li - 1];
    Level& C = s->lv[li];
    hipLaunchKernelGGL(k_coarsen_E<double>, dim3(ew_grid(C.Ne)), dim3(EW_THREADS), 0, s->stream,
                       C.nx, C.ny, C.nz, F.E64, C.E64, s->opt.e_coarsen);
    LAUNCHCHK();
    hipLaunchKernelGGL((k_convert<double, float>), dim3(ew_grid(C.Ne)), dim3(EW_THREADS), 0,
                       s->stream, C.Ne, C.E64, C.E32);
    LAUNCHCHK();
  }
  for (auto& L : s->lv) {
    hipLaunchKernelGGL(k_effective_mask, dim3(ew_grid(L.Nn)), dim3(EW_THREADS), 0, s->stream,
                       L.nx, L.ny, L.nz, L.E64, L.mask_fixed, active_sys ? 1 : 0, L.mask);
    LAUNCHCHK();
    const dim3 g3((L.Lx + 63) / 64, L.Ly, L.Lz);
    hipLaunchKernelGGL(k_node_minv<double>, g3, dim3(64), 0, s->stream, L.nx, L.ny, L.nz, L.E64,
                       s->kd, L.mi64);
    LAUNCHCHK();
    hipLaunchKernelGGL(k_node_minv<float>, g3, dim3(64), 0, s->stream, L.nx, L.ny, L.nz, L.E32,
                       (float)s->kd, L.mi32);
    LAUNCHCHK();
  }
  s->density_set = true;
  // The captured V-cycles hold the level buffers (unchanged here) and the
  // coarsest-level treatment: whether the direct solve runs, its size and
  // its buffers.  Only a change of those invalidates them.
  const bool cd_ok0 = s->cd_ok;
  const int cd_nf0 = s->cd_nf;
  const void* const cd_bufs0[3] = {s->cd_inv, s->cd_pos, s->cd_idx};
  s->cd_ok = false;
  if (s->opt.coarse_mode == VX_COARSE_DIRECT) CHK(build_coarse_direct(s));
  if (s->cd_ok != cd_ok0 ||
      (s->cd_ok && (s->cd_nf != cd_nf0 || s->cd_inv != cd_bufs0[0] || s->cd_pos != cd_bufs0[1] ||
                    s->cd_idx != cd_bufs0[2])))
    drop_graphs(s);
  if (s->corr) return corr_features(s, s->corr_level);
  return VX_OK;
}

}  // extern "C"

static int alloc_pcg64(vx_solver* s) {
  const size_t n = 3 * (size_t)s->lv[0].Nn * s->opt.max_batch;
  CHK(dalloc(s, &s->r64, n));
  CHK(dalloc(s, &s->z64, n));
  CHK(dalloc(s, &s->p64, n));
  CHK(dalloc(s, &s->Ap64, n));
  CHK(dalloc(s, &s->fm64, n));
  return alloc_level_vectors<double>(s);
}

static int alloc_pcg32(vx_solver* s) {
  const size_t n = 3 * (size_t)s->lv[0].Nn * s->opt.max_batch;
  CHK(dalloc(s, &s->r32, n));
  CHK(dalloc(s, &s->z32, n));
  CHK(dalloc(s, &s->p32, n));
  CHK(dalloc(s, &s->Ap32, n));
  CHK(dalloc(s, &s->d32, n));
  return alloc_level_vectors<float>(s);
}

extern "C" {

}  // extern "C"

// Solve of nb right-hand sides (vectors [nb][3 Nn]); vx_solve is nb = 1.
static int solve_impl(vx_solver* s, int nb, const double* F, double* U, const double* X0,
                      double tol, int max_iter, vx_progress_cb cb, void* user, double* hist,
                      int hist_cap, vx_info* infos) {
  if (!s || !F || !U) return fail(s, VX_ERR_ARG, "null solve argument");
  if (nb < 1 || nb > s->opt.max_batch)
    return fail(s, VX_ERR_ARG, "batch size outside [1, max_batch]");
  CHK(require_ready(s));
  if (max_iter <= 0) max_iter = 1000;   // core/multigrid.py:481
  const bool mixed = s->opt.precision == VX_MIXED;
  CHK(alloc_pcg64(s));
  if (mixed) CHK(alloc_pcg32(s));
  s->nb = nb;
  Level& L0 = s->lv[0];
  const long long n = 3 * L0.Nn;
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  HIPCHK(hipEventRecord(e0, s->stream));

  // f masked; fnorm = ||f_free|| per right-hand side (core/multigrid.py:460-471)
  hipLaunchKernelGGL((k_mask_copy<double, double>), dim3(EW_BLOCKS, nb), dim3(EW_THREADS), 0,
                     s->stream, L0.Nn, L0.mask, F, s->fm64, s->partial);
  LAUNCHCHK();
  CHK(finalize(s, EW_BLOCKS, F_FNORM));
  if (X0) {
    hipLaunchKernelGGL((k_mask_copy<double, double>), dim3(EW_BLOCKS, nb), dim3(EW_THREADS), 0,
                       s->stream, L0.Nn, L0.mask, X0, U, nullptr);
    LAUNCHCHK();
  } else {
    HIPCHK(hipMemsetAsync(U, 0, sizeof(double) * n * nb, s->stream));
  }
  CHK(sync_scalars(s));
  std::vector<double> fnorm(nb);
  std::vector<int> zero(nb, 0);
  for (int q = 0; q < nb; ++q) {
    fnorm[q] = s->hscal[q].fnorm;
    zero[q] = fnorm[q] == 0.0;
  }
  std::vector<vx_info> inf(nb);
  for (int q = 0; q < nb; ++q) {
    inf[q] = vx_info{};
    inf[q].n_levels = (int)s->lv.size();
    if (zero[q]) {   // zero load: u = 0 (core/multigrid.py:471-473)
      inf[q].converged = 1;
      inf[q].resid_ratio = inf[q].true_resid_ratio = 0.0;
      HIPCHK(hipMemsetAsync(U + q * n, 0, sizeof(double) * n, s->stream));
    }
  }
  int rc = VX_OK;
  std::vector<PcgResult> pr(nb);
  std::vector<int> stalled(nb, 0);
  // corrector: 0 none, 1 used throughout, 2 fell back to the plain V-cycle
  s->corr_on = mixed && s->corr != nullptr && s->corr_enabled;
  int corr_state = s->corr_on ? 1 : 0;
  int np = 0;
  if (!std::all_of(zero.begin(), zero.end(), [](int z) { return z; })) {
    if (!mixed) {
      // r = f - A u  (or f when u = 0)
      if (X0) rc = launch_tile<double, M_RESID, false>(s, 0, U, s->fm64, s->r64, nullptr, nullptr);
      else rc = copy_vec<double>(s, s->r64, s->fm64, n * nb);
      std::vector<double> one(nb, 1.0);
      std::vector<int> off(nb, 0);
      if (rc == VX_OK)
        rc = pcg<double>(s, U, s->r64, s->z64, s->p64, s->Ap64, L0.t64,
                         std::vector<double>(nb, tol), max_iter, true, one, tol, off, cb, user,
                         hist, hist_cap, pr, zero);
      for (int q = 0; q < nb; ++q) {
        if (zero[q]) continue;
        inf[q].iters = pr[q].iters;
        inf[q].converged = pr[q].converged ? 1 : 0;
        inf[q].resid_ratio = pr[q].ratio;
        stalled[q] = pr[q].stalled;
      }
    } else {
      // Iterative refinement: FP64 true residual, FP32 inner MG-PCG.  With
      // the learned corrector, each inner solve must cut the worst true
      // residual at least in half; otherwise (or on a non-finite inner
      // ratio) the corrector is switched off for the rest of the solve and
      // the plain V-cycle takes over.
      std::vector<int> total(nb, 0), done(zero);
      std::vector<double> true_ratio(nb, 0.0);
      // keep_direction: the previous inner solve's residual scale (true ratio)
      // and whether it left a direction
      std::vector<double> prev_scale(nb, 0.0);
      bool have_dir = false;
      double prev_worst = INFINITY;
      bool learned_step = false;
      for (int outer = 0; outer <= s->opt.max_outer && rc == VX_OK; ++outer) {
        if (outer == 0 && !X0) {
          rc = copy_vec<double>(s, s->r64, s->fm64, n * nb);
          if (rc == VX_OK) rc = dot_partials<double>(s, n, s->r64, s->r64, &np);
        } else {
          rc = launch_tile<double, M_RESID, true>(s, 0, U, s->fm64, s->r64, s->partial, &np);
        }
        if (rc == VX_OK) rc = finalize(s, np, F_SET_TMP);
        if (rc == VX_OK) rc = sync_scalars(s);
        if (rc != VX_OK) break;
        int open = 0, most = 0;
        double worst = 0.0;
        for (int q = 0; q < nb; ++q) {
          if (zero[q]) continue;
          true_ratio[q] = std::sqrt(s->hscal[q].tmp) / fnorm[q];
          inf[q].outer_iters = outer;
          if (true_ratio[q] < tol) done[q] = 1;
          if (!done[q]) {
            ++open;
            worst = std::max(worst, true_ratio[q]);
          }
          most = std::max(most, total[q]);
        }
        if (learned_step && open && !(worst < 0.5 * prev_worst)) {
          s->corr_on = false;
          corr_state = 2;
        }
        prev_worst = worst;
        if (!open || outer == s->opt.max_outer || most >= max_iter) break;
        // mark finished right-hand sides (tmp2) so their inner input is 0.  A
        // continued direction p (keep_direction, FP32 inner solves without the
        // corrector) is in the previous inner system's units: scaled by
        // old over new residual norm through beta (pscale; 0 starts afresh)
        const bool keep = s->opt.keep_direction && have_dir && !s->corr_on;
        for (int q = 0; q < nb; ++q) {
          s->hscal[q].tmp2 = done[q] ? 1.0 : 0.0;
          s->hscal[q].drift_max = s->opt.keep_drift;
          s->hscal[q].pscale = (keep && !done[q] && prev_scale[q] > 0.0 && true_ratio[q] > 0.0)
                                   ? prev_scale[q] / true_ratio[q]
                                   : 0.0;
        }
        HIPCHK(hipMemcpyAsync(s->dscal, s->hscal, sizeof(Scalars) * nb, hipMemcpyHostToDevice,
                              s->stream));
        if (keep) {
          hipLaunchKernelGGL(k_scale_to_f32_drift, dim3(EW_BLOCKS, nb), dim3(EW_THREADS), 0,
                             s->stream, n, s->dscal, s->r64, s->r32, s->partial);
          LAUNCHCHK();
          rc = finalize(s, EW_BLOCKS, F_DRIFT);
          if (rc != VX_OK) break;
        } else {
          hipLaunchKernelGGL(k_scale_to_f32, dim3(EW_BLOCKS, nb), dim3(EW_THREADS), 0, s->stream,
                             n, s->dscal, s->r64, s->r32);
          LAUNCHCHK();
        }
        HIPCHK(hipMemsetAsync(s->d32, 0, sizeof(float) * n * nb, s->stream));
        // each inner right-hand side has unit norm
        hipLaunchKernelGGL(k_set_fnorm, dim3(nb), dim3(1), 0, s->stream, s->dscal, 1.0);
        LAUNCHCHK();
        // each right-hand side's inner solve stops where its own true residual
        // would reach half the tolerance, but not beyond inner_tol
        std::vector<double> want(nb, s->opt.inner_tol);
        for (int q = 0; q < nb; ++q)
          if (!done[q]) want[q] = std::max(s->opt.inner_tol, 0.5 * tol / true_ratio[q]);
        learned_step = s->corr_on;
        int cap = max_iter - most;
        if (learned_step && s->corr_budget > 0) cap = std::min(cap, s->corr_budget);
        rc = pcg<float>(s, s->d32, s->r32, s->z32, s->p32, s->Ap32, L0.t32, want, cap,
                        false, true_ratio, tol, total, cb, user, hist, hist_cap, pr, done, 0,
                        keep);
        if (rc != VX_OK) break;
        if (learned_step && s->corr_budget > 0) {
          // over budget without reaching the inner tolerance: keep the (energy-
          // decreasing) partial update, continue with the plain V-cycle
          for (int q = 0; q < nb; ++q)
            if (!done[q] && !pr[q].converged && pr[q].iters >= cap) {
              s->corr_on = false;
              corr_state = 2;
            }
        }
        bool finite = true;
        for (int q = 0; q < nb; ++q)
          if (!done[q] && pr[q].iters > 0 && !std::isfinite(pr[q].ratio)) finite = false;
        if (!finite) {
          if (!learned_step) {
            rc = fail(s, VX_ERR_HIP, "non-finite residual in the FP32 inner solve");
            break;
          }
          // drop the inner result, retry from the same iterate without the corrector
          s->corr_on = false;
          corr_state = 2;
          learned_step = false;
          prev_worst = INFINITY;
          have_dir = false;
          continue;
        }
        // a direction is continued only from a plain inner solve, for the
        // right-hand sides that iterated to its end (one frozen earlier keeps
        // updating p with alpha = 0, which no longer is a CG direction)
        int last = 0;
        for (int q = 0; q < nb; ++q)
          if (!done[q]) last = std::max(last, pr[q].iters);
        for (int q = 0; q < nb; ++q)
          prev_scale[q] = (!done[q] && pr[q].iters > 0 && pr[q].iters == last) ? true_ratio[q]
                                                                                : 0.0;
        have_dir = !learned_step;
        bool progressed = false;
        for (int q = 0; q < nb; ++q) {
          if (done[q]) continue;
          total[q] += pr[q].iters;
          progressed = progressed || pr[q].iters > 0;
        }
        hipLaunchKernelGGL(k_add_scaled, dim3(EW_BLOCKS, nb), dim3(EW_THREADS), 0, s->stream, n,
                           s->dscal, s->d32, U);
        LAUNCHCHK();
        if (!progressed) break;
      }
      for (int q = 0; q < nb; ++q) {
        if (zero[q]) continue;
        inf[q].iters = total[q];
        inf[q].resid_ratio = inf[q].true_resid_ratio = true_ratio[q];
        inf[q].converged = true_ratio[q] < tol ? 1 : 0;
      }
    }
  }
  // final mask of u (core/multigrid.py:547) and the true FP64 residual
  if (rc == VX_OK) {
    hipLaunchKernelGGL((k_mask_copy<double, double>), dim3(EW_BLOCKS, nb), dim3(EW_THREADS), 0,
                       s->stream, L0.Nn, L0.mask, U, U, nullptr);
    HIPCHK(hipEventRecord(e1, s->stream));
    if (!mixed) {
      rc = launch_tile<double, M_RESID, true>(s, 0, U, s->fm64, s->z64, s->partial, &np);
      if (rc == VX_OK) rc = finalize(s, np, F_SET_TMP);
    }
    if (rc == VX_OK) rc = sync_scalars(s);
    if (rc == VX_OK) {
      float ms = 0.f;
      (void)hipEventElapsedTime(&ms, e0, e1);
      for (int q = 0; q < nb; ++q) {
        if (!mixed && !zero[q]) inf[q].true_resid_ratio = std::sqrt(s->hscal[q].tmp) / fnorm[q];
        inf[q].solve_ms = ms;
      }
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  for (int q = 0; q < nb; ++q) {
    inf[q].corrector = corr_state;
    inf[q].coarse_direct = s->cd_ok ? 1 : 0;
  }
  s->corr_on = false;
  if (infos)
    for (int q = 0; q < nb; ++q) infos[q] = inf[q];
  s->nb = 1;
  if (rc != VX_OK) return rc;
  if (std::any_of(stalled.begin(), stalled.end(), [](int v) { return v; })) return VX_ERR_STALL;
  for (int q = 0; q < nb; ++q)
    if (!inf[q].converged) return VX_ERR_NOT_CONVERGED;
  return VX_OK;
}

extern "C" {

int vx_solve(vx_handle s, const double* f, double* u, const double* x0, double tol,
             int32_t max_iter, vx_progress_cb cb, void* user, double* hist, int32_t hist_cap,
             vx_info* info) {
  return solve_impl(s, 1, f, u, x0, tol, max_iter, cb, user, hist, hist_cap, info);
}

int vx_solve_batch(vx_handle s, int32_t nrhs, const double* F, double* U, double tol,
                   int32_t max_iter, vx_progress_cb cb, void* user, double* hist,
                   int32_t hist_cap, vx_info* infos) {
  return solve_impl(s, nrhs, F, U, nullptr, tol, max_iter, cb, user, hist, hist_cap, infos);
}

// ---------------------------------------------------------------------------
// Jacobi-PCG fallback on the active-DOF system (core/fea.py:497-593).
//
// The reference compresses the system to the DOFs of active elements and
// drops void elements entirely (no e_min).  The same system is solved here
// on the full grid: inactive elements get E = 0 and DOFs that no active
// element touches are masked like fixed DOFs, so every dot product and
// update runs over exactly the reference's reduced DOF set.
// ---------------------------------------------------------------------------
int vx_solve_jacobi(vx_handle s, const double* E, const uint8_t* active, const double* f,
                    double* u, const double* x0, double tol, int32_t max_iter,
                    vx_progress_cb cb, void* user, double* hist, int32_t hist_cap,
                    vx_info* info) {
  if (!s || !E || !f || !u) return fail(s, VX_ERR_ARG, "null Jacobi argument");
  s->nb = 1;
  if (!s->fixed_set) return fail(s, VX_ERR_ARG, "call vx_set_fixed() before solving");
  CHK(alloc_pcg64(s));
  Level& L0 = s->lv[0];
  Level& J = s->jac;
  if (!J.mask) {
    J = L0;
    J.mask = nullptr; J.E64 = nullptr; J.E32 = nullptr;
    J.b64 = J.x64 = J.t64 = nullptr;
    J.b32 = J.x32 = J.t32 = nullptr;
    CHK(dalloc(s, &J.mask, J.Nn));
    CHK(dalloc(s, &J.E64, J.Ne));
    CHK(dalloc(s, &J.E32, J.Ne));
    CHK(dalloc(s, &s->jac_minv, J.Nn));
  }
  const long long n = 3 * J.Nn;
  hipEvent_t e0, e1, ev;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  HIPCHK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
  HIPCHK(hipEventRecord(e0, s->stream));
  hipLaunchKernelGGL(k_import_E, dim3(ew_grid(J.Ne)), dim3(EW_THREADS), 0, s->stream, J.nx, J.ny,
                     J.nz, E, active, 0.0, J.E64, J.E32);
  LAUNCHCHK();
  hipLaunchKernelGGL(k_jacobi_setup, dim3(ew_grid(J.Nn)), dim3(EW_THREADS), 0, s->stream, J.nx,
                     J.ny, J.nz, active, L0.mask_fixed, J.E64, s->kd, J.mask, s->jac_minv);
  LAUNCHCHK();
  // f masked to the reduced free set; fnorm; count of free DOFs
  hipLaunchKernelGGL((k_mask_copy<double, double>), dim3(EW_BLOCKS), dim3(EW_THREADS), 0,
                     s->stream, J.Nn, J.mask, f, s->fm64, s->partial);
  LAUNCHCHK();
  CHK(finalize(s, EW_BLOCKS, F_FNORM));
  hipLaunchKernelGGL(k_count_free, dim3(EW_BLOCKS), dim3(EW_THREADS), 0, s->stream, J.Nn, J.mask,
                     s->partial);
  LAUNCHCHK();
  CHK(finalize(s, EW_BLOCKS, F_SET_TMP));
  HIPCHK(hipMemsetAsync(u, 0, sizeof(double) * n, s->stream));
  HIPCHK(hipMemcpyAsync(s->hscal, s->dscal, sizeof(Scalars), hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  const double fnorm = s->hscal->fnorm;
  const long long n_free = (long long)llround(s->hscal->tmp);
  vx_info inf{};
  inf.n_levels = 1;
  int rc = VX_OK;
  PcgResult res;
  if (fnorm == 0.0) {
    inf.converged = 1;
  } else {
    if (max_iter <= 0) max_iter = (int)std::min<long long>(4 * n_free, 2000000000LL);
    double* r = s->r64;
    double* z = s->z64;
    double* p = s->p64;
    double* Ap = s->Ap64;
    int np = 0;
    if (x0) {
      // warm start (core/fea.py:529-535): u = x0 on the reduced free DOFs,
      // r = f - K u there (M_RESID masks K u to the same set)
      hipLaunchKernelGGL((k_mask_copy<double, double>), dim3(EW_BLOCKS), dim3(EW_THREADS), 0,
                         s->stream, J.Nn, J.mask, x0, u, nullptr);
      LAUNCHCHK();
      CHK((launch_tile_L<double, M_RESID, false>(s, J, u, s->fm64, r, nullptr, nullptr)));
    } else {
      HIPCHK(hipMemcpyAsync(r, s->fm64, sizeof(double) * n, hipMemcpyDeviceToDevice, s->stream));
    }
    hipLaunchKernelGGL(k_diag_precond, dim3(EW_BLOCKS), dim3(EW_THREADS), 0, s->stream, J.Nn,
                       J.mask, s->jac_minv, r, z, s->partial);
    LAUNCHCHK();
    CHK(finalize(s, EW_BLOCKS, F_RZ0));
    HIPCHK(hipMemcpyAsync(p, z, sizeof(double) * n, hipMemcpyDeviceToDevice, s->stream));
    for (int it = 0; it < max_iter; ++it) {
      rc = launch_tile_L<double, M_MATVEC, true>(s, J, p, nullptr, Ap, s->partial, &np);
      if (rc == VX_OK) rc = finalize(s, np, F_ALPHA);
      if (rc != VX_OK) break;
      hipLaunchKernelGGL(k_cg_update<double>, dim3(EW_BLOCKS), dim3(EW_THREADS), 0, s->stream, n,
                         s->dscal, u, r, p, Ap, s->partial);
      rc = finalize(s, EW_BLOCKS, F_RATIO);
      if (rc == VX_OK) rc = read_scalars(s, ev);
      if (rc != VX_OK) break;
      hipLaunchKernelGGL(k_diag_precond, dim3(EW_BLOCKS), dim3(EW_THREADS), 0, s->stream, J.Nn,
                         J.mask, s->jac_minv, r, z, s->partial);
      rc = finalize(s, EW_BLOCKS, F_BETA);
      if (rc != VX_OK) break;
      hipLaunchKernelGGL(k_p_update<double>, dim3(EW_BLOCKS), dim3(EW_THREADS), 0, s->stream, n,
                         s->dscal, z, p);
      if (hipEventSynchronize(ev) != hipSuccess) { rc = fail(s, VX_ERR_HIP, "event sync"); break; }
      const Scalars sc = *s->hscal;
      if (sc.pAp == 0.0) {
        res.converged = res.ratio < tol;
        break;
      }
      res.iters = it + 1;
      res.ratio = sc.ratio;
      if (hist && it < hist_cap) hist[it] = sc.ratio;
      if (cb) cb(it + 1, sc.ratio, tol, user);
      if (sc.ratio < tol) {
        res.converged = true;
        break;
      }
    }
    inf.iters = res.iters;
    inf.converged = res.converged ? 1 : 0;
    inf.resid_ratio = res.ratio;
  }
  if (rc == VX_OK) {
    hipLaunchKernelGGL((k_mask_copy<double, double>), dim3(EW_BLOCKS), dim3(EW_THREADS), 0,
                       s->stream, J.Nn, J.mask, u, u, nullptr);
    HIPCHK(hipEventRecord(e1, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    inf.solve_ms = ms;
    inf.true_resid_ratio = NAN;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipEventDestroy(ev);
  if (info) *info = inf;
  if (rc != VX_OK) return rc;
  return inf.converged ? VX_OK : VX_ERR_NOT_CONVERGED;
}

}  // extern "C"

extern "C" int vx_post_fields_impl(int nx, int ny, int nz, void* stream, const double* u,
                                   const double* vm, const uint8_t* active, double vsize,
                                   double yield, float* stress3d, float* safety3d, float* disp3d,
                                   double* out5, char* err, int errlen);

extern "C" {

// Post-solve fields of the solve worker (core/solve_worker.py:358-431).
int vx_post_fields(vx_handle s, const double* u, const double* E, const uint8_t* active,
                   double vsize, double yield_pa, float* stress3d, float* safety3d,
                   float* disp3d, double* scalars5) {
  if (!s || !u || !E || !stress3d || !safety3d || !disp3d || !scalars5 || !(vsize > 0))
    return fail(s, VX_ERR_ARG, "bad post-field arguments");
  Level& L = s->lv[0];
  double* vm = nullptr;
  HIPCHK(hipMalloc(&vm, L.Ne * sizeof(double)));
  hipLaunchKernelGGL(k_von_mises, dim3(ew_grid(L.Ne)), dim3(EW_THREADS), 0, s->stream, L.nx, L.ny,
                     L.nz, s->nu, u, E, active, vm);
  hipError_t le = hipGetLastError();
  if (le != hipSuccess) {
    (void)hipFree(vm);
    return fail(s, VX_ERR_HIP, std::string("von Mises launch: ") + hipGetErrorString(le));
  }
  char err[256] = {0};
  const int rc = vx_post_fields_impl(L.nx, L.ny, L.nz, s->stream, u, vm, active, vsize, yield_pa,
                                     stress3d, safety3d, disp3d, scalars5, err, sizeof err);
  (void)hipFree(vm);
  if (rc != VX_OK) return fail(s, rc, err);
  return VX_OK;
}

// element_strain_energy (core/fea.py:595-602): u_e^T KE u_e with unit E.
int vx_strain_energy(vx_handle s, const double* u, const uint8_t* active, double* ce) {
  if (!s || !u || !ce) return fail(s, VX_ERR_ARG, "null strain-energy argument");
  Level& L = s->lv[0];
  hipLaunchKernelGGL(k_strain_energy, dim3(ew_grid(L.Ne)), dim3(EW_THREADS), 0, s->stream, L.nx,
                     L.ny, L.nz, s->K64, u, active, ce);
  LAUNCHCHK();
  HIPCHK(hipStreamSynchronize(s->stream));
  return VX_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// single-operator entry points (float64 in/out)
// ---------------------------------------------------------------------------
static int check_level(vx_solver* s, int l) {
  if (!s) return VX_ERR_ARG;
  s->nb = 1;
  if (l < 0 || l >= (int)s->lv.size()) return fail(s, VX_ERR_ARG, "level out of range");
  return require_ready(s);
}

template <typename T>
static int to_T(vx_solver* s, const double* a, T* b, long long n) {
  hipLaunchKernelGGL((k_convert<double, T>), dim3(ew_grid(n)), dim3(EW_THREADS), 0, s->stream, n,
                     a, b);
  LAUNCHCHK();
  return VX_OK;
}
template <typename T>
static int from_T(vx_solver* s, const T* a, double* b, long long n) {
  hipLaunchKernelGGL((k_convert<T, double>), dim3(ew_grid(n)), dim3(EW_THREADS), 0, s->stream, n,
                     a, b);
  LAUNCHCHK();
  return VX_OK;
}

extern "C" {

int vx_matvec(vx_handle s, int32_t l, int32_t prec, const double* u, double* y) {
  CHK(check_level(s, l));
  Level& L = s->lv[l];
  const long long n = 3 * L.Nn;
  if (prec == VX_FP64) {
    CHK((launch_tile<double, M_MATVEC, false>(s, l, u, nullptr, y, nullptr, nullptr)));
  } else {
    CHK(alloc_level_vectors<float>(s));
    CHK(to_T<float>(s, u, L.x32, n));
    CHK((launch_tile<float, M_MATVEC, false>(s, l, L.x32, nullptr, L.t32, nullptr, nullptr)));
    CHK(from_T<float>(s, L.t32, y, n));
  }
  HIPCHK(hipStreamSynchronize(s->stream));
  return VX_OK;
}

int vx_smooth(vx_handle s, int32_t l, int32_t prec, const double* u, const double* b,
              int32_t iters, double* out) {
  CHK(check_level(s, l));
  if (iters < 1) return fail(s, VX_ERR_ARG, "iters < 1");
  Level& L = s->lv[l];
  const long long n = 3 * L.Nn;
  if (prec == VX_FP64) {
    CHK(alloc_level_vectors<double>(s));
    CHK(smooth_step<double>(s, l, u, b, iters, L.x64, L.t64, L.w64, false, nullptr));
    HIPCHK(hipMemcpyAsync(out, L.x64, sizeof(double) * n, hipMemcpyDeviceToDevice, s->stream));
  } else {
    CHK(alloc_pcg32(s));
    float* bb = s->d32;
    float* uu = s->r32;
    CHK(to_T<float>(s, b, bb, n));
    if (u) CHK(to_T<float>(s, u, uu, n));
    CHK(smooth_step<float>(s, l, u ? uu : nullptr, bb, iters, L.x32, L.t32, L.w32, false,
                           nullptr));
    CHK(from_T<float>(s, L.x32, out, n));
  }
  HIPCHK(hipStreamSynchronize(s->stream));
  return VX_OK;
}

int vx_restrict(vx_handle s, int32_t l, const double* rf, double* rc) {
  CHK(check_level(s, l));
  if (l + 1 >= (int)s->lv.size()) return fail(s, VX_ERR_ARG, "no coarser level");
  CHK(restrict_to<double>(s, s->lv[l + 1], rf, rc, 1));
  HIPCHK(hipStreamSynchronize(s->stream));
  return VX_OK;
}

int vx_prolong(vx_handle s, int32_t l, const double* ec, double* ef) {
  CHK(check_level(s, l));
  if (l + 1 >= (int)s->lv.size()) return fail(s, VX_ERR_ARG, "no coarser level");
  Level& L = s->lv[l];
  HIPCHK(hipMemsetAsync(ef, 0, sizeof(double) * 3 * L.Nn, s->stream));
  CHK(prolong_add<double>(s, l, ec, ef, 1));
  HIPCHK(hipStreamSynchronize(s->stream));
  return VX_OK;
}

int vx_vcycle(vx_handle s, int32_t prec, const double* b, double* z) {
  CHK(check_level(s, 0));
  const long long n = 3 * s->lv[0].Nn;
  if (prec == VX_FP64) {
    CHK(alloc_pcg64(s));
    HIPCHK(hipMemcpyAsync(s->r64, b, sizeof(double) * n, hipMemcpyDeviceToDevice, s->stream));
    CHK(vcycle<double>(s, s->r64, s->z64, s->lv[0].t64, s->lv[0].w64, false, nullptr, 0));
    HIPCHK(hipMemcpyAsync(z, s->z64, sizeof(double) * n, hipMemcpyDeviceToDevice, s->stream));
  } else {
    CHK(alloc_pcg32(s));
    CHK(to_T<float>(s, b, s->r32, n));
    CHK(vcycle<float>(s, s->r32, s->z32, s->lv[0].t32, s->lv[0].w32, false, nullptr, 0));
    CHK(from_T<float>(s, s->z32, z, n));
  }
  HIPCHK(hipStreamSynchronize(s->stream));
  return VX_OK;
}

int vx_inv_diag(vx_handle s, int32_t l, double* minv) {
  CHK(check_level(s, l));
  Level& L = s->lv[l];
  hipLaunchKernelGGL(k_expand3, dim3(ew_grid(L.Nn)), dim3(EW_THREADS), 0, s->stream, L.Nn,
                     L.mi64, minv);
  LAUNCHCHK();
  HIPCHK(hipStreamSynchronize(s->stream));
  return VX_OK;
}

int vx_von_mises(vx_handle s, const double* u, const double* E, const uint8_t* active,
                 double* vm) {
  if (!s || !u || !E || !vm) return fail(s, VX_ERR_ARG, "null von Mises argument");
  Level& L = s->lv[0];
  hipLaunchKernelGGL(k_von_mises, dim3(ew_grid(L.Ne)), dim3(EW_THREADS), 0, s->stream, L.nx, L.ny,
                     L.nz, s->nu, u, E, active, vm);
  LAUNCHCHK();
  HIPCHK(hipStreamSynchronize(s->stream));
  return VX_OK;
}

}  // extern "C"

namespace {
// nf[d / 3] = 1 for each fixed DOF d (already range-checked)
__global__ void k_mark_fixed_nodes(long long n, const int64_t* __restrict__ fixed,
                                   uint8_t* __restrict__ nf) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    nf[fixed[i] / 3] = 1;
}

// one dilation pass along each axis (x, y, z); *changed is set when any
// element was reached
int island_sweeps(vx_solver* s, const Level& L, const uint8_t* active, uint8_t* kept,
                  int* changed) {
  const long long lines[3] = {(long long)L.ny * L.nz, (long long)L.nx * L.nz,
                              (long long)L.nx * L.ny};
  for (int axis = 0; axis < 3; ++axis)
    hipLaunchKernelGGL(k_island_sweep, dim3(ew_grid(lines[axis])), dim3(EW_THREADS), 0, s->stream,
                       L.nx, L.ny, L.nz, axis, active, kept, changed);
  LAUNCHCHK();
  return VX_OK;
}
}  // namespace

extern "C" {

int vx_drop_islands(vx_handle s, const uint8_t* active, const int64_t* fixed, int64_t n_fixed,
                    uint8_t* kept, int64_t* n_removed) {
  if (!s || !active || !kept || (n_fixed > 0 && !fixed))
    return fail(s, VX_ERR_ARG, "bad island arguments");
  Level& L = s->lv[0];
  for (int64_t i = 0; i < n_fixed; ++i)
    if (fixed[i] < 0 || fixed[i] >= 3 * L.Nn) return fail(s, VX_ERR_ARG, "fixed DOF out of range");
  // the fixed-node flags are marked on the device from the DOF list (a host
  // array of Nn flags and its copy cost more than the fill at 256^3)
  uint8_t* dnf = nullptr;
  int64_t* dfix = nullptr;
  int* dchg = nullptr;
  unsigned long long* dcnt = nullptr;
  HIPCHK(hipMalloc(&dnf, L.Nn));
  HIPCHK(hipMalloc(&dfix, sizeof(int64_t) * std::max<int64_t>(n_fixed, 1)));
  HIPCHK(hipMalloc(&dchg, sizeof(int)));
  HIPCHK(hipMalloc(&dcnt, 2 * sizeof(unsigned long long)));
  auto done = [&]() {
    (void)hipFree(dnf);
    (void)hipFree(dfix);
    (void)hipFree(dchg);
    (void)hipFree(dcnt);
  };
  HIPCHK(hipMemsetAsync(dnf, 0, L.Nn, s->stream));
  if (n_fixed > 0) {
    HIPCHK(hipMemcpyAsync(dfix, fixed, sizeof(int64_t) * n_fixed, hipMemcpyHostToDevice, s->stream));
    hipLaunchKernelGGL(k_mark_fixed_nodes, dim3(ew_grid(n_fixed)), dim3(EW_THREADS), 0, s->stream,
                       (long long)n_fixed, dfix, dnf);
    LAUNCHCHK();
  }
  hipLaunchKernelGGL(k_island_seed, dim3(ew_grid(L.Ne)), dim3(EW_THREADS), 0, s->stream, L.nx,
                     L.ny, L.nz, active, dnf, kept);
  int hchg = 1;
  long long passes = 0;
  while (hchg) {
    HIPCHK(hipMemsetAsync(dchg, 0, sizeof(int), s->stream));
    CHK(island_sweeps(s, L, active, kept, dchg));
    HIPCHK(hipMemcpyAsync(&hchg, dchg, sizeof(int), hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    if (++passes > 4 * (L.Ne + 1)) {
      done();
      return fail(s, VX_ERR_HIP, "island flood fill did not terminate");
    }
  }
  HIPCHK(hipMemsetAsync(dcnt, 0, 2 * sizeof(unsigned long long), s->stream));
  hipLaunchKernelGGL(k_count_u8, dim3(ew_grid(L.Ne)), dim3(EW_THREADS), 0, s->stream, L.Ne, active,
                     dcnt);
  hipLaunchKernelGGL(k_count_u8, dim3(ew_grid(L.Ne)), dim3(EW_THREADS), 0, s->stream, L.Ne, kept,
                     dcnt + 1);
  unsigned long long hc[2] = {0, 0};
  HIPCHK(hipMemcpyAsync(hc, dcnt, sizeof hc, hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  if (hc[1] == 0) {
    // nothing reaches a support: leave the part untouched (core/fea.py:220-225)
    HIPCHK(hipMemcpyAsync(kept, active, L.Ne, hipMemcpyDeviceToDevice, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    if (n_removed) *n_removed = 0;
  } else if (n_removed) {
    *n_removed = (int64_t)(hc[0] - hc[1]);
  }
  done();
  return VX_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Island removal without device allocation (vx_drop_islands_noalloc).
//
// The per-call buffers of vx_drop_islands -- the per-node "touches a fixed
// DOF" flags, the changed flag of a dilation pass, the two element counts and
// an argument-error flag -- live in one file-scope device array, `scratch`,
// sized for grids of at most NOALLOC_MAX_NODES nodes.  Every GPU the library
// runs on holds its own copy until the process ends.  Host code never uses
// the array's name (a host-side stand-in): the device address is resolved
// once per device with hipGetSymbolAddress and passed to the kernels.  All
// calls on one GPU share the array and it keeps the last call's contents, so
// every call first clears what must start at zero, and calls that overlap on
// one GPU race.
// ---------------------------------------------------------------------------
constexpr long long NOALLOC_MAX_NODES = 1LL << 25;   // 321^3 nodes: grids up to 320^3
constexpr int NOALLOC_HEAD = 256;                    // changed, error, two counts
static __device__ __align__(256) uint8_t scratch[NOALLOC_HEAD + NOALLOC_MAX_NODES];

namespace {

struct NoallocHead {
  int changed;
  int error;
  unsigned long long count[2];   // active elements, kept elements
};

constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex noalloc_mu;
uint8_t* noalloc_ptr[NOALLOC_MAX_DEVICES] = {};

// device address of `scratch` on the current device (resolved on first use)
int noalloc_scratch(vx_solver* s, uint8_t** out) {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= NOALLOC_MAX_DEVICES)
    return fail(s, VX_ERR_ARG, "device index beyond the no-allocation table");
  std::lock_guard<std::mutex> lock(noalloc_mu);
  if (!noalloc_ptr[dev]) {
    void* p = nullptr;
    HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    noalloc_ptr[dev] = static_cast<uint8_t*>(p);
  }
  *out = noalloc_ptr[dev];
  return VX_OK;
}

// clear the head and the first nn node flags
__global__ void k_noalloc_clear(uint8_t* __restrict__ sc, long long nn) {
  NoallocHead* h = reinterpret_cast<NoallocHead*>(sc);
  uint8_t* flags = sc + NOALLOC_HEAD;
  const long long i0 = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i0 == 0) {
    h->changed = 0;
    h->error = 0;
    h->count[0] = h->count[1] = 0ULL;
  }
  for (long long i = i0; i < nn; i += (long long)gridDim.x * blockDim.x) flags[i] = 0;
}

// flags[d / 3] = 1 for every fixed DOF id d; an id outside [0, 3 nn) sets error
__global__ void k_noalloc_mark_fixed(long long n_fixed, const int64_t* __restrict__ fixed,
                                     long long nn, uint8_t* __restrict__ sc) {
  NoallocHead* h = reinterpret_cast<NoallocHead*>(sc);
  uint8_t* flags = sc + NOALLOC_HEAD;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n_fixed;
       i += (long long)gridDim.x * blockDim.x) {
    const long long d = fixed[i];
    if (d < 0 || d >= 3 * nn) atomicOr(&h->error, 1);
    else flags[d / 3] = 1;
  }
}

__global__ void k_noalloc_reset_changed(uint8_t* sc) {
  reinterpret_cast<NoallocHead*>(sc)->changed = 0;
}

__global__ void k_copy_u8(long long n, const uint8_t* __restrict__ a, uint8_t* __restrict__ b) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    b[i] = a[i];
}

int noalloc_read_head(vx_solver* s, const uint8_t* sc, NoallocHead* h) {
  HIPCHK(hipMemcpyAsync(h, sc, sizeof(NoallocHead), hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  return VX_OK;
}

}  // namespace

extern "C" {

int vx_drop_islands_noalloc(vx_handle s, const uint8_t* active, const int64_t* fixed_dev,
                            int64_t n_fixed, uint8_t* kept, int64_t* n_removed) {
  if (!s || !active || !kept || n_fixed < 0 || (n_fixed > 0 && !fixed_dev))
    return fail(s, VX_ERR_ARG, "bad island arguments");
  Level& L = s->lv[0];
  if (L.Nn > NOALLOC_MAX_NODES)
    return fail(s, VX_ERR_ARG, "grid of " + std::to_string(L.Nn) +
                                   " nodes exceeds the no-allocation limit of " +
                                   std::to_string(NOALLOC_MAX_NODES) + " nodes");
  uint8_t* sc = nullptr;
  CHK(noalloc_scratch(s, &sc));
  NoallocHead* dh = reinterpret_cast<NoallocHead*>(sc);
  uint8_t* flags = sc + NOALLOC_HEAD;
  hipLaunchKernelGGL(k_noalloc_clear, dim3(ew_grid(L.Nn)), dim3(EW_THREADS), 0, s->stream, sc,
                     L.Nn);
  if (n_fixed > 0)
    hipLaunchKernelGGL(k_noalloc_mark_fixed, dim3(ew_grid(n_fixed)), dim3(EW_THREADS), 0,
                       s->stream, (long long)n_fixed, fixed_dev, L.Nn, sc);
  LAUNCHCHK();
  NoallocHead h{};
  CHK(noalloc_read_head(s, sc, &h));
  if (h.error) return fail(s, VX_ERR_ARG, "fixed DOF out of range");
  hipLaunchKernelGGL(k_island_seed, dim3(ew_grid(L.Ne)), dim3(EW_THREADS), 0, s->stream, L.nx,
                     L.ny, L.nz, active, flags, kept);
  LAUNCHCHK();
  long long passes = 0;
  do {
    hipLaunchKernelGGL(k_noalloc_reset_changed, dim3(1), dim3(1), 0, s->stream, sc);
    CHK(island_sweeps(s, L, active, kept, &dh->changed));
    CHK(noalloc_read_head(s, sc, &h));
    if (++passes > 4 * (L.Ne + 1)) return fail(s, VX_ERR_HIP, "island flood fill did not terminate");
  } while (h.changed);
  hipLaunchKernelGGL(k_count_u8, dim3(ew_grid(L.Ne)), dim3(EW_THREADS), 0, s->stream, L.Ne, active,
                     &dh->count[0]);
  hipLaunchKernelGGL(k_count_u8, dim3(ew_grid(L.Ne)), dim3(EW_THREADS), 0, s->stream, L.Ne, kept,
                     &dh->count[1]);
  LAUNCHCHK();
  CHK(noalloc_read_head(s, sc, &h));
  if (h.count[1] == 0) {
    // nothing reaches a support: leave the part untouched (core/fea.py:220-225)
    hipLaunchKernelGGL(k_copy_u8, dim3(ew_grid(L.Ne)), dim3(EW_THREADS), 0, s->stream, L.Ne,
                       active, kept);
    LAUNCHCHK();
    HIPCHK(hipStreamSynchronize(s->stream));
    if (n_removed) *n_removed = 0;
  } else if (n_removed) {
    *n_removed = (int64_t)(h.count[0] - h.count[1]);
  }
  return VX_OK;
}

int vx_time_kernel(vx_handle s, int32_t kind, int32_t l, int32_t prec, int32_t reps,
                   double* avg_ms) {
  CHK(check_level(s, l));
  if (reps < 1 || !avg_ms || kind < 0 || kind > 6 || ((kind == 3 || kind == 4) && prec == VX_FP64) ||
      (kind == 4 && l + 1 >= (int)s->lv.size()))
    return fail(s, VX_ERR_ARG, "bad timing arguments");
  Level& L = s->lv[l];
  hipEvent_t a, b;
  HIPCHK(hipEventCreate(&a));
  HIPCHK(hipEventCreate(&b));
  int rc = VX_OK;
  auto run = [&](auto* x, auto* y, auto* bb, auto* up) -> int {
    using T = std::remove_pointer_t<decltype(x)>;
    SmoothCoef sc;
    sc.omega = s->opt.omega;
    sc.up = up;
    sc.cu = 0.5;
    sc.cr = 0.1;
    if (kind == 0) return launch_tile<T, M_MATVEC, false>(s, l, x, nullptr, y, nullptr, nullptr);
    if (kind == 1) return launch_tile<T, M_JACOBI, false>(s, l, x, bb, y, nullptr, nullptr, &sc);
    if (kind == 3)   // two Chebyshev steps from zero, b -> y
      return launch_cheb_z2(s, l, 2, reinterpret_cast<const float*>(bb), reinterpret_cast<float*>(y));
    if (kind == 4)   // the residual b - K x restricted into level l + 1
      return launch_resr(s, l, reinterpret_cast<const float*>(x), reinterpret_cast<const float*>(bb),
                         s->lv[l + 1].b32, nullptr, 0.0);
    if (kind == 5) {   // the last smoothing step of a pass, with its dot partials
      int np = 0;
      return launch_tile<T, M_CHEB, true>(s, l, x, bb, up, s->partial, &np, &sc);
    }
    if (kind == 6) {   // the same as a damped-Jacobi step (FP32 two-step passes)
      int np = 0;
      return launch_tile<T, M_JACOBI, true>(s, l, x, bb, y, s->partial, &np, &sc);
    }
    // as inside the smoother, the new iterate overwrites the previous one
    return launch_tile<T, M_CHEB, false>(s, l, x, bb, up, nullptr, nullptr, &sc);
  };
  if (prec == VX_FP64) {
    CHK(alloc_pcg64(s));
    double* x = l == 0 ? s->p64 : L.x64;
    double* y = l == 0 ? s->Ap64 : L.t64;
    double* bb = l == 0 ? s->r64 : L.b64;
    double* up = L.w64;
    HIPCHK(hipMemsetAsync(x, 0, sizeof(double) * 3 * L.Nn, s->stream));
    HIPCHK(hipMemsetAsync(bb, 0, sizeof(double) * 3 * L.Nn, s->stream));
    HIPCHK(hipMemsetAsync(up, 0, sizeof(double) * 3 * L.Nn, s->stream));
    for (int w = 0; w < 2 && rc == VX_OK; ++w) rc = run(x, y, bb, up);
    HIPCHK(hipEventRecord(a, s->stream));
    for (int i = 0; i < reps && rc == VX_OK; ++i) rc = run(x, y, bb, up);
    HIPCHK(hipEventRecord(b, s->stream));
  } else {
    CHK(alloc_pcg32(s));
    float* x = l == 0 ? s->p32 : L.x32;
    float* y = l == 0 ? s->Ap32 : L.t32;
    float* bb = l == 0 ? s->r32 : L.b32;
    float* up = L.w32;
    HIPCHK(hipMemsetAsync(x, 0, sizeof(float) * 3 * L.Nn, s->stream));
    HIPCHK(hipMemsetAsync(bb, 0, sizeof(float) * 3 * L.Nn, s->stream));
    HIPCHK(hipMemsetAsync(up, 0, sizeof(float) * 3 * L.Nn, s->stream));
    for (int w = 0; w < 2 && rc == VX_OK; ++w) rc = run(x, y, bb, up);
    HIPCHK(hipEventRecord(a, s->stream));
    for (int i = 0; i < reps && rc == VX_OK; ++i) rc = run(x, y, bb, up);
    HIPCHK(hipEventRecord(b, s->stream));
  }
  HIPCHK(hipEventSynchronize(b));
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, a, b));
  *avg_ms = ms / reps;
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return rc;
}

}  // extern "C"

// autotuner: candidates within this fraction of the fastest count as ties
constexpr float TUNE_TOL = 0.02f;
// a candidate's timed launches cover about this many node updates at least
// (reps grow up to 16x on the small levels, whose launches take a few us and
// would otherwise be timed over a span comparable to the timer's jitter)
constexpr long long TUNE_MIN_NODES = 2000000;

// ===========================================================================
// autotuner of the tile kernels' z-chunk
// ===========================================================================
extern "C" int vx_autotune(vx_handle s, int32_t precision, int32_t reps) {
  if (!s) return VX_ERR_ARG;
  CHK(require_ready(s));
  if (reps < 1) reps = 5;
  drop_graphs(s);   // the z-chunks change
  const int pi = precision == VX_FP64 ? 0 : 1;
  if (pi == 0) CHK(alloc_pcg64(s));
  else CHK(alloc_pcg32(s));
  hipEvent_t a, b;
  HIPCHK(hipEventCreate(&a));
  HIPCHK(hipEventCreate(&b));
  const int cands[] = {1, 2, 3, 4, 6, 8, 10, 12, 14, 16, 18, 20, 24, 29, 33, 43, 65};
  // grow the partial buffer to the largest work-group count a candidate gives
  long long need = s->n_partial;
  for (auto& L : s->lv)
    for (int c : cands) {
      const dim3 g = tile_grid(s, L, std::min(c, L.Lz));
      need = std::max(need, (long long)g.x * g.y * g.z * s->opt.max_batch);
    }
  if (need > s->n_partial) {
    s->partial = nullptr;   // the smaller buffer stays owned by s->allocs
    CHK(dalloc(s, &s->partial, (size_t)need));
    s->n_partial = (int)need;
  }
  int rc = VX_OK;
  // timed at the handle's full batch: the work-group count that the z-chunk
  // trades against scales with the right-hand sides of a launch
  s->nb = s->opt.max_batch;
  for (size_t li = 0; li < s->lv.size() && rc == VX_OK; ++li) {
    Level& L = s->lv[li];
    const long long per = (long long)L.Nn * s->nb;
    const int lreps = reps * (int)std::min<long long>(16, std::max<long long>(
                                                             1, TUNE_MIN_NODES / std::max(1LL, per)));
    for (int cls = 0; cls < 5 && rc == VX_OK; ++cls) {
      L.zc_tuned[pi][cls] = 0;
      // class 3 (the two Chebyshev steps from zero in one launch) and class 4
      // (the residual restricted in the tiles, a level with a coarser one) are FP32 only
      if (cls == 3 && (pi == 0 || s->opt.smoother != VX_SMOOTH_CHEBYSHEV)) continue;
      if (cls == 4 && (pi == 0 || li + 1 >= s->lv.size())) continue;
      std::vector<int> cs;
      cs.push_back(zchunk(s, L));   // the rule's choice competes too
      for (int c : cands)
        if (c < L.Lz && std::find(cs.begin(), cs.end(), c) == cs.end()) cs.push_back(c);
      float best = INFINITY;
      int best_zc = 0;
      std::vector<std::pair<int, float>> times;
      for (int zc : cs) {
        // the partial buffer must hold this launch's work-groups
        const dim3 g = tile_grid(s, L, zc);
        if ((long long)g.x * g.y * g.z * s->opt.max_batch > s->n_partial) continue;
        L.zc_tuned[pi][cls] = zc;
        // the variants the solve launches: class 0 the PCG operator (with the
        // p.Ap partials), 1 a smoothing step, 2 the last smoothing step of the
        // V-cycle (with its dot-product partials), 3 the first two steps of a
        // Chebyshev pass from zero (M_CHEB_Z2), 4 the residual restricted in
        // the tiles (M_RESR; its work-groups own (zc + 1) / 2 coarse planes)
        auto variant = [&](auto* x, auto* y, auto* bb, auto* w) -> int {
          using T = std::remove_pointer_t<decltype(x)>;
          SmoothCoef sc;
          sc.omega = s->opt.omega;
          sc.cu = 0.5;
          sc.cr = 0.1;
          sc.up = w;
          int np = 0;
          if (cls == 0) return launch_tile_L<T, M_MATVEC, true>(s, L, x, nullptr, y, s->partial, &np);
          // class 1 on an FP32 level with two-step passes is only the first
          // post-smoothing step (the pre-smoothing runs as class 3): time that
          if (cls == 1 && pi == 1 && s->opt.smoother == VX_SMOOTH_CHEBYSHEV &&
              level_ns(s, (int)li) == 2)
            return launch_tile<T, M_JACOBI, false>(s, (int)li, x, bb, w, nullptr, nullptr, &sc);
          if (cls == 1)
            return launch_tile<T, M_CHEB, false>(s, (int)li, x, bb, w, nullptr, nullptr, &sc);
          if (cls == 3)
            return launch_cheb_z2(s, (int)li, 2, reinterpret_cast<const float*>(bb),
                                  reinterpret_cast<float*>(w));
          if (cls == 4)
            return launch_resr(s, (int)li, reinterpret_cast<const float*>(x),
                               reinterpret_cast<const float*>(bb), s->lv[li + 1].b32, nullptr, 0.0);
          // class 2 on such a level: the last post-smoothing step is a
          // damped-Jacobi step with the dot partials
          if (pi == 1 && post_pass_jacobi(s, level_ns(s, (int)li)))
            return launch_tile<T, M_JACOBI, true>(s, (int)li, x, bb, y, s->partial, &np, &sc);
          return launch_tile<T, M_CHEB, true>(s, (int)li, x, bb, w, s->partial, &np, &sc);
        };
        auto launch = [&]() -> int {
          if (pi == 0)
            return variant(li == 0 ? s->p64 : L.x64, li == 0 ? s->Ap64 : L.t64,
                           li == 0 ? s->r64 : L.b64, L.w64);
          return variant(li == 0 ? s->p32 : L.x32, li == 0 ? s->Ap32 : L.t32,
                         li == 0 ? s->r32 : L.b32, L.w32);
        };
        rc = launch();   // warm-up
        if (rc != VX_OK) break;
        if (hipEventRecord(a, s->stream) != hipSuccess) { rc = fail(s, VX_ERR_HIP, "event"); break; }
        for (int r = 0; r < lreps && rc == VX_OK; ++r) rc = launch();
        if (rc != VX_OK) break;
        float ms = 0.f;
        if (hipEventRecord(b, s->stream) != hipSuccess || hipEventSynchronize(b) != hipSuccess ||
            hipEventElapsedTime(&ms, a, b) != hipSuccess) {
          rc = fail(s, VX_ERR_HIP, "event timing");
          break;
        }
        times.push_back({zc, ms});
        best = std::min(best, ms);
      }
      // the first candidate (the rule's choice first, then ascending) within
      // TUNE_TOL of the fastest: near-ties, where timing noise would decide,
      // resolve the same way on every run.  The z-chunk also fixes how the
      // dot-product partials are grouped, so a stable choice keeps the
      // solve's rounding, and its iteration count, the same run to run.
      for (const auto& zt : times)
        if (zt.second <= best * (1.0f + TUNE_TOL)) {
          best_zc = zt.first;
          break;
        }
      L.zc_tuned[pi][cls] = best_zc;
    }
    // FP32 smoothing passes of a small level: one tile launch per step
    // against the blocked smoother's box edges and steps per launch, each
    // timed on the level's own pass (its steps from zero)
    const int nl = (int)s->lv.size();
    const int kp = li + 1 == (size_t)nl ? s->opt.coarse_sweeps : level_ns(s, (int)li);
    if (pi == 1 && rc == VX_OK && li > 0 && L.v32a && kp >= 2 && kp <= COARSE_MAX_STEPS) {
      std::vector<std::pair<int, int>> cand{{-1, 0}};
      for (int S = 2; S <= std::min(4, kp); ++S)
        for (int B = 2; B <= 8; ++B)
          if (block_fits(L, B, S)) cand.push_back({B, S});
      std::vector<float> tms;
      float best = INFINITY;
      for (const auto& c : cand) {
        L.blk_B = c.first;
        L.blk_S = c.second;
        auto pass = [&]() {
          return smooth_step<float>(s, (int)li, nullptr, L.b32, kp, L.x32, L.t32, L.w32, false,
                                    nullptr);
        };
        rc = pass();   // warm-up
        if (rc != VX_OK) break;
        if (hipEventRecord(a, s->stream) != hipSuccess) { rc = fail(s, VX_ERR_HIP, "event"); break; }
        // passes of 20-200 us: up to 8x the reps on the smallest levels, so
        // that plans a tenth apart are told apart run to run
        for (int r = 0; r < std::min(lreps, 8 * reps) && rc == VX_OK; ++r) rc = pass();
        if (rc != VX_OK) break;
        float ms = 0.f;
        if (hipEventRecord(b, s->stream) != hipSuccess || hipEventSynchronize(b) != hipSuccess ||
            hipEventElapsedTime(&ms, a, b) != hipSuccess) {
          rc = fail(s, VX_ERR_HIP, "event timing");
          break;
        }
        tms.push_back(ms);
        best = std::min(best, ms);
      }
      L.blk_B = 0;
      L.blk_S = 0;
      // the first candidate (the tile launches, then ascending S and B) within
      // TUNE_TOL of the fastest
      for (size_t i = 0; i < tms.size() && rc == VX_OK; ++i)
        if (tms[i] <= best * (1.0f + TUNE_TOL)) {
          L.blk_B = cand[i].first;
          L.blk_S = cand[i].second;
          break;
        }
    }
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  s->nb = 1;
  return rc;
}

extern "C" int vx_tuned_zchunk(vx_handle s, int32_t level, int32_t precision, int32_t cls) {
  if (!s || level < 0 || level >= (int)s->lv.size() || cls < 0 || cls > 4) return VX_ERR_ARG;
  const Level& L = s->lv[level];
  return zchunk(s, L, precision == VX_FP64 ? 0 : 1, cls);
}

extern "C" int vx_block_plan(vx_handle s, int32_t level, int32_t steps, int32_t* edge,
                             int32_t* per_launch) {
  if (!s || level < 0 || level >= (int)s->lv.size() || !edge || !per_launch) return VX_ERR_ARG;
  *edge = 0;
  *per_launch = 0;
  const int saved = s->nb;
  s->nb = 1;
  if (block_eligible(s, level, steps, false)) {
    const BlockPlan p = block_plan(s, s->lv[level], steps);
    *edge = p.B;
    *per_launch = p.S;
  }
  s->nb = saved;
  return VX_OK;
}

// ===========================================================================
// learned coarse-grid corrector: attach, capture, targets, training
// ===========================================================================
extern "C" {

int vx_set_corrector(vx_handle s, vxn_handle net, int32_t level, float gain, int32_t enabled,
                     int32_t iter_budget) {
  if (!s) return VX_ERR_ARG;
  if (!net) {
    s->corr = nullptr;
    s->corr_level = -1;
    s->corr_enabled = false;
    return VX_OK;
  }
  if (s->opt.precision != VX_MIXED)
    return fail(s, VX_ERR_ARG, "the corrector runs inside the FP32 V-cycle: precision VX_MIXED");
  if (level < 0 || level >= (int)s->lv.size()) return fail(s, VX_ERR_ARG, "bad corrector level");
  CHK(corr_buffers(s, level));
  const Level& L = s->lv[level];
  s->corr = net;
  s->corr_level = level;
  s->corr_gain = gain > 0.f ? gain : (float)std::max(L.Lx, std::max(L.Ly, L.Lz));
  s->corr_enabled = enabled != 0;
  s->corr_budget = std::max(0, (int)iter_budget);
  if (s->density_set) CHK(corr_features(s, level));
  return VX_OK;
}

int vx_corrector_eval(vx_handle s, int32_t nsamp, const float* T, const float* V,
                      double* gain_host) {
  if (!s || !T || !V || !gain_host || nsamp < 1)
    return fail(s, VX_ERR_ARG, "bad corrector_eval arguments");
  if (!s->corr) return fail(s, VX_ERR_ARG, "attach a corrector first (vx_set_corrector)");
  CHK(require_ready(s));
  CHK(alloc_pcg32(s));
  Level& L = s->lv[s->corr_level];
  const long long n3 = 3 * L.Nn;
  s->nb = 1;
  std::vector<double> m(5 * EW_BLOCKS);
  for (int i = 0; i < nsamp; ++i) {
    const float* t = T + (long long)i * n3;
    const float* v = V + (long long)i * n3;
    int np = 0;
    CHK(dot_partials<float>(s, n3, t, t, &np));
    CHK(finalize(s, np, F_SET_CS));
    hipLaunchKernelGGL(k_net_rms, dim3(1), dim3(64), 0, s->stream, s->dscal, 1, (double)n3,
                       s->corr_rms);
    LAUNCHCHK();
    CHK(corrector_net(s, s->corr_level, t, v, s->corr_y, s->corr_rms));
    CHK((launch_tile_L<float, M_MATVEC, false>(s, L, v, nullptr, s->corr_Av, nullptr, nullptr)));
    CHK((launch_tile_L<float, M_MATVEC, false>(s, L, s->corr_y, nullptr, s->corr_Ay, nullptr,
                                               nullptr)));
    hipLaunchKernelGGL(k_dot5, dim3(EW_BLOCKS, 1), dim3(EW_THREADS), 0, s->stream, n3, v,
                       s->corr_y, s->corr_Av, s->corr_Ay, t, s->corr_part);
    LAUNCHCHK();
    HIPCHK(hipMemcpyAsync(m.data(), s->corr_part, sizeof(double) * 5 * EW_BLOCKS,
                          hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    double d[5] = {0, 0, 0, 0, 0};
    for (int k = 0; k < 5; ++k)
      for (int j = 0; j < EW_BLOCKS; ++j) d[k] += m[k * EW_BLOCKS + j];
    const double a = d[0], b = d[1], c = d[2], vt = d[3], yt = d[4];
    // energy error of c: e.Ae - 2 c.t + c.Ac; gain of the best alpha v + beta y over v
    double al = a > 0 ? vt / a : 0.0, be = 0.0;
    const double det = a * c - b * b;
    if (a > 0 && c > 0 && det > 1e-10 * a * c) {
      al = (c * vt - b * yt) / det;
      be = (a * yt - b * vt) / det;
    }
    const double e_v = -2.0 * vt + a;
    const double e_c = -2.0 * (al * vt + be * yt) + al * al * a + 2 * al * be * b + be * be * c;
    gain_host[i] = vt > 0 ? (e_v - e_c) / vt : 0.0;
  }
  return VX_OK;
}

int vx_capture(vx_handle s, int32_t level, float* buf, int32_t cap, int32_t every) {
  if (!s) return VX_ERR_ARG;
  if (!buf || cap <= 0) {
    s->cap_buf = nullptr;
    s->cap_cap = 0;
    return VX_OK;
  }
  const int nl = (int)s->lv.size();
  if (level < 0 || level >= nl || (level == nl - 1 && level != s->corr_level))
    return fail(s, VX_ERR_ARG, "capture level must be one whose residual the V-cycle forms");
  s->cap_buf = buf;
  s->cap_level = level;
  s->cap_cap = cap;
  s->cap_every = std::max(1, (int)every);
  s->cap_count = 0;
  s->cap_calls = 0;
  return VX_OK;
}

int vx_capture_count(vx_handle s) { return s ? s->cap_count : VX_ERR_ARG; }

int vx_corrector_prepare(vx_handle s, int32_t nsamp, const float* T, float* V, float* R,
                         double* vt) {
  if (!s || !T || !V || !R || !vt || nsamp < 1)
    return fail(s, VX_ERR_ARG, "bad corrector_prepare arguments");
  if (!s->corr) return fail(s, VX_ERR_ARG, "attach a corrector first (vx_set_corrector)");
  CHK(require_ready(s));
  CHK(alloc_pcg32(s));
  const int l = s->corr_level;
  Level& L = s->lv[l];
  const long long n3 = 3 * L.Nn;
  const bool was_on = s->corr_on;
  s->corr_on = false;
  s->nb = 1;
  int rc = VX_OK;
  for (int i = 0; i < nsamp && rc == VX_OK; ++i) {
    const float* t = T + (long long)i * n3;
    float* v = V + (long long)i * n3;
    float* r = R + (long long)i * n3;
    rc = plain_coarse<float>(s, l, t, v, s->corr_Av, s->corr_Ay);
    if (rc == VX_OK)
      rc = launch_tile_L<float, M_MATVEC, false>(s, L, v, nullptr, s->corr_Av, nullptr, nullptr);
    if (rc != VX_OK) break;
    hipLaunchKernelGGL(k_sub, dim3(ew_grid(n3)), dim3(EW_THREADS), 0, s->stream, n3, t,
                       s->corr_Av, r);
    if (hipGetLastError() != hipSuccess) { rc = fail(s, VX_ERR_HIP, "k_sub"); break; }
    int np = 0;
    rc = dot_partials<float>(s, n3, v, t, &np);
    if (rc == VX_OK) rc = finalize(s, np, F_SET_CS);
    if (rc == VX_OK &&
        hipMemcpyAsync(vt + i, reinterpret_cast<char*>(s->dscal) + offsetof(Scalars, cs),
                       sizeof(double), hipMemcpyDeviceToDevice, s->stream) != hipSuccess)
      rc = fail(s, VX_ERR_HIP, "v.t copy");
  }
  s->corr_on = was_on;
  if (rc != VX_OK) return rc;
  HIPCHK(hipStreamSynchronize(s->stream));
  return VX_OK;
}

int vx_corrector_train(vx_handle s, int32_t nsamp, const float* T, const float* V, const float* R,
                       const double* vt, int32_t batch, int32_t steps, float lr, uint64_t seed,
                       float* loss_host) {
  if (!s || !T || !V || !R || !vt || nsamp < 1 || batch < 1 || steps < 1)
    return fail(s, VX_ERR_ARG, "bad corrector_train arguments");
  if (!s->corr) return fail(s, VX_ERR_ARG, "attach a corrector first (vx_set_corrector)");
  if (!s->density_set) return fail(s, VX_ERR_ARG, "call vx_set_density() before training");
  Level& L = s->lv[s->corr_level];
  const long long Nn = L.Nn, n3 = 3 * Nn;
  const long long np = vxnet_api::num_params(s->corr);
  // training buffers live for this call only
  std::vector<void*> mine;
  bool ok = true;
  auto get = [&](size_t bytes) -> void* {
    void* p = nullptr;
    if (hipMalloc(&p, std::max<size_t>(bytes, 16)) != hipSuccess) {
      ok = false;
      return nullptr;
    }
    mine.push_back(p);
    return p;
  };
  float* tt = (float*)get(sizeof(float) * batch * n3);
  float* tv = (float*)get(sizeof(float) * batch * n3);
  float* tr = (float*)get(sizeof(float) * batch * n3);
  float* ty = (float*)get(sizeof(float) * batch * n3);
  float* tAy = (float*)get(sizeof(float) * batch * n3);
  float* tAv = (float*)get(sizeof(float) * batch * n3);
  float* tgy = (float*)get(sizeof(float) * batch * n3);
  float* tout = (float*)get(sizeof(float) * batch * Nn * 16);
  float* tdout = (float*)get(sizeof(float) * batch * Nn * 16);
  float* tgrad = (float*)get(sizeof(float) * np);
  float* trms = (float*)get(sizeof(float) * batch);
  double* tvt = (double*)get(sizeof(double) * batch);
  double* part = (double*)get(sizeof(double) * 5 * batch * EW_BLOCKS);
  double* tcoef = (double*)get(sizeof(double) * 2 * batch);
  double* lossd = (double*)get(sizeof(double) * steps);
  auto release = [&]() {
    (void)hipStreamSynchronize(s->stream);
    for (void* p : mine) (void)hipFree(p);
  };
  if (!ok) {
    release();
    return fail(s, VX_ERR_NOMEM, "training buffers");
  }
  int rc = VX_OK;
  uint64_t x = seed ? seed : 0x9E3779B97F4A7C15ull;
  s->nb = 1;
  const double pi = 3.14159265358979323846;
  for (int step = 1; step <= steps && rc == VX_OK; ++step) {
    for (int b = 0; b < batch && rc == VX_OK; ++b) {
      x = x * 6364136223846793005ull + 1442695040888963407ull;
      const long long i = (long long)((x >> 33) % (uint64_t)nsamp);
      if (hipMemcpyAsync(tt + b * n3, T + i * n3, sizeof(float) * n3, hipMemcpyDeviceToDevice,
                         s->stream) != hipSuccess ||
          hipMemcpyAsync(tv + b * n3, V + i * n3, sizeof(float) * n3, hipMemcpyDeviceToDevice,
                         s->stream) != hipSuccess ||
          hipMemcpyAsync(tr + b * n3, R + i * n3, sizeof(float) * n3, hipMemcpyDeviceToDevice,
                         s->stream) != hipSuccess ||
          hipMemcpyAsync(tvt + b, vt + i, sizeof(double), hipMemcpyDeviceToDevice,
                         s->stream) != hipSuccess)
        rc = fail(s, VX_ERR_HIP, "sample copy");
    }
    if (rc != VX_OK) break;
    hipLaunchKernelGGL(k_dot<float>, dim3(EW_BLOCKS, batch), dim3(EW_THREADS), 0, s->stream, n3,
                       tt, tt, part);
    hipLaunchKernelGGL(k_rms_from_partials, dim3(batch), dim3(256), 0, s->stream, part, EW_BLOCKS,
                       (double)n3, trms);
    if (hipGetLastError() != hipSuccess) { rc = fail(s, VX_ERR_HIP, "rms"); break; }
    rc = vxnet_api::prepare(s->corr, s->stream, batch, L.Lx, L.Ly, L.Lz);
    if (rc != VX_OK) { fail(s, rc, vxnet_api::error(s->corr)); break; }
    hipLaunchKernelGGL(k_net_input, dim3(ew_grid(Nn), batch), dim3(EW_THREADS), 0, s->stream, Nn,
                       tt, tv, s->corr_feat, L.mi32, trms, s->corr_gain,
                       vxnet_api::input(s->corr));
    rc = vxnet_api::forward_bf16(s->corr, tout);
    if (rc != VX_OK) { fail(s, rc, vxnet_api::error(s->corr)); break; }
    hipLaunchKernelGGL(k_net_apply, dim3(ew_grid(Nn), batch), dim3(EW_THREADS), 0, s->stream, Nn,
                       tout, L.mi32, L.mask, trms, s->corr_gain, ty, 0);
    if (hipGetLastError() != hipSuccess) { rc = fail(s, VX_ERR_HIP, "apply"); break; }
    for (int b = 0; b < batch && rc == VX_OK; ++b)
      rc = launch_tile_L<float, M_MATVEC, false>(s, L, ty + b * n3, nullptr, tAy + b * n3,
                                                 nullptr, nullptr);
    if (rc != VX_OK) break;
    // A v = t - r; the five products, the applied combination, loss, gradient
    hipLaunchKernelGGL(k_sub, dim3(ew_grid(batch * n3)), dim3(EW_THREADS), 0, s->stream,
                       batch * n3, tt, tr, tAv);
    hipLaunchKernelGGL(k_dot5, dim3(EW_BLOCKS, batch), dim3(EW_THREADS), 0, s->stream, n3, tv, ty,
                       tAv, tAy, tt, part);
    hipLaunchKernelGGL(k_combine_coef, dim3(batch), dim3(256), 0, s->stream, part, EW_BLOCKS,
                       tcoef);
    hipLaunchKernelGGL(k_train_loss, dim3(1), dim3(256), 0, s->stream, part, EW_BLOCKS, batch,
                       tcoef, lossd + (step - 1));
    hipLaunchKernelGGL(k_train_gy, dim3(ew_grid(n3), batch), dim3(EW_THREADS), 0, s->stream, n3,
                       tt, tAv, tAy, tcoef, tvt, tgy);
    hipLaunchKernelGGL(k_net_apply_bwd, dim3(ew_grid(Nn), batch), dim3(EW_THREADS), 0, s->stream,
                       Nn, tgy, L.mi32, L.mask, trms, s->corr_gain, tdout);
    if (hipGetLastError() != hipSuccess) { rc = fail(s, VX_ERR_HIP, "loss"); break; }
    rc = vxnet_api::backward(s->corr, tdout, tgrad);
    if (rc != VX_OK) { fail(s, rc, vxnet_api::error(s->corr)); break; }
    // cosine decay to 5 % of lr
    const double f = 0.05 + 0.95 * 0.5 * (1.0 + std::cos(pi * (step - 1) / steps));
    rc = vxnet_api::adam(s->corr, tgrad, (float)(lr * f), 0.9f, 0.999f, 1e-8f, step);
    if (rc != VX_OK) { fail(s, rc, vxnet_api::error(s->corr)); break; }
  }
  if (rc == VX_OK && loss_host) {
    std::vector<double> l(steps);
    if (hipMemcpyAsync(l.data(), lossd, sizeof(double) * steps, hipMemcpyDeviceToHost,
                       s->stream) != hipSuccess ||
        hipStreamSynchronize(s->stream) != hipSuccess)
      rc = fail(s, VX_ERR_HIP, "loss readback");
    else
      for (int i = 0; i < steps; ++i) loss_host[i] = (float)l[i];
  }
  release();
  return rc;
}

}  // extern "C"

#include "parts.h"
