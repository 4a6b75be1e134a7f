// HIP kernels of the hex8 MG-PCG solver (gfx950, wave64).
//
// Data layout in HBM (all levels):
//   vectors  : [Nn][3] interleaved, node x-fastest -- the reference's DOF
//              order 3*n + c (core/fea.py:129-130), so caller buffers are
//              used in place;
//   E        : per element, already multiplied by the level scale 2^l
//              (core/multigrid.py:161-162, 292), stored x-fastest
//              e = ix + nx*(iy + ny*iz) (NOT the reference's z-fastest
//              order, which is converted once per solve);
//   mask     : one byte per node, bit c set when DOF 3*n + c is free.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "elem.h"

namespace vx {

constexpr int TILE = 16;              // element columns per tile edge
constexpr int OWN = TILE - 1;         // owned node columns per tile edge
constexpr int FOOT = TILE + 1;        // node footprint per tile edge
constexpr int NT = TILE * TILE;       // threads per tile work-group
constexpr int FOOTVALS = FOOT * FOOT * 3;
constexpr int PRE = (FOOTVALS + NT - 1) / NT;

enum { M_MATVEC = 0, M_RESID = 1, M_JACOBI = 2, M_CHEB = 3, M_CHEB_Z2 = 4, M_RESR = 5 };

// PCG scalars of one right-hand side, in device memory.
struct Scalars {
  double rz;       // r.z of the current direction
  double pAp;
  double alpha;
  double rr;       // ||r||^2
  double ratio;    // sqrt(rr)/fnorm
  double rz_new;
  double beta;
  double fnorm;
  double tmp;
  double tmp2;
  double cs;       // corrector: sum of squares of its input residual
  double rz_a;     // the r.z the last step's alpha used
  double pscale;   // continued direction: (old over new) residual scale of an inner system
  double drift;    // ||true - recurrence residual|| at a refresh, in units of the new one
  double drift_max;  // continued direction only while drift <= drift_max (0: no limit)
  int frozen;      // 1: this right-hand side has stopped (alpha = 0)
  int pad;
};

template <typename T>
struct TileArgs {
  int nx, ny, nz;        // elements
  int Lx, Ly, Lz;        // nodes
  int zc;                // owned node planes per work-group
  int gx, gy, gz;        // tile grid (the launch is 1D, XCD-ordered)
  int nb;                // right-hand sides in the launch (vectors [nb][vstride])
  long long vstride;     // elements between consecutive right-hand sides
  const T* __restrict__ u;
  const T* __restrict__ E;
  const uint8_t* __restrict__ mask;
  const T* __restrict__ b;
  T* __restrict__ out;
  T omega;               // M_JACOBI weight
  const T* __restrict__ up;   // M_CHEB: previous iterate (nullptr = zero)
  T cu, cr;              // M_CHEB: out = u + cu (u - up) + cr D^-1 (b - K u)
  T kd;                  // unit-E diagonal of KE
  ElemConst<T> K;
  double* __restrict__ partial;   // one double per work-group (DOT variants)
  // M_CHEB_Z2: D^-1 per node and the weight of the first step from zero
  const T* __restrict__ minv = nullptr;
  T c1 = T(0);
  // M_RESR: the coarse level (nodes Lcx x Lcy x Lcz, zc counts its planes per
  // work-group), its right-hand side rc, mask and D^-1, and (cx1 set) its
  // first smoothing step from zero w0 D^-1 rc
  int Lcx = 0, Lcy = 0, Lcz = 0;
  long long cvstride = 0;
  T* __restrict__ rc = nullptr;
  T* __restrict__ cx1 = nullptr;
  const uint8_t* __restrict__ cmask = nullptr;
  const T* __restrict__ cminv = nullptr;
  T w0 = T(0);
};

__device__ __forceinline__ float fmaT(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fmaT(double a, double b, double c) { return __builtin_fma(a, b, c); }

// Raw buffer loads (gfx9 buffer resource, stride 0): a load whose byte
// offset is at or beyond the resource's size returns 0, so the tile kernel
// stages halo values outside the grid, and skips non-owner epilogue loads,
// with no branch and no per-load 64-bit address arithmetic.  BUF_OOB is the
// offset used for "no value" (as composable_kernel does).
constexpr int BUF_WORD3 = 0x00020000;
constexpr int BUF_OOB = (int)0x80000000u;

__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void* base, int bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, bytes, BUF_WORD3);
}
__device__ __forceinline__ float buf_ld(__amdgpu_buffer_rsrc_t r, int off, float) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
}
__device__ __forceinline__ double buf_ld(__amdgpu_buffer_rsrc_t r, int off, double) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
}
// Three consecutive values (one node's components) in one 96-bit global
// load; the backend splits a 96-bit raw buffer load into dwords.
struct V3f {
  float x, y, z;
};
__device__ __forceinline__ void ld3(const float* p, float (&v)[3]) {
  const V3f w = *reinterpret_cast<const V3f*>(p);
  v[0] = w.x;
  v[1] = w.y;
  v[2] = w.z;
}
__device__ __forceinline__ void ld3(const double* p, double (&v)[3]) {
  v[0] = p[0];
  v[1] = p[1];
  v[2] = p[2];
}
__device__ __forceinline__ unsigned buf_ld_u8(__amdgpu_buffer_rsrc_t r, int off) {
  return (unsigned)__builtin_amdgcn_raw_buffer_load_b8(r, off, 0, 0);
}

// lane i <- lane i-1 inside each row of 16 lanes (DPP row_shr:1); the first
// lane of a row receives 0.  One tile row of element columns is one DPP row.
__device__ __forceinline__ float dpp_shr1(float v) {
  return __builtin_bit_cast(
      float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x111, 0xf, 0xf, true));
}
__device__ __forceinline__ double dpp_shr1(double v) {
  const long long w = __builtin_bit_cast(long long, v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(w & 0xffffffffLL), 0x111, 0xf, 0xf, true);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(w >> 32), 0x111, 0xf, 0xf, true);
  return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned int)lo);
}

// Inverse Jacobi diagonal 1/d (1 where d == 0, core/multigrid.py:303-304):
// exact division in FP64 (the reference's arithmetic); the hardware
// reciprocal (1 ulp) in FP32, where the smoother runs inside refinement.
__device__ __forceinline__ double inv_diag(double d) { return 1.0 / (d == 0.0 ? 1.0 : d); }
__device__ __forceinline__ float inv_diag(float d) {
  return d == 0.0f ? 1.0f : __builtin_amdgcn_rcpf(d);
}

// Deterministic work-group sum of one double; thread 0 gets the result.
__device__ inline double block_sum(double v, double* sred) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 0) sred[w] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0) {
    const int nw = (blockDim.x + 63) >> 6;
    for (int i = 0; i < nw; ++i) s += sred[i];
  }
  return s;
}

// ---------------------------------------------------------------------------
// Operator tile kernel.
//
// A work-group owns a 15x15 column of nodes and a chunk of zc node planes.
// Its 256 threads map to the 16x16 element columns that touch those nodes
// (one halo column on the low side).  The kernel marches up z: each step
// stages one node plane of the 17x17 footprint in LDS (coalesced rows; the
// plane after next is prefetched into registers), every thread applies its
// element of the current layer in the Hadamard basis, the 2D modes of the
// forces of the bottom plane are completed with the top-plane half kept from
// the previous layer, and the node sums over the four element columns are
// formed by a DPP row shift in x and one LDS exchange in y (double-buffered,
// so one barrier per layer).  No atomics: every node's sum is formed in a
// fixed order, so the result is deterministic.  Work-groups are numbered so
// that each XCD (work-groups are dealt to the 8 XCDs round-robin) receives a
// contiguous run of x-neighbouring tiles, whose shared halo then stays in
// that XCD's L2.
//
// MODE  M_MATVEC: out = mask ? K u : 0                (core/multigrid.py:318-341)
//       M_RESID : out = b - (mask ? K u : 0)          (:431)
//       M_JACOBI: out = u + omega*(mask ? (b-Ku)/d : 0), d = kd*sum E
//                                                     (:367-402, :300-304)
//       M_CHEB  : out = u + mask ? cu (u - up) + cr (b-Ku)/d : 0, one step of
//                 the Chebyshev iteration on D^-1 A (three-term recurrence)
//       M_CHEB_Z2: the first two steps of a Chebyshev pass from zero in one
//                 launch.  u holds b; the first step x1 = mask c1 D^-1 b is
//                 formed pointwise from each staged value (D^-1 from minv),
//                 and the second, out = x1 + mask (cu x1 + cr (b - K x1)/d),
//                 takes b at the node from the same staged plane
//       M_RESR  : the residual b - (mask ? K u : 0) restricted at once to the
//                 coarse level (R = P^T, the coarse mask, optionally the coarse
//                 first step): tiles step by 14 fine nodes from -1, so a tile's
//                 15 owned fine columns hold the full 3-wide stencils of 7
//                 coarse columns (neighbouring tiles both compute the shared
//                 column), and work-groups own whole coarse planes; the fine
//                 residual goes through LDS, plane by plane, never to memory
// DOT   adds the work-group partial of  sum out*u (MATVEC) or sum out*b
//       (the smoothers) to partial[block]; RESID: sum out*out.
// ---------------------------------------------------------------------------
// FP32: at least 4 waves per SIMD (at most 128 registers: no spills with the
// layer loop unrolled by two; forcing 5 spills and is slower); FP64 keeps
// what its registers give (forcing 3 spills and is slower)
// The tile work of one work-group: tile vt of the grid described by a, on the
// vectors au / ab / aup / aout of one right-hand side (or part); its DOT
// partial goes to a.partial[slot].
// A staged node plane lives in LDS in its global order (row-major, the three
// components of a node together), so a thread's staging slot is its linear
// index and needs no table.  Four plane buffers rotate: layer j reads its
// bottom plane (buffer j % 4) and top plane, and stages the plane after the
// top into buffer (j + 2) % 4; the bottom plane is overwritten only by layer
// j + 2, after the barrier that ends layer j + 1, so the epilogue of layer j
// reads the owner's own node value there instead of carrying it in registers.
constexpr int NBUF = 4;
template <typename T, int MODE, bool DOT>
__device__ __forceinline__ void tile_run(const TileArgs<T>& a, int vt, int slot,
                                         const T* __restrict__ au, const T* __restrict__ ab,
                                         const T* __restrict__ aup, T* __restrict__ aout,
                                         T* __restrict__ arc = nullptr,
                                         T* __restrict__ ax1 = nullptr) {
  constexpr bool RUP = MODE == M_CHEB_Z2;
  constexpr bool RES = MODE == M_RESR;
  constexpr int RSTEP = 2 * (OWN / 2);          // M_RESR: fine nodes per tile step (14)
  constexpr int RCOL = RSTEP / 2;               // coarse columns per tile edge (7)
  constexpr int FN = FOOT * FOOT;               // footprint nodes of a plane
  constexpr int NPN = (FN + NT - 1) / NT;       // of them per thread
  __shared__ T sP[NBUF][FOOTVALS];
  __shared__ T sQ[2][4][TILE][TILE + 1];
  __shared__ double sred[NT / 64];
  // M_CHEB_Z2: c1 D^-1 and the mask byte of each staged node
  __shared__ T sW[RUP ? NBUF : 1][RUP ? FN : 1];
  __shared__ uint8_t sM[RUP ? NBUF : 1][RUP ? FN : 1];
  // M_RESR: the residual of the owned nodes of one fine plane
  __shared__ T sRes[RES ? OWN * OWN * 3 : 1];
  const int bx = vt % a.gx;
  const int byz = vt / a.gx;
  const int by = byz % a.gy;
  const int bz = byz / a.gy;

  const int tid = threadIdx.x;
  const int tx = tid & (TILE - 1), ty = tid / TILE;
  // owned nodes X0 .. X0+14 (M_RESR: from -1 in steps of 14, see above) and
  // output planes Z0 .. Z1-1 (M_RESR: fine planes 2 cz0 - 1 .. 2 (cz0 + zc) - 1
  // for the coarse planes cz0 .. cz0 + zc - 1); planes below 0 are skipped
  const int X0 = RES ? bx * RSTEP - 1 : bx * OWN, Y0 = RES ? by * RSTEP - 1 : by * OWN;
  const int Z0 = RES ? 2 * bz * a.zc - 1 : bz * a.zc;
  const int Z1 = RES ? min(2 * (bz * a.zc + a.zc), a.Lz) : min(Z0 + a.zc, a.Lz);
  const int Zlo = max(Z0, 0);
  const int ex = X0 - 1 + tx, ey = Y0 - 1 + ty;
  const bool col_in = ex >= 0 && ex < a.nx && ey >= 0 && ey < a.ny;
  const int e0 = col_in ? ey * a.nx + ex : 0;
  const int nxy = a.nx * a.ny;
  const bool owner = tx >= 1 && ty >= 1 && ex >= 0 && ey >= 0 && ex < a.Lx && ey < a.Ly;
  const int n0 = ey * a.Lx + ex;
  const int plane = a.Lx * a.Ly;
  const int plane3 = 3 * plane;

  // staging descriptors: byte offset inside a plane (BUF_OOB: outside the
  // grid) of the (up to) PRE values this thread moves per plane; value i of
  // the footprint goes to LDS slot i
  int goff[PRE];
#pragma unroll
  for (int k = 0; k < PRE; ++k) {
    const int i = tid + k * NT;
    goff[k] = BUF_OOB;
    if (i < FOOTVALS) {
      const int row = i / (FOOT * 3);
      const int col = i - row * (FOOT * 3);
      const int xn = col / 3;
      const int c = col - xn * 3;
      const int x = X0 - 1 + xn, y = Y0 - 1 + row;
      if (x >= 0 && x < a.Lx && y >= 0 && y < a.Ly)
        goff[k] = ((y * a.Lx + x) * 3 + c) * (int)sizeof(T);
    }
  }

  // one plane's buffer resource (size 0 outside the grid: every load is 0)
  auto plane_rsrc = [&](const T* v, int pz) {
    const bool zin = pz >= 0 && pz < a.Lz;
    return buf_rsrc(v + (long long)(zin ? pz : 0) * plane3, zin ? plane3 * (int)sizeof(T) : 0);
  };
  // M_CHEB_Z2: node nd = tid + k NT of the footprint, its offset in a plane
  // (BUF_OOB: outside the grid or the footprint)
  int noff[RUP ? NPN : 1];
  if (RUP) {
#pragma unroll
    for (int k = 0; k < NPN; ++k) {
      const int nd = tid + k * NT;
      const int row = nd / FOOT, xn = nd - row * FOOT;
      const int x = X0 - 1 + xn, y = Y0 - 1 + row;
      noff[k] = nd < FN && x >= 0 && x < a.Lx && y >= 0 && y < a.Ly ? y * a.Lx + x : BUF_OOB;
    }
  }
  T pre[PRE];
  // M_CHEB_Z2: D^-1 and the mask byte of each staged node, prefetched with
  // the plane
  T pw[RUP ? NPN : 1];
  unsigned pm[RUP ? NPN : 1];
  // planes above Z1 (the last layer's top) are never used: a zero-size
  // resource makes those prefetches of the last two layers cost no traffic
  auto fetch = [&](int pz) {
    const int pzu = pz <= Z1 ? pz : -1;
    const auto r = plane_rsrc(au, pzu);
#pragma unroll
    for (int k = 0; k < PRE; ++k) pre[k] = buf_ld(r, goff[k], T(0));
    if (RUP) {
      const bool zin = pzu >= 0 && pzu < a.Lz;
      const long long pb = (long long)(zin ? pzu : 0) * plane;
      const auto rw = buf_rsrc(a.minv + pb, zin ? plane * (int)sizeof(T) : 0);
      const auto rm = buf_rsrc(a.mask + pb, zin ? plane : 0);
#pragma unroll
      for (int k = 0; k < NPN; ++k) {
        pw[k] = buf_ld(rw, noff[k] == BUF_OOB ? BUF_OOB : noff[k] * (int)sizeof(T), T(0));
        pm[k] = buf_ld_u8(rm, noff[k]);
      }
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int k = 0; k < PRE; ++k)
      if (k * NT + NT <= FOOTVALS || tid + k * NT < FOOTVALS)
        sP[buf][tid + k * NT] = pre[k];
    if (RUP) {
#pragma unroll
      for (int k = 0; k < NPN; ++k)
        if (k * NT + NT <= FN || tid + k * NT < FN) {
          sW[buf][tid + k * NT] = a.c1 * pw[k];
          sM[buf][tid + k * NT] = (uint8_t)pm[k];
        }
    }
  };
  // this thread's element column: node (ty, tx) of the footprint, component c
  const int nbase = ty * FOOT + tx;
  const int cbase = 3 * nbase;
  // M_CHEB_Z2: the four corner nodes' c1 D^-1 of a plane, read once per
  // layer; the element sees the first step x1 = mask c1 D^-1 b of its
  // corners.  b is zero at every fixed DOF (the PCG residual, the masked
  // restriction and the masked operator keep it so), so x1 = c1 D^-1 b needs
  // no mask test; the node's own mask still gates its output.
  T cw[RUP ? 4 : 1];
  auto corner_scales = [&](int buf) {
    if (RUP) {
      const int nd[4] = {nbase, nbase + 1, nbase + FOOT, nbase + FOOT + 1};
#pragma unroll
      for (int q = 0; q < 4; ++q) cw[q] = sW[buf][nd[q]];
    }
  };
  auto corners = [&](int buf, int c, T& u00, T& u10, T& u01, T& u11) {
    const T* P = &sP[buf][cbase + c];
    u00 = P[0];
    u10 = P[3];
    u01 = P[3 * FOOT];
    u11 = P[3 * FOOT + 3];
    if (RUP) {
      u00 = cw[0] * u00;
      u10 = cw[1] * u10;
      u01 = cw[2] * u01;
      u11 = cw[3] * u11;
    }
  };

  T Pb[3][4], acc[3][4];
  T Eprev = T(0);
  double dacc = 0.0;
  T ra = T(0), rh = T(0);   // M_RESR: this thread's coarse value and odd-plane partial

  if (RES) {   // the slots of nodes outside the grid are never written
    for (int i = tid; i < OWN * OWN * 3; i += NT) sRes[i] = T(0);
  }
  // prologue: plane Z0-1 (bottom of the first layer) and plane Z0
  fetch(Z0 - 1);
  store(0);
  fetch(Z0);
  __syncthreads();
  corner_scales(0);
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    T u00, u10, u01, u11;
    corners(0, c, u00, u10, u01, u11);
    had2(u00, u10, u01, u11, Pb[c]);
#pragma unroll
    for (int m = 0; m < 4; ++m) acc[c][m] = T(0);
  }
  store(1);
  fetch(Z0 + 1);
  __syncthreads();

  const int J = Z1 - Z0 + 1;   // layers Z0-1 .. Z1-1
  // software pipeline: E of the next layer and the owner's b / up / mask of
  // the next output plane are loaded one step ahead
  constexpr bool NEEDB = MODE != M_MATVEC && !RUP;   // M_CHEB_Z2: b from LDS
  constexpr bool NEEDUP = MODE == M_CHEB;
  const bool has_up = NEEDUP && aup != nullptr;
  const int eoff = col_in ? e0 * (int)sizeof(T) : BUF_OOB;
  auto loadE = [&](int iz) -> T {
    const bool zin = iz >= 0 && iz < a.nz;
    const auto r = buf_rsrc(a.E + (long long)(zin ? iz : 0) * nxy, zin ? nxy * (int)sizeof(T) : 0);
    return buf_ld(r, eoff, T(0));
  };
  // epilogue operands of the owners (node n0 of plane pz).  b and up are
  // read as one 96-bit load per node; a non-owner, or a plane outside the
  // chunk, reads node 0 of plane Zlo (a valid address) and never uses it:
  // only the next layer's owners of planes Zlo .. Z1-1 do.
  const int moff = owner ? n0 : BUF_OOB;
  const int vnode = 3 * (owner ? n0 : 0);
  unsigned mk_n = 0;
  T b_n[3] = {T(0), T(0), T(0)}, up_n[3] = {T(0), T(0), T(0)};
  auto loadEpi = [&](int pz) {
    if (RUP) return;   // the node's mask, b and x1 come from LDS
    const bool zin = pz >= Zlo && pz < Z1;
    const int pzc = zin ? pz : 0;
    const auto rm = buf_rsrc(a.mask + (long long)pzc * plane, zin ? plane : 0);
    mk_n = buf_ld_u8(rm, moff);
    const long long voff = (long long)(zin ? pz : Zlo) * plane3 + vnode;
    if (NEEDB) ld3(ab + voff, b_n);
    if (NEEDUP && has_up) ld3(aup + voff, up_n);
  };
  T E_n = loadE(Z0 - 1);
  loadEpi(Z0);
  // one element layer: the bottom-plane modes and force halves come in
  // (Pb, acc), the top plane's go out (Pt, Yt).  The layer loop is unrolled
  // by two with the two state sets swapped, so that no register copies carry
  // the state from one layer to the next.
  auto layer = [&](int j, const T (&Pb)[3][4], const T (&acc)[3][4], T (&Pt)[3][4],
                   T (&Yt)[3][4]) {
    const int iz = Z0 - 1 + j;
    const int bot = j & (NBUF - 1), top = (j + 1) & (NBUF - 1);
    const T Ee = E_n;
    E_n = loadE(iz + 1 < Z1 ? iz + 1 : -1);   // the last layer is Z1 - 1
    corner_scales(top);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      T u00, u10, u01, u11;
      corners(top, c, u00, u10, u01, u11);
      had2(u00, u10, u01, u11, Pt[c]);
    }
    T Yb[3][4];
    elem_modes(Pb, Pt, Ee, a.K, Yb, Yt);
    // the diagonal D = kd * (sum of the node's <= 8 element E) of the
    // smoothing epilogues: z pair in registers, x pair by DPP, y pair by LDS
    constexpr bool NEEDD = MODE == M_JACOBI || MODE == M_CHEB || RUP;
    const T Es = Eprev + Ee;
    if (NEEDD) Eprev = Ee;
    T A0[3];
    const int q = j & 1;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      T Y[4];
#pragma unroll
      for (int m = 0; m < 4; ++m) Y[m] = acc[c][m] + Yb[c][m];
      T y00, y10, y01, y11;
      ihad2(Y, y00, y10, y01, y11);
      A0[c] = y00 + dpp_shr1(y10);
      sQ[q][c][ty][tx] = y01 + dpp_shr1(y11);
    }
    const T EX = NEEDD ? Es + dpp_shr1(Es) : T(0);
    if (NEEDD) sQ[q][3][ty][tx] = EX;
    // stage the plane after the current top; prefetch the next one
    store((j + 2) & (NBUF - 1));
    fetch(iz + 3);
    // this plane's epilogue operands (loaded one step ago)
    const unsigned mk = mk_n;
    T bb[3], upv[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      bb[c] = b_n[c];
      upv[c] = up_n[c];
    }
    if (iz >= Z0) loadEpi(iz + 1);
    __syncthreads();

    if (iz >= Zlo && owner) {
      // output plane base (uniform, 64-bit) + this node's 32-bit offset
      // (M_RESR: the node's slot in sRes)
      T* const op = RES ? &sRes[((ty - 1) * OWN + (tx - 1)) * 3]
                        : aout + (long long)iz * plane3 + 3 * n0;
      T y[3], ub[3];
      unsigned mke = mk;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        y[c] = A0[c] + sQ[q][c][ty - 1][tx];
        ub[c] = sP[bot][cbase + c];   // u at this node (the bottom plane's corner 00)
      }
      if (RUP) {
        // b at the node is the staged value (sP); the node's own u is x1,
        // the step before it 0
        mke = sM[bot][nbase];
        const T w = sW[bot][nbase];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          bb[c] = ub[c];
          ub[c] = ((mke >> c) & 1u) ? w * bb[c] : T(0);
          upv[c] = T(0);
        }
      }
      if (MODE == M_MATVEC) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const T o = ((mke >> c) & 1u) ? y[c] : T(0);
          op[c] = o;
          if (DOT) dacc = fmaT((double)o, (double)ub[c], dacc);
        }
      } else if (MODE == M_RESID || RES) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const T o = bb[c] - (((mke >> c) & 1u) ? y[c] : T(0));
          op[c] = o;
          if (DOT) dacc = fmaT((double)o, (double)o, dacc);
        }
      } else {
        const T d = a.kd * (EX + sQ[q][3][ty - 1][tx]);
        const T minv = inv_diag(d);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          // explicit fused multiply-adds and no other contractible mul+add:
          // the two unrolled copies of the layer body must round alike
          T o;
          const bool fr = (mke >> c) & 1u;
          if (MODE == M_JACOBI) {
            o = fr ? fmaT(a.omega, minv * (bb[c] - y[c]), ub[c]) : ub[c];
          } else {
            const T du = ub[c] - upv[c];
            o = fr ? ub[c] + fmaT(a.cu, du, a.cr * (minv * (bb[c] - y[c]))) : ub[c];
          }
          op[c] = o;
          if (DOT) dacc = fmaT((double)o, (double)bb[c], dacc);
        }
      }
    }
    if (RES && iz >= Zlo) {
      // restriction of this fine plane's residual into the coarse columns of
      // the tile, in the tap order of k_restrict_z: an even plane 2 jz adds
      // its weighted taps to h (the odd plane below, weight 1/2) and, if it is
      // the grid's last plane, completes coarse plane jz; an odd plane 2 jz + 1
      // completes jz with its half-weighted taps (unless it is the chunk's
      // first plane) and starts h of jz + 1
      __syncthreads();
      if (tid < RCOL * RCOL * 3) {
        const int col = tid / 3, c = tid - 3 * col;
        const int cjy = col / RCOL, cjx = col - cjy * RCOL;
        const int jx = bx * RCOL + cjx, jy = by * RCOL + cjy;
        const bool odd = iz & 1;
        const T so = odd ? T(0.5) : T(1);
        T acc_a = odd ? ra : rh;   // the chain this plane continues
        T acc_h = T(0);
        // fine node 2 jx + dx sits at slot 2 cjx + 1 + dx of sRes (0 .. 14); a
        // tap outside the grid reads a slot that no owner writes, 0 since the
        // kernel's start, so it adds 0 as its weight 0 did
        const T* const q0 = &sRes[((2 * cjy + 1) * OWN + 2 * cjx + 1) * 3 + c];
#pragma unroll
        for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
          for (int dx = -1; dx <= 1; ++dx) {
            const T wk = so * ((dx == 0 ? T(1) : T(0.5)) * (dy == 0 ? T(1) : T(0.5)));
            const T qv = q0[(dy * OWN + dx) * 3];
            acc_a = fmaT(wk, qv, acc_a);
            acc_h = fmaT(wk, qv, acc_h);
          }
        const int jz = iz >> 1;
        const bool done = odd ? iz > Z0 : iz + 1 >= a.Lz;
        if (odd) rh = acc_h;
        else ra = acc_a;
        if (done && jx < a.Lcx && jy < a.Lcy) {
          const long long J = ((long long)jz * a.Lcy + jy) * a.Lcx + jx;
          const bool fr = (a.cmask[J] >> c) & 1u;
          arc[3 * J + c] = fr ? acc_a : T(0);
          if (ax1) ax1[3 * J + c] = fr ? (a.w0 * a.cminv[J]) * acc_a : T(0);
        }
      }
    }
  };
  T Pb2[3][4], acc2[3][4];
  int j = 0;
  for (; j + 1 < J; j += 2) {
    layer(j, Pb, acc, Pb2, acc2);
    layer(j + 1, Pb2, acc2, Pb, acc);
  }
  if (j < J) layer(j, Pb, acc, Pb2, acc2);

  if (DOT) {
    const double s = block_sum(dacc, sred);
    if (tid == 0) a.partial[slot] = s;
  }
}

// Waves per SIMD a tile kernel is compiled for at least.  FP32: 4, which
// the compiler raises to 5 or 6 where 96 or 80 registers suffice; the
// pre-smoothing pass (97 registers unconstrained) is held to 5, which it
// reaches without spilling (170 against 174-178 us per level-0 launch at
// 256^3).  Holding the operator with its p.Ap partials to 6 spills its FP64
// accumulator and gains nothing.  FP64: 1.
template <typename T, int MODE, bool DOT>
constexpr int tile_waves() {
  return sizeof(T) != 4 ? 1 : MODE == M_CHEB_Z2 ? 5 : 4;
}

template <typename T, int MODE, bool DOT>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(tile_waves<T, MODE, DOT>())))
void k_tile(TileArgs<T> a) {
  // XCD-aware tile numbering over (right-hand side, tile)
  const int G = a.gx * a.gy * a.gz;
  const int GT = G * a.nb;
  const int per = (GT + 7) >> 3;
  const int v = (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
  if (v >= GT) return;
  const int rhs = v / G;
  const int vt = v - rhs * G;
  const long long voff = (long long)rhs * a.vstride;
  if (MODE == M_RESR)
    tile_run<T, MODE, DOT>(a, vt, v, a.u + voff, a.b + voff, nullptr, nullptr,
                           a.rc + rhs * a.cvstride, a.cx1 ? a.cx1 + rhs * a.cvstride : nullptr);
  else
    tile_run<T, MODE, DOT>(a, vt, v, a.u + voff, a.b ? a.b + voff : nullptr,
                           a.up ? a.up + voff : nullptr, a.out + voff);
}

// Inverse Jacobi diagonal per node of a level: 1/(kd * sum of the <= 8
// adjacent E), 1 where that is 0 (core/multigrid.py:300-304; the diagonal of
// KE is the same for all 24 DOFs).  One thread per node, 3D launch.
template <typename T>
__global__ void k_node_minv(int nx, int ny, int nz, const T* __restrict__ E, T kd,
                            T* __restrict__ minv) {
  const int ix = blockIdx.x * blockDim.x + threadIdx.x;
  const int iy = blockIdx.y, iz = blockIdx.z;
  const int Lx = nx + 1, Ly = ny + 1;
  if (ix >= Lx) return;
  T es = T(0);
#pragma unroll
  for (int dz = -1; dz <= 0; ++dz)
#pragma unroll
    for (int dy = -1; dy <= 0; ++dy)
#pragma unroll
      for (int dx = -1; dx <= 0; ++dx) {
        const int x = ix + dx, y = iy + dy, z = iz + dz;
        if (x >= 0 && x < nx && y >= 0 && y < ny && z >= 0 && z < nz)
          es += E[((long long)z * ny + y) * nx + x];
      }
  minv[((long long)iz * Ly + iy) * Lx + ix] = inv_diag(kd * es);
}

// ---------------------------------------------------------------------------
// Coarsest-level smoother in one work-group: k smoothing steps from zero
// (core/multigrid.py:428-429) on a grid small enough to keep x, the previous
// iterate, b and every element's 24 forces in LDS.  Per step, one thread per
// element applies its operator into its LDS slot, then one thread per DOF
// sums the <= 8 slots of its node in a fixed order (no atomics):
//   j = 0          : x1 = w0 * D^-1 b                      (A * 0 = 0)
//   Jacobi, j > 0  : x  = x + w0 * D^-1 (b - K x)
//   Chebyshev j > 0: x  = x + cu_j (x - x_prev) + cr_j D^-1 (b - K x)
// (masked DOFs stay 0).
// ---------------------------------------------------------------------------
// LDS budget (160 KiB): 24 forces per element + E per element, x, the previous
// x and b per DOF, D^-1 per node, the mask byte per node.
template <typename T>
constexpr int coarse_max_elems() { return sizeof(T) == 8 ? 512 : 1024; }
template <typename T>
constexpr int coarse_max_nodes() { return sizeof(T) == 8 ? 736 : 1400; }
constexpr int COARSE_MAX_STEPS = 64;
// elements and nodes per thread of the one-work-group smoother: FP32 levels
// up to 1400 nodes take two per thread; FP64 ones (at most 736 nodes) one
template <typename T>
constexpr int cs_slots() { return sizeof(T) == 8 ? 1 : 2; }

struct CoarseCoef {
  int steps;
  int cheb;
  double w0;
  double cu[COARSE_MAX_STEPS];
  double cr[COARSE_MAX_STEPS];
};

template <typename T>
__device__ __forceinline__ void coarse_smooth_run(int nx, int ny, int nz, const T* __restrict__ E,
                                                  const uint8_t* __restrict__ mask,
                                                  const T* __restrict__ minv,
                                                  const T* __restrict__ b, T* __restrict__ out,
                                                  const ElemConst<T>& K, const CoarseCoef& cc) {
  constexpr int MAXN = coarse_max_nodes<T>(), MAXE = coarse_max_elems<T>();
  __shared__ T sye[MAXE * 24];
  __shared__ T sx[3 * MAXN], sxp[3 * MAXN], sb[3 * MAXN];
  __shared__ T smi[MAXN];
  __shared__ T sE[MAXE];
  __shared__ uint8_t smk[MAXN];
  const int Lx = nx + 1, Ly = ny + 1;
  const int Nn = Lx * Ly * (nz + 1);
  const int Ne = nx * ny * nz;
  const int tid = threadIdx.x, nt = blockDim.x;
  for (int n = tid; n < Nn; n += nt) {
    smk[n] = mask[n];
    smi[n] = minv[n];
  }
  for (int e = tid; e < Ne; e += nt) sE[e] = E[e];   // read once, used every step
  for (int i = tid; i < 3 * Nn; i += nt) sb[i] = b[i];
  __syncthreads();
  for (int i = tid; i < 3 * Nn; i += nt) {
    const int n = i / 3, c = i - 3 * n;
    sx[i] = ((smk[n] >> c) & 1u) ? (T)cc.w0 * smi[n] * sb[i] : T(0);
    sxp[i] = T(0);
  }
  __syncthreads();
  constexpr int CSN = cs_slots<T>();
  if (Ne <= CSN * nt && Nn <= CSN * nt) {
    // at most CSN elements and nodes per thread: their LDS indices are
    // computed once instead of once per step.  The forces are stored
    // component-major (slot (3q + c) * MAXE + e: consecutive elements in
    // consecutive banks), and a node reads its 8 slots without branches (a
    // missing element reads slot 0 and adds 0); b, D^-1, the mask and the
    // previous iterate of a node stay in its thread's registers.
    int en[CSN][8];
    T Ee[CSN];
    int sl[CSN][8];
    unsigned okm[CSN], mk[CSN];
    T mi[CSN], bi[CSN][3], xp[CSN][3];
#pragma unroll
    for (int k = 0; k < CSN; ++k) {
      const int e = tid + k * nt;
      const int ee = e < Ne ? e : 0;
      const int ex = ee % nx, ey = (ee / nx) % ny, ez = ee / (nx * ny);
#pragma unroll
      for (int q = 0; q < 8; ++q)
        en[k][q] = 3 * (((ez + (q >> 2)) * Ly + (ey + ((q >> 1) & 1))) * Lx + (ex + (q & 1)));
      Ee[k] = sE[ee];
      const int n = tid + k * nt;
      const bool nv = n < Nn;
      const int nn = nv ? n : 0;
      const int ix = nn % Lx, iy = (nn / Lx) % Ly, iz = nn / (Lx * Ly);
      okm[k] = 0u;
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int fx = ix - (q & 1), fy = iy - ((q >> 1) & 1), fz = iz - (q >> 2);
        const bool ok = ((unsigned)fx < (unsigned)nx) & ((unsigned)fy < (unsigned)ny) &
                        ((unsigned)fz < (unsigned)nz);
        okm[k] |= ok ? (1u << q) : 0u;
        sl[k][q] = ok ? 3 * q * MAXE + (fz * ny + fy) * nx + fx : 0;
      }
      mk[k] = nv ? (unsigned)smk[nn] : 0u;
      mi[k] = smi[nn];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        bi[k][c] = sb[3 * nn + c];
        xp[k][c] = T(0);
      }
    }
    for (int st = 1; st < cc.steps; ++st) {
#pragma unroll
      for (int k = 0; k < CSN; ++k) {
        const int e = tid + k * nt;
        if (e < Ne) {
          T u[8][3], y[8][3];
#pragma unroll
          for (int q = 0; q < 8; ++q)
#pragma unroll
            for (int c = 0; c < 3; ++c) u[q][c] = sx[en[k][q] + c];
          elem_apply<T>(u, Ee[k], K, y);
#pragma unroll
          for (int q = 0; q < 8; ++q)
#pragma unroll
            for (int c = 0; c < 3; ++c) sye[(3 * q + c) * MAXE + e] = y[q][c];
        }
      }
      __syncthreads();
      const T w0 = (T)cc.w0, cu = (T)cc.cu[st], cr = (T)cc.cr[st];
#pragma unroll
      for (int k = 0; k < CSN; ++k) {
        const int n = tid + k * nt;
        if (n < Nn) {
          T v[8][3];
#pragma unroll
          for (int q = 0; q < 8; ++q)
#pragma unroll
            for (int c = 0; c < 3; ++c) v[q][c] = sye[sl[k][q] + c * MAXE];
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            T kx = T(0);
#pragma unroll
            for (int q = 0; q < 8; ++q) kx += ((okm[k] >> q) & 1u) ? v[q][c] : T(0);
            const T x = sx[3 * n + c];
            T xn = x;
            if ((mk[k] >> c) & 1u) {
              const T r = mi[k] * (bi[k][c] - kx);
              xn = cc.cheb ? x + cu * (x - xp[k][c]) + cr * r : x + w0 * r;
            }
            xp[k][c] = x;
            sx[3 * n + c] = xn;
          }
        }
      }
      __syncthreads();
    }
    for (int i = tid; i < 3 * Nn; i += nt) out[i] = sx[i];
    return;
  }
  for (int st = 1; st < cc.steps; ++st) {
    for (int e = tid; e < Ne; e += nt) {
      const int ex = e % nx, ey = (e / nx) % ny, ez = e / (nx * ny);
      T u[8][3], y[8][3];
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int n = ((ez + (q >> 2)) * Ly + (ey + ((q >> 1) & 1))) * Lx + (ex + (q & 1));
#pragma unroll
        for (int c = 0; c < 3; ++c) u[q][c] = sx[3 * n + c];
      }
      elem_apply<T>(u, sE[e], K, y);
#pragma unroll
      for (int q = 0; q < 8; ++q)
#pragma unroll
        for (int c = 0; c < 3; ++c) sye[e * 24 + 3 * q + c] = y[q][c];
    }
    __syncthreads();
    const T w0 = (T)cc.w0, cu = (T)cc.cu[st], cr = (T)cc.cr[st];
    for (int i = tid; i < 3 * Nn; i += nt) {
      const int n = i / 3, c = i - 3 * n;
      const int ix = n % Lx, iy = (n / Lx) % Ly, iz = n / (Lx * Ly);
      T kx = T(0);
#pragma unroll
      for (int q = 0; q < 8; ++q) {   // element at (node - corner offset q), corner q
        const int ex = ix - (q & 1), ey = iy - ((q >> 1) & 1), ez = iz - (q >> 2);
        if (ex >= 0 && ex < nx && ey >= 0 && ey < ny && ez >= 0 && ez < nz)
          kx += sye[((ez * ny + ey) * nx + ex) * 24 + 3 * q + c];
      }
      const T x = sx[i];
      T xn = x;
      if ((smk[n] >> c) & 1u) {
        const T r = smi[n] * (sb[i] - kx);
        xn = cc.cheb ? x + cu * (x - sxp[i]) + cr * r : x + w0 * r;
      }
      sxp[i] = x;
      sx[i] = xn;
    }
    __syncthreads();
  }
  for (int i = tid; i < 3 * Nn; i += nt) out[i] = sx[i];
}

template <typename T>
__global__ __launch_bounds__(1024) void k_coarse_smooth(int nx, int ny, int nz,
                                                        const T* __restrict__ E,
                                                        const uint8_t* __restrict__ mask,
                                                        const T* __restrict__ minv,
                                                        const T* __restrict__ b,
                                                        T* __restrict__ out, ElemConst<T> K,
                                                        CoarseCoef cc) {
  const long long off = (long long)blockIdx.x * 3 * (nx + 1) * (ny + 1) * (nz + 1);
  coarse_smooth_run<T>(nx, ny, nz, E, mask, minv, b + off, out + off, K, cc);
}

// ---------------------------------------------------------------------------
// The two bottom levels of the V-cycle in one single-work-group launch per
// right-hand side (FP32): on the second-to-last level f (held in LDS), kf
// smoothing steps from zero, the residual, the full-weighting restriction to
// the coarsest level c, kc smoothing steps from zero there, the trilinear
// prolongation back (x += mask * P e) and kf smoothing steps from that x; the
// result goes to x.  This is the sequence smooth_step / M_RESID /
// k_restrict_z / k_coarse_smooth / k_prolong_f / smooth_step launches on
// these levels (core/multigrid.py:426-436), with the same step arithmetic
// (cf: the step coefficients of level f, cc: of the coarsest level);
// only the order of the element sums into a node differs.  The levels'
// launches are latency-bound, and a variable V-cycle gives them tens of steps.
// ---------------------------------------------------------------------------
constexpr int BOT_MAXE = 640, BOT_MAXN = 900;     // level f: elements, nodes
constexpr int BOT_MAXEC = 100, BOT_MAXNC = 160;   // coarsest level
constexpr int BOT_THREADS = 1024;                 // >= every count above: one
                                                  // element and one node per thread

struct BottomArgs {
  int nx, ny, nz;                       // level f elements (even; c has half)
  const float* E;                       // level f scaled E, x-fastest
  const uint8_t* mask;
  const float* minv;
  const float* b;                       // [nb][3 Nn_f]
  float* x;                             // [nb][3 Nn_f]
  const float* Ec;
  const uint8_t* maskc;
  const float* minvc;
  int kf, kc;                           // smoothing steps on f and on c
  ElemConst<float> K;
};

__global__ __launch_bounds__(BOT_THREADS) void k_bottom(BottomArgs a, CoarseCoef cf, CoarseCoef cc) {
  __shared__ float sye[BOT_MAXE * 24];
  __shared__ float sx[3 * BOT_MAXN], sxp[3 * BOT_MAXN], sb[3 * BOT_MAXN];
  __shared__ float smi[BOT_MAXN];
  __shared__ float sE[BOT_MAXE];
  __shared__ uint8_t smk[BOT_MAXN];
  __shared__ float cx[3 * BOT_MAXNC], cxp[3 * BOT_MAXNC], cb[3 * BOT_MAXNC];
  __shared__ float cmi[BOT_MAXNC];
  __shared__ float cE[BOT_MAXEC];
  __shared__ uint8_t cmk[BOT_MAXNC];
  const int tid = threadIdx.x, nt = blockDim.x;
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const int Lx = nx + 1, Ly = ny + 1, Nn = Lx * Ly * (nz + 1), Ne = nx * ny * nz;
  const int cnx = nx / 2, cny = ny / 2, cnz = nz / 2;
  const int CLx = cnx + 1, CLy = cny + 1, Nc = CLx * CLy * (cnz + 1), Nec = cnx * cny * cnz;
  const float* b = a.b + (long long)blockIdx.x * 3 * Nn;
  float* xo = a.x + (long long)blockIdx.x * 3 * Nn;
  for (int n = tid; n < Nn; n += nt) {
    smk[n] = a.mask[n];
    smi[n] = a.minv[n];
  }
  for (int e = tid; e < Ne; e += nt) sE[e] = a.E[e];
  for (int i = tid; i < 3 * Nn; i += nt) sb[i] = b[i];
  for (int n = tid; n < Nc; n += nt) {
    cmk[n] = a.maskc[n];
    cmi[n] = a.minvc[n];
  }
  for (int e = tid; e < Nec; e += nt) cE[e] = a.Ec[e];
  __syncthreads();

  // Per-thread tables of a level in LDS (at most one element and one node
  // per thread): the element's 8 corner nodes, and the node's 8 element
  // slots in sye (-1: no element), in corner order q = x + 2y + 4z.
  struct Tab {
    int en[8];
    int sl[8];     // component-0 slot of the node's element at corner q (0 if none)
    unsigned ok;   // bit q: that element exists
  };
  auto make_tab = [&](int ex_, int ey_, int ez_) {
    Tab t;
    const int lx = ex_ + 1, ly = ey_ + 1;
    const int ne = ex_ * ey_ * ez_, nn = lx * ly * (ez_ + 1);
    const int e = tid < ne ? tid : 0;
    const int ex = e % ex_, ey = (e / ex_) % ey_, ez = e / (ex_ * ey_);
    const int n = tid < nn ? tid : 0;
    const int ix = n % lx, iy = (n / lx) % ly, iz = n / (lx * ly);
    t.ok = 0u;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      t.en[q] = ((ez + (q >> 2)) * ly + (ey + ((q >> 1) & 1))) * lx + (ex + (q & 1));
      const int fx = ix - (q & 1), fy = iy - ((q >> 1) & 1), fz = iz - (q >> 2);
      const bool ok = ((unsigned)fx < (unsigned)ex_) & ((unsigned)fy < (unsigned)ey_) &
                      ((unsigned)fz < (unsigned)ez_);
      t.ok |= ok ? (1u << q) : 0u;
      t.sl[q] = ok ? 3 * q * BOT_MAXE + (fz * ey_ + fy) * ex_ + fx : 0;
    }
    return t;
  };
  const Tab tf = make_tab(nx, ny, nz);
  const Tab tc = make_tab(cnx, cny, cnz);
  // K X: element forces into sye (one slot per element, corner, component)
  auto elem_pass = [&](const Tab& t, int ne, const float* Es, const float* X) {
    if (tid < ne) {
      float u[8][3], y[8][3];
#pragma unroll
      for (int q = 0; q < 8; ++q)
#pragma unroll
        for (int c = 0; c < 3; ++c) u[q][c] = X[3 * t.en[q] + c];
      elem_apply<float>(u, Es[tid], a.K, y);
#pragma unroll
      for (int q = 0; q < 8; ++q)
#pragma unroll
        for (int c = 0; c < 3; ++c) sye[(3 * q + c) * BOT_MAXE + tid] = y[q][c];
    }
  };
  // node n = tid: the sums of its <= 8 element slots per component, in corner
  // order (the slots are component-major, so consecutive elements sit in
  // consecutive banks; all 24 are read before the sums, a missing element
  // adds 0)
  auto node_sums = [&](const Tab& t, float (&kx)[3]) {
    float v[8][3];
#pragma unroll
    for (int q = 0; q < 8; ++q)
#pragma unroll
      for (int c = 0; c < 3; ++c) v[q][c] = sye[t.sl[q] + c * BOT_MAXE];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      kx[c] = 0.f;
#pragma unroll
      for (int q = 0; q < 8; ++q) kx[c] += ((t.ok >> q) & 1u) ? v[q][c] : 0.f;
    }
  };
  // k smoothing steps on a level in LDS; from zero (x1 = w0 D^-1 b, as
  // k_zero_step), or from X (first step x + w0 D^-1 r, as M_JACOBI); then
  // the Chebyshev steps of M_CHEB (or further Jacobi steps)
  auto smooth = [&](const CoarseCoef& q, const Tab& t, int ne, int nn, const float* Es, float* X,
                    float* XP, const float* B, const float* MI, const uint8_t* MK, int k,
                    bool from_zero) {
    const float w0 = (float)q.w0;
    int st = 0;
    if (from_zero) {
      if (tid < nn) {
        const unsigned mk = MK[tid];
        const float mi = MI[tid];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          X[3 * tid + c] = ((mk >> c) & 1u) ? w0 * mi * B[3 * tid + c] : 0.f;
          XP[3 * tid + c] = 0.f;
        }
      }
      __syncthreads();
      st = 1;
    }
    for (; st < k; ++st) {
      elem_pass(t, ne, Es, X);
      __syncthreads();
      const bool jac = !q.cheb || st == 0;
      const float cu = (float)q.cu[st], cr = (float)q.cr[st];
      if (tid < nn) {
        const unsigned mk = MK[tid];
        const float mi = MI[tid];
        float kx[3];
        node_sums(t, kx);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const int i = 3 * tid + c;
          const float x = X[i];
          float xn = x;
          if ((mk >> c) & 1u) {
            const float r = mi * (B[i] - kx[c]);
            xn = jac ? fmaT(w0, r, x) : x + fmaT(cu, x - XP[i], cr * r);
          }
          XP[i] = x;
          X[i] = xn;
        }
      }
      __syncthreads();
    }
  };

  // pre-smoothing on f
  smooth(cf, tf, Ne, Nn, sE, sx, sxp, sb, smi, smk, a.kf, true);
  // residual on f into sxp
  elem_pass(tf, Ne, sE, sx);
  __syncthreads();
  if (tid < Nn) {
    const unsigned mk = smk[tid];
    float kx[3];
    node_sums(tf, kx);
#pragma unroll
    for (int c = 0; c < 3; ++c) sxp[3 * tid + c] = sb[3 * tid + c] - (((mk >> c) & 1u) ? kx[c] : 0.f);
  }
  __syncthreads();
  // restriction: taps in (z, y, x) order, weights 1 or 1/2 per axis
  for (int i = tid; i < 3 * Nc; i += nt) {
    const int J = i / 3, c = i - 3 * J;
    const int jx = J % CLx, jy = (J / CLx) % CLy, jz = J / (CLx * CLy);
    float acc = 0.f;
    for (int dz = -1; dz <= 1; ++dz) {
      const int z = 2 * jz + dz;
      if (z < 0 || z > nz) continue;
      const float sz = dz == 0 ? 1.f : 0.5f;
      for (int dy = -1; dy <= 1; ++dy) {
        const int y = 2 * jy + dy;
        if (y < 0 || y > ny) continue;
        for (int dx = -1; dx <= 1; ++dx) {
          const int x = 2 * jx + dx;
          if (x < 0 || x > nx) continue;
          const float wk = sz * ((dx == 0 ? 1.f : 0.5f) * (dy == 0 ? 1.f : 0.5f));
          acc += wk * sxp[3 * ((z * Ly + y) * Lx + x) + c];
        }
      }
    }
    cb[i] = ((cmk[J] >> c) & 1u) ? acc : 0.f;
  }
  __syncthreads();
  // the coarsest level: kc steps from zero
  smooth(cc, tc, Nec, Nc, cE, cx, cxp, cb, cmi, cmk, a.kc, true);
  // prolongation x += mask * P e (the tap order of k_prolong_f)
  for (int i = tid; i < 3 * Nn; i += nt) {
    const int n = i / 3, c = i - 3 * n;
    if (!((smk[n] >> c) & 1u)) continue;
    const int ix = n % Lx, iy = (n / Lx) % Ly, iz = n / (Lx * Ly);
    const int ox = ix & 1, oy = iy & 1, oz = iz & 1;
    const int t00 = 3 * ((iy >> 1) * CLx + (ix >> 1)) + c;
    const int t01 = t00 + 3, t10 = t00 + 3 * CLx, t11 = t10 + 3;
    const int cplane = 3 * CLx * CLy;
    auto cont = [&](int cz, float s) {
      const float* P = cx + cz * cplane;
      s += P[t00];
      if (ox) s += P[t01];
      if (oy) {
        s += P[t10];
        if (ox) s += P[t11];
      }
      return s;
    };
    const float* P0 = cx + (iz >> 1) * cplane;
    float s = P0[t00];
    if (ox) s += P0[t01];
    if (oy) {
      s += P0[t10];
      if (ox) s += P0[t11];
    }
    const float wxy = (ox ? 0.5f : 1.f) * (oy ? 0.5f : 1.f);
    if (oz) s = cont((iz >> 1) + 1, s);
    sx[i] += s * (oz ? 0.5f * wxy : wxy);
  }
  __syncthreads();
  // post-smoothing on f from x
  smooth(cf, tf, Ne, Nn, sE, sx, sxp, sb, smi, smk, a.kf, false);
  for (int i = tid; i < 3 * Nn; i += nt) xo[i] = sx[i];
}

// ---------------------------------------------------------------------------
// Several smoothing steps of a small level in one launch (FP32).  A launch on
// a level of a few tens of nodes per edge is latency-bound: one step costs
// the dispatch and a few dependent memory round trips, not its bytes.  Here
// each work-group owns a box of B^3 nodes and loads the iterate on that box
// grown by `steps` nodes per side (clipped to the grid) into LDS.  Step j
// (0-based) updates the box grown by steps-1-j nodes: one thread per element
// of the elements touching those nodes applies the element operator into the
// element's 24 LDS slots, then one thread per node sums its <= 8 slots in
// corner order q = x + 2y + 4z (as k_bottom) and updates the node.  After the
// last step the owned box holds the result, written out together with the
// iterate before it (the next launch's previous iterate).  Every node's sum
// has the same terms in the same order whatever B and the step count, so the
// result does not depend on them; the halo work is repeated by neighbouring
// work-groups.  Step forms:
//   BLK_ZERO: x = mask ? w D^-1 b : 0                  (start from zero)
//   BLK_JAC : x = x + (mask ? w D^-1 (b - K x) : 0)
//   BLK_CHEB: x = x + (mask ? cu (x - x_prev) + cr D^-1 (b - K x) : 0)
// LDS: 3 values per region node, E per region element and 24 force slots per
// element of the first step (at most 160 KiB; the host sizes B and steps).
// ---------------------------------------------------------------------------
constexpr int BLK_THREADS = 1024;
constexpr int BLK_NPT = 2;            // region nodes per thread (region <= 2048 nodes)
constexpr int BLK_MAX_STEPS = 8;
constexpr int BLK_LDS_BYTES = 160 * 1024;
enum { BLK_ZERO = 0, BLK_JAC = 1, BLK_CHEB = 2 };

struct BlockArgs {
  int nx, ny, nz, Lx, Ly, Lz;
  int B;                      // owned box edge in nodes
  int gx, gy, gz;             // boxes per axis
  int steps;
  long long vstride;          // values per right-hand side (3 Nn)
  const float* E;             // scaled E per element
  const uint8_t* mask;        // free-DOF bits per node
  const float* minv;          // D^-1 per node
  const float* b;
  const float* x;             // iterate before the launch (nullptr: 0)
  const float* xp;            // the one before it (nullptr: 0)
  float* xo;                  // iterate after the launch
  float* xpo;                 // the one before it (nullptr: not written)
  ElemConst<float> K;
  int form[BLK_MAX_STEPS];
  float c0[BLK_MAX_STEPS];    // BLK_ZERO, BLK_JAC: w; BLK_CHEB: cu
  float c1[BLK_MAX_STEPS];    // BLK_CHEB: cr
};

// q = n / d for 0 <= n < 2^21 and d >= 1, from rd = 1/d in FP32: (n + 0.5) / d
// lies at least 0.5 / d from an integer, and the two roundings (of 1/d and of
// the product) move it by at most 1.2e-7 (n + 0.5) / d < 0.5 / d
__device__ __forceinline__ int qdiv(int n, float rd) { return (int)(((float)n + 0.5f) * rd); }

__global__ __launch_bounds__(BLK_THREADS) void k_smooth_block(BlockArgs a) {
  extern __shared__ float blk_lds[];
  const int tid = threadIdx.x, nt = blockDim.x;
  const int blk = blockIdx.x;
  const int S = a.steps, B = a.B;
  const int bq = blk / a.gx, bz = bq / a.gy;
  const int o0x = (blk - bq * a.gx) * B, o0y = (bq - bz * a.gy) * B, o0z = bz * B;
  const int o1x = min(o0x + B, a.Lx), o1y = min(o0y + B, a.Ly), o1z = min(o0z + B, a.Lz);
  const int rx0 = max(0, o0x - S), ry0 = max(0, o0y - S), rz0 = max(0, o0z - S);
  const int RX = min(a.Lx, o1x + S) - rx0, RY = min(a.Ly, o1y + S) - ry0,
            RZ = min(a.Lz, o1z + S) - rz0;
  const int NR = RX * RY * RZ;
  const int QX = RX - 1, QY = RY - 1;       // region elements per row / column
  const int NER = QX * QY * (RZ - 1);
  float* sx = blk_lds;                      // [3 NR]
  float* sE = sx + 3 * NR;                  // [NER]
  float* sye = sE + NER;                    // [24 NE] of the current step
  const long long voff = (long long)blockIdx.y * a.vstride;
  const float* b = a.b + voff;
  const float r3x = 1.f / (float)(3 * RX), rry = 1.f / (float)RY, rrx = 1.f / (float)RX;
  const float rqx = 1.f / (float)QX, rqy = 1.f / (float)QY;

  // Every global load of the set-up is issued before the first wait (one
  // memory round trip), as raw buffer loads whose out-of-range offset reads 0
  // (no branches; the level's vectors are far below 2 GiB): the iterate on
  // the region (rows of 3 RX contiguous values; 3 NR <= 3 BLK_NPT nt), E on
  // its elements (NER < NR), and b, D^-1, the mask and the previous iterate
  // of this thread's nodes, which stay in registers (only the node's own
  // thread reads them).
  const int Nn = a.Lx * a.Ly * a.Lz;
  const int vbytes = (int)(4 * a.vstride);
  const __amdgpu_buffer_rsrc_t rX = buf_rsrc(a.x ? a.x + voff : b, a.x ? vbytes : 0);
  const __amdgpu_buffer_rsrc_t rP = buf_rsrc(a.xp ? a.xp + voff : b, a.xp ? vbytes : 0);
  const __amdgpu_buffer_rsrc_t rB = buf_rsrc(b, vbytes);
  const __amdgpu_buffer_rsrc_t rE = buf_rsrc(a.E, 4 * a.nx * a.ny * a.nz);
  const __amdgpu_buffer_rsrc_t rM = buf_rsrc(a.minv, 4 * Nn);
  const __amdgpu_buffer_rsrc_t rK = buf_rsrc(a.mask, Nn);
  constexpr int XL = 3 * BLK_NPT;
  float vx[XL], ve[BLK_NPT];
#pragma unroll
  for (int k = 0; k < XL; ++k) {
    const int i = k * nt + tid;
    const int row = qdiv(i, r3x), w = i - row * 3 * RX;
    const int lz = qdiv(row, rry), ly = row - lz * RY;
    const int g3 = 3 * (((rz0 + lz) * a.Ly + (ry0 + ly)) * a.Lx + rx0) + w;
    vx[k] = buf_ld(rX, i < 3 * NR ? 4 * g3 : BUF_OOB, 0.f);
  }
#pragma unroll
  for (int k = 0; k < BLK_NPT; ++k) {
    const int e = k * nt + tid;
    const int row = qdiv(e, rqx), lx = e - row * QX;
    const int lz = qdiv(row, rqy), ly = row - lz * QY;
    const int ge = ((rz0 + lz) * a.ny + (ry0 + ly)) * a.nx + (rx0 + lx);
    ve[k] = buf_ld(rE, e < NER ? 4 * ge : BUF_OOB, 0.f);
  }
  int gxn[BLK_NPT], gyn[BLK_NPT], gzn[BLK_NPT], gn[BLK_NPT];
  float rb[BLK_NPT][3], rxp[BLK_NPT][3], rmi[BLK_NPT];
  unsigned rmk[BLK_NPT];
#pragma unroll
  for (int k = 0; k < BLK_NPT; ++k) {
    const int n = tid + k * nt;
    const bool in = n < NR;
    const int row = qdiv(n, rrx), lz = qdiv(row, rry);
    gxn[k] = in ? rx0 + n - row * RX : -1;
    gyn[k] = ry0 + row - lz * RY;
    gzn[k] = rz0 + lz;
    gn[k] = (gzn[k] * a.Ly + gyn[k]) * a.Lx + (rx0 + n - row * RX);
    rmi[k] = buf_ld(rM, in ? 4 * gn[k] : BUF_OOB, 0.f);
    rmk[k] = buf_ld_u8(rK, in ? gn[k] : BUF_OOB);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      rb[k][c] = buf_ld(rB, in ? 4 * (3 * gn[k] + c) : BUF_OOB, 0.f);
      rxp[k][c] = buf_ld(rP, in ? 4 * (3 * gn[k] + c) : BUF_OOB, 0.f);
    }
  }
#pragma unroll
  for (int k = 0; k < XL; ++k)
    if (k * nt + tid < 3 * NR) sx[k * nt + tid] = vx[k];
#pragma unroll
  for (int k = 0; k < BLK_NPT; ++k)
    if (k * nt + tid < NER) sE[k * nt + tid] = ve[k];
  __syncthreads();

  // Element slots are indexed by the element's place in the region (the same
  // in every step), so each thread's elements and each node's eight slots are
  // found once per launch: thread t applies region elements t and t + nt (those
  // inside a step's element box), a node reads slot 3 q NER + its element q.
  int el[BLK_NPT][3], en0[BLK_NPT];
#pragma unroll
  for (int k = 0; k < BLK_NPT; ++k) {
    const int e = tid + k * nt;
    const int row = qdiv(e, rqx), lz = qdiv(row, rqy);
    el[k][0] = e < NER ? e - row * QX : -1;   // -1: no element (never inside a box)
    el[k][1] = row - lz * QY;
    el[k][2] = lz;
    en0[k] = (lz * RY + el[k][1]) * RX + el[k][0];
  }
  int nsl[BLK_NPT];        // the slot of a node's element 0 (q = 0), other q at fixed offsets
  unsigned nok[BLK_NPT];   // bit q: the node's element q lies in the region
#pragma unroll
  for (int k = 0; k < BLK_NPT; ++k) {
    const int lx = gxn[k] - rx0, ly = gyn[k] - ry0, lz = gzn[k] - rz0;
    nsl[k] = (lz * QY + ly) * QX + lx;
    unsigned okb = 0;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int ex = lx - (q & 1), ey = ly - ((q >> 1) & 1), ez = lz - (q >> 2);
      const bool ok = ((unsigned)ex < (unsigned)QX) & ((unsigned)ey < (unsigned)QY) &
                      ((unsigned)ez < (unsigned)(RZ - 1));
      okb |= ok ? 1u << q : 0u;
    }
    nok[k] = okb;
  }
  const int QXY = QX * QY;
  for (int j = 0; j < S; ++j) {
    const int h = S - 1 - j;   // the step updates the owned box grown by h
    const int ux0 = max(0, o0x - h), ux1 = min(a.Lx, o1x + h);
    const int uy0 = max(0, o0y - h), uy1 = min(a.Ly, o1y + h);
    const int uz0 = max(0, o0z - h), uz1 = min(a.Lz, o1z + h);
    const int form = a.form[j];
    // elements touching the updated nodes (in region coordinates)
    const int bx0 = max(0, ux0 - 1) - rx0, by0 = max(0, uy0 - 1) - ry0, bz0 = max(0, uz0 - 1) - rz0;
    const int EX = min(a.nx, ux1) - rx0 - bx0, EY = min(a.ny, uy1) - ry0 - by0,
              EZ = min(a.nz, uz1) - rz0 - bz0;
    if (form != BLK_ZERO) {
#pragma unroll
      for (int k = 0; k < BLK_NPT; ++k) {
        if (!(((unsigned)(el[k][0] - bx0) < (unsigned)EX) & ((unsigned)(el[k][1] - by0) < (unsigned)EY) &
              ((unsigned)(el[k][2] - bz0) < (unsigned)EZ)))
          continue;
        const int e = tid + k * nt;
        const int n0 = en0[k];
        float u[8][3], y[8][3];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
          const int n = n0 + (q >> 2) * RX * RY + ((q >> 1) & 1) * RX + (q & 1);
#pragma unroll
          for (int c = 0; c < 3; ++c) u[q][c] = sx[3 * n + c];
        }
        elem_apply<float>(u, sE[e], a.K, y);
#pragma unroll
        for (int q = 0; q < 8; ++q)
#pragma unroll
          for (int c = 0; c < 3; ++c) sye[(3 * q + c) * NER + e] = y[q][c];
      }
      __syncthreads();
    }
    const float c0 = a.c0[j], c1 = a.c1[j];
#pragma unroll
    for (int k = 0; k < BLK_NPT; ++k) {
      const int X = gxn[k], Y = gyn[k], Z = gzn[k];
      if (!(((unsigned)(X - ux0) < (unsigned)(ux1 - ux0)) & ((unsigned)(Y - uy0) < (unsigned)(uy1 - uy0)) &
            ((unsigned)(Z - uz0) < (unsigned)(uz1 - uz0))))
        continue;
      const int n = tid + k * nt;
      float kx[3] = {0.f, 0.f, 0.f};
      if (form != BLK_ZERO) {
        // the node's <= 8 element slots, read without branches (a missing
        // element reads the node's q = 0 slot and adds +0)
        float v[8][3];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
          const bool ok = (nok[k] >> q) & 1u;
          const int sl = 3 * q * NER + (ok ? nsl[k] - (q & 1) - ((q >> 1) & 1) * QX - (q >> 2) * QXY : 0);
#pragma unroll
          for (int c = 0; c < 3; ++c) v[q][c] = sye[sl + c * NER];
        }
#pragma unroll
        for (int q = 0; q < 8; ++q)
#pragma unroll
          for (int c = 0; c < 3; ++c) kx[c] += ((nok[k] >> q) & 1u) ? v[q][c] : 0.f;
      }
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const float x = sx[3 * n + c];
        const bool fr = (rmk[k] >> c) & 1u;
        float xn;
        if (form == BLK_ZERO) {
          xn = fr ? c0 * rmi[k] * rb[k][c] : 0.f;
        } else {
          const float r = rmi[k] * (rb[k][c] - kx[c]);
          const float t = form == BLK_JAC ? fmaT(c0, r, x) : x + fmaT(c0, x - rxp[k][c], c1 * r);
          xn = fr ? t : x;
        }
        rxp[k][c] = x;
        sx[3 * n + c] = xn;
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int k = 0; k < BLK_NPT; ++k) {
    const int X = gxn[k], Y = gyn[k], Z = gzn[k];
    if (!(((unsigned)(X - o0x) < (unsigned)(o1x - o0x)) & ((unsigned)(Y - o0y) < (unsigned)(o1y - o0y)) &
          ((unsigned)(Z - o0z) < (unsigned)(o1z - o0z))))
      continue;
    const int n = tid + k * nt;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      a.xo[voff + 3 * gn[k] + c] = sx[3 * n + c];
      if (a.xpo) a.xpo[voff + 3 * gn[k] + c] = rxp[k][c];
    }
  }
}

// per-node values -> per-DOF (3 copies)
__global__ void k_expand3(long long Nn, const double* __restrict__ a, double* __restrict__ b) {
  for (long long n = blockIdx.x * (long long)blockDim.x + threadIdx.x; n < Nn;
       n += (long long)gridDim.x * blockDim.x) {
    b[3 * n] = a[n];
    b[3 * n + 1] = a[n];
    b[3 * n + 2] = a[n];
  }
}

// Smoothing step from a zero start: out = c * mask * minv * b (A*0 = 0).
// FP32 with 16-byte aligned rows (vec4): four nodes per thread through
// float4 / uint32 accesses, the tail node by node.
template <typename T>
__device__ __forceinline__ void zero_step_run(long long Nn, const uint8_t* __restrict__ mask,
                                              const T* __restrict__ minv, const T* __restrict__ b,
                                              T* __restrict__ out, T c, int vec4) {
  long long start = 0;
  if (std::is_same<T, float>::value && vec4) {
    const long long ng = Nn >> 2;
    for (long long g = blockIdx.x * (long long)blockDim.x + threadIdx.x; g < ng;
         g += (long long)gridDim.x * blockDim.x) {
      const unsigned mk4 = reinterpret_cast<const unsigned*>(mask)[g];
      const float4 mv = reinterpret_cast<const float4*>(minv)[g];
      const float4* b4 = reinterpret_cast<const float4*>(b) + 3 * g;
      float4* o4 = reinterpret_cast<float4*>(out) + 3 * g;
      float v[12];
      const float4 q0 = b4[0], q1 = b4[1], q2 = b4[2];
      v[0] = q0.x; v[1] = q0.y; v[2] = q0.z; v[3] = q0.w;
      v[4] = q1.x; v[5] = q1.y; v[6] = q1.z; v[7] = q1.w;
      v[8] = q2.x; v[9] = q2.y; v[10] = q2.z; v[11] = q2.w;
      const float w4[4] = {(float)c * mv.x, (float)c * mv.y, (float)c * mv.z, (float)c * mv.w};
#pragma unroll
      for (int i = 0; i < 12; ++i) {
        const int nd = i / 3, k = i - 3 * (i / 3);
        const unsigned mk = (mk4 >> (8 * nd)) & 0xFFu;
        v[i] = ((mk >> k) & 1u) ? w4[nd] * v[i] : 0.f;
      }
      o4[0] = make_float4(v[0], v[1], v[2], v[3]);
      o4[1] = make_float4(v[4], v[5], v[6], v[7]);
      o4[2] = make_float4(v[8], v[9], v[10], v[11]);
    }
    start = ng << 2;
  }
  for (long long n = start + blockIdx.x * (long long)blockDim.x + threadIdx.x; n < Nn;
       n += (long long)gridDim.x * blockDim.x) {
    const unsigned mk = mask[n];
    const T w = c * minv[n];
#pragma unroll
    for (int k = 0; k < 3; ++k) out[3 * n + k] = ((mk >> k) & 1u) ? w * b[3 * n + k] : T(0);
  }
}

template <typename T>
__global__ void k_zero_step(long long Nn, const uint8_t* __restrict__ mask,
                            const T* __restrict__ minv, const T* __restrict__ b,
                            T* __restrict__ out, T c, int vec4) {
  const long long off = blockIdx.y * 3 * Nn;
  zero_step_run<T>(Nn, mask, minv, b + off, out + off, c, vec4);
}

// Grid transfers marching in z.  One thread per column of the 2-D plane
// grid, z split into chunks so that about TR_TARGET_GROUPS work-groups run;
// the 1-D launch is numbered XCD-aware (blocks are dealt to the 8 XCDs
// round-robin; each XCD receives a contiguous run of columns, whose shared
// rows then stay in its L2).  Work-groups: groups_xy * chunks * nb, the grid
// rounded up to a multiple of 8 (the extra blocks exit at once).
constexpr int TR_THREADS = 256;
constexpr int TR_TARGET_GROUPS = 2048;

struct TrGrid {
  int gxy, nchunk, per_chunk;   // per_chunk: planes (coarse / fine) per chunk
};

__device__ __forceinline__ bool tr_block(int total, int& v) {
  const int per = (total + 7) >> 3;
  v = (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
  return v < total;
}

// R = P^T (27-point full weighting, unscaled) then the coarse mask
// (core/multigrid.py:404-416), z-marching: the
// thread of coarse column (jx, jy) reads each fine plane once.  Each coarse
// value is the tap-by-tap sum over (z, y, x) in that order, the odd plane
// below carried from the previous coarse plane as its own partial sum (the
// weights are powers of 2, so that partial is exact to reuse).  With x1 set,
// also the coarse level's first smoothing step from zero, w0 * minv * rc.
// work-group v (of g.gxy * g.nchunk) of one right-hand side's restriction
template <typename T>
__device__ __forceinline__ void restrict_run(int Lcx, int Lcy, int Lcz, const TrGrid& g, int v,
                                             const T* __restrict__ rf,
                                             const uint8_t* __restrict__ cmask,
                                             T* __restrict__ rc, const T* __restrict__ cminv,
                                             T w0, T* __restrict__ x1) {
  const int bxy = v % g.gxy;
  const int chunk = v / g.gxy;
  const int t = bxy * TR_THREADS + (int)threadIdx.x;
  const int jx = t % Lcx, jy = t / Lcx;
  if (jy >= Lcy) return;
  const int Lfx = 2 * Lcx - 1, Lfy = 2 * Lcy - 1, Lfz = 2 * Lcz - 1;
  const long long fplane = 3LL * Lfx * Lfy;
  const int jz0 = chunk * g.per_chunk;
  const int jz1 = min(jz0 + g.per_chunk, Lcz);
  // x-y taps of this column (y outer, x inner): fine offsets and weights;
  // a tap outside the grid reads the column's own node with weight 0, so
  // every tap is loaded without a branch (per-lane branches around loads made
  // each load wait for the one before it)
  int off[9];
  T w[9];
#pragma unroll
  for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
    for (int dx = -1; dx <= 1; ++dx) {
      const int k = (dy + 1) * 3 + (dx + 1);
      const int x = 2 * jx + dx, y = 2 * jy + dy;
      const bool in = x >= 0 && x < Lfx && y >= 0 && y < Lfy;
      off[k] = in ? 3 * (y * Lfx + x) : 3 * (2 * jy * Lfx + 2 * jx);
      w[k] = in ? (dx == 0 ? T(1) : T(0.5)) * (dy == 0 ? T(1) : T(0.5)) : T(0);
    }
  auto load_plane = [&](int z, T (&q)[9][3]) {
    const T* P = rf + (long long)z * fplane;
#pragma unroll
    for (int k = 0; k < 9; ++k)
#pragma unroll
      for (int c = 0; c < 3; ++c) q[k][c] = P[off[k] + c];
  };
  T h[3] = {T(0), T(0), T(0)};   // partial sum of plane 2 jz - 1 (weight 1/2)
  if (jz0 > 0) {
    T q[9][3];
    load_plane(2 * jz0 - 1, q);
#pragma unroll
    for (int k = 0; k < 9; ++k)
#pragma unroll
      for (int c = 0; c < 3; ++c) h[c] = fmaT(T(0.5) * w[k], q[k][c], h[c]);
  }
  for (int jz = jz0; jz < jz1; ++jz) {
    // the coarse node's mask and D^-1 and both fine planes' taps are loaded
    // before the first wait.  Past the grid's last plane the odd plane is the
    // even one again, with weight 0 (the loop's last step: h is not used).
    const bool odd = 2 * jz + 1 < Lfz;   // uniform
    const long long J = ((long long)jz * Lcy + jy) * Lcx + jx;
    const unsigned mk = cmask[J];
    const T cmi = x1 ? cminv[J] : T(0);
    T q0[9][3], q1[9][3];
    load_plane(2 * jz, q0);
    load_plane(odd ? 2 * jz + 1 : 2 * jz, q1);
    T a[3] = {h[0], h[1], h[2]};
#pragma unroll
    for (int k = 0; k < 9; ++k)
#pragma unroll
      for (int c = 0; c < 3; ++c) a[c] = fmaT(w[k], q0[k][c], a[c]);
    const T so = odd ? T(0.5) : T(0);
    h[0] = h[1] = h[2] = T(0);
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      const T wk = so * w[k];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        a[c] = fmaT(wk, q1[k][c], a[c]);
        h[c] = fmaT(wk, q1[k][c], h[c]);
      }
    }
    rc[3 * J] = (mk & 1u) ? a[0] : T(0);
    rc[3 * J + 1] = (mk & 2u) ? a[1] : T(0);
    rc[3 * J + 2] = (mk & 4u) ? a[2] : T(0);
    if (x1) {
      // the coarse level's first smoothing step from zero, as k_zero_step
      // forms it from rc
      const T wm = w0 * cmi;
      x1[3 * J] = (mk & 1u) ? wm * a[0] : T(0);
      x1[3 * J + 1] = (mk & 2u) ? wm * a[1] : T(0);
      x1[3 * J + 2] = (mk & 4u) ? wm * a[2] : T(0);
    }
  }
}

template <typename T>
__global__ __launch_bounds__(TR_THREADS) void k_restrict_z(int Lcx, int Lcy, int Lcz, TrGrid g,
                                                           int nb, const T* __restrict__ rf,
                                                           const uint8_t* __restrict__ cmask,
                                                           T* __restrict__ rc,
                                                           const T* __restrict__ cminv, T w0,
                                                           T* __restrict__ x1) {
  int v;
  if (!tr_block(g.gxy * g.nchunk * nb, v)) return;
  const int per_rhs = g.gxy * g.nchunk;
  const int rhs = v / per_rhs;
  const long long fstride = 3LL * (2 * Lcx - 1) * (2 * Lcy - 1) * (2 * Lcz - 1);
  const long long cstride = 3LL * Lcx * Lcy * Lcz;
  restrict_run<T>(Lcx, Lcy, Lcz, g, v - rhs * per_rhs, rf + rhs * fstride, cmask,
                  rc + rhs * cstride, cminv, w0, x1 ? x1 + rhs * cstride : nullptr);
}

// x += fine mask * (P e) (core/multigrid.py:418-424, 434) with one thread per
// fine value (node, component) of a plane: a wave's fine loads and stores
// cover one contiguous span (whole cache lines), its coarse loads a
// contiguous span about half as long.  The thread marches in z over its
// chunk's fine planes; per coarse plane it forms its interpolant's x-y sum
// once.  Each interpolant is the sum of its coarse values in (z, y, x) order
// times the product of the 1/2 weights (a power of 2, so applying it last is
// exact).
// work-group v (of g.gxy * g.nchunk) of one right-hand side's prolongation
template <typename T>
__device__ __forceinline__ void prolong_run(int Lcx, int Lcy, int Lcz, const TrGrid& g, int v,
                                            const T* __restrict__ ec,
                                            const uint8_t* __restrict__ fmask,
                                            T* __restrict__ x) {
  const int bxy = v % g.gxy;
  const int chunk = v / g.gxy;
  const int Lfx = 2 * Lcx - 1, Lfy = 2 * Lcy - 1, Lfz = 2 * Lcz - 1;
  const int fplane = Lfx * Lfy;   // nodes per fine plane
  const int f = bxy * TR_THREADS + (int)threadIdx.x;   // value index in a plane
  if (f >= 3 * fplane) return;
  const int node = f / 3, c = f - 3 * node;
  const int iy = node / Lfx, ix = node - iy * Lfx;
  const int cx = ix >> 1, cy = iy >> 1;
  const bool ox = ix & 1, oy = iy & 1;
  const long long cplane = 3LL * Lcx * Lcy;
  // the coarse taps of this value inside a coarse plane, (ky, kx) order; a tap
  // the value does not use repeats t00 and adds 0, so all four are loaded
  // without a branch (per-lane branches around the loads made each load wait
  // for the one before it)
  const int t00 = 3 * (cy * Lcx + cx) + c;
  const int t01 = t00 + (ox ? 3 : 0), t10 = t00 + (oy ? 3 * Lcx : 0);
  const int t11 = t10 + (ox ? 3 : 0);
  const bool oxy = ox && oy;
  auto taps = [&](int cz, T (&v)[4]) {
    const T* P = ec + (long long)cz * cplane;
    v[0] = P[t00];
    v[1] = P[t01];
    v[2] = P[t10];
    v[3] = P[t11];
  };
  auto cont = [&](const T (&v)[4], T sacc) -> T {   // sacc continued with a plane's taps
    sacc += v[0];
    sacc += ox ? v[1] : T(0);
    sacc += oy ? v[2] : T(0);
    sacc += oxy ? v[3] : T(0);
    return sacc;
  };
  auto csum = [&](const T (&v)[4]) -> T {
    T sacc = v[0];
    sacc += ox ? v[1] : T(0);
    sacc += oy ? v[2] : T(0);
    sacc += oxy ? v[3] : T(0);
    return sacc;
  };
  const T wxy = (ox ? T(0.5) : T(1)) * (oy ? T(0.5) : T(1));
  const int iz0 = chunk * g.per_chunk;
  const int iz1 = min(iz0 + g.per_chunk, Lfz);
  const long long fp3 = 3LL * fplane;
  T lo;
  {
    T v[4];
    taps(iz0 >> 1, v);
    lo = csum(v);
  }
  // four fine planes (two coarse planes) per step, every load issued before
  // the first store: more loads in flight per wave
  for (int iz = iz0; iz < iz1; iz += 4) {
    const int cz = iz >> 1;
    const bool h1 = iz + 1 < iz1, h2 = iz + 2 < iz1, h3 = iz + 3 < iz1;   // uniform
    const long long ia = (long long)iz * fp3 + f;
    const long long im = (long long)iz * fplane + node;
    const unsigned ma = fmask[im];
    const unsigned mb = h1 ? fmask[im + fplane] : 0u;
    const unsigned mc = h2 ? fmask[im + 2LL * fplane] : 0u;
    const unsigned md = h3 ? fmask[im + 3LL * fplane] : 0u;
    T va = x[ia];
    const T vb = h1 ? x[ia + fp3] : T(0);
    const T vc = h2 ? x[ia + 2 * fp3] : T(0);
    const T vd = h3 ? x[ia + 3 * fp3] : T(0);
    T s1 = T(0), v1 = T(0), s2 = T(0), v3 = T(0);
    T p1[4], p2[4];
    if (h1) taps(cz + 1, p1);
    if (h3) taps(cz + 2, p2);
    if (h1) {
      s1 = csum(p1);        // plane cz+1's own sum
      v1 = cont(p1, lo);    // plane cz's sum continued with plane cz+1
    }
    if (h3) {
      s2 = csum(p2);
      v3 = cont(p2, s1);
    }
    // every value is stored (unchanged where masked): whole cache lines
    if ((ma >> c) & 1u) va += lo * wxy;
    x[ia] = va;
    if (h1) {
      T w = vb;
      if ((mb >> c) & 1u) w += v1 * (T(0.5) * wxy);
      x[ia + fp3] = w;
    }
    if (h2) {
      T w = vc;
      if ((mc >> c) & 1u) w += s1 * wxy;
      x[ia + 2 * fp3] = w;
    }
    if (h3) {
      T w = vd;
      if ((md >> c) & 1u) w += v3 * (T(0.5) * wxy);
      x[ia + 3 * fp3] = w;
      lo = s2;
    }
  }
}

template <typename T>
__global__ __launch_bounds__(TR_THREADS) void k_prolong_f(int Lcx, int Lcy, int Lcz, TrGrid g,
                                                          int nb, const T* __restrict__ ec,
                                                          const uint8_t* __restrict__ fmask,
                                                          T* __restrict__ x) {
  int v;
  if (!tr_block(g.gxy * g.nchunk * nb, v)) return;
  const int per_rhs = g.gxy * g.nchunk;
  const int rhs = v / per_rhs;
  const long long cstride = 3LL * Lcx * Lcy * Lcz;
  const long long fstride = 3LL * (2 * Lcx - 1) * (2 * Lcy - 1) * (2 * Lcz - 1);
  prolong_run<T>(Lcx, Lcy, Lcz, g, v - rhs * per_rhs, ec + rhs * cstride, fmask, x + rhs * fstride);
}

// E of level l+1 = 2 * mean of the 8 children (scaled E carries 2^l;
// core/multigrid.py:287-292).  x-fastest layout on both levels.
// mode 0: arithmetic mean (the reference rule); 1: geometric mean; 2:
// harmonic mean (0 when any child is 0).  The factor 2 of the level scale is
// applied to all three.
template <typename T>
__global__ void k_coarsen_E(int cnx, int cny, int cnz, const T* __restrict__ Ef, T* __restrict__ Ec,
                            int mode) {
  const int fnx = 2 * cnx, fny = 2 * cny;
  const long long Nc = (long long)cnx * cny * cnz;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < Nc;
       e += (long long)gridDim.x * blockDim.x) {
    const int x = (int)(e % cnx);
    const int y = (int)((e / cnx) % cny);
    const int z = (int)(e / ((long long)cnx * cny));
    if (mode == 0) {
      T s = T(0);
      for (int dz = 0; dz < 2; ++dz)
        for (int dy = 0; dy < 2; ++dy) {
          const long long r = ((long long)(2 * z + dz) * fny + (2 * y + dy)) * fnx + 2 * x;
          s += Ef[r] + Ef[r + 1];
        }
      Ec[e] = s * T(0.25);
    } else {
      double acc = 0.0;
      bool zero = false;
      for (int dz = 0; dz < 2; ++dz)
        for (int dy = 0; dy < 2; ++dy)
          for (int dx = 0; dx < 2; ++dx) {
            const double v = (double)Ef[((long long)(2 * z + dz) * fny + (2 * y + dy)) * fnx + 2 * x + dx];
            if (v <= 0.0) zero = true;
            else acc += (mode == 1) ? log(v) : 1.0 / v;
          }
      double m = 0.0;
      if (mode == 1) m = zero ? 0.0 : exp(acc / 8.0);
      else m = zero ? 0.0 : 8.0 / acc;
      Ec[e] = (T)(2.0 * m);
    }
  }
}

// Reference element order (z fastest) -> x-fastest, with the void rule
// E*e_min where active == 0 (core/fea.py:467-469).  Writes FP64 and FP32.
__global__ void k_import_E(int nx, int ny, int nz, const double* __restrict__ Eref,
                           const uint8_t* __restrict__ active, double emin,
                           double* __restrict__ E64, float* __restrict__ E32) {
  const long long Ne = (long long)nx * ny * nz;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < Ne;
       e += (long long)gridDim.x * blockDim.x) {
    const int x = (int)(e % nx);
    const int y = (int)((e / nx) % ny);
    const int z = (int)(e / ((long long)nx * ny));
    const long long r = (long long)z + (long long)nz * (y + (long long)ny * x);
    double v = Eref[r];
    if (active && !active[r]) v = v * emin;
    E64[e] = v;
    if (E32) E32[e] = (float)v;
  }
}

template <typename Tin, typename Tout>
__global__ void k_convert(long long n, const Tin* __restrict__ a, Tout* __restrict__ b) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    b[i] = (Tout)a[i];
}

// Effective free mask of a level: the fixed-DOF mask, and in the active-DOF
// system also 0 where no adjacent element carries stiffness.
__global__ void k_effective_mask(int nx, int ny, int nz, const double* __restrict__ E,
                                 const uint8_t* __restrict__ mfixed, int drop_unloaded,
                                 uint8_t* __restrict__ mask) {
  const int Lx = nx + 1, Ly = ny + 1, Lz = nz + 1;
  const long long Nn = (long long)Lx * Ly * Lz;
  for (long long n = blockIdx.x * (long long)blockDim.x + threadIdx.x; n < Nn;
       n += (long long)gridDim.x * blockDim.x) {
    uint8_t m = mfixed[n];
    if (drop_unloaded) {
      const int ix = (int)(n % Lx);
      const int iy = (int)((n / Lx) % Ly);
      const int iz = (int)(n / ((long long)Lx * Ly));
      bool any = false;
      for (int dz = -1; dz <= 0; ++dz)
        for (int dy = -1; dy <= 0; ++dy)
          for (int dx = -1; dx <= 0; ++dx) {
            const int x = ix + dx, y = iy + dy, z = iz + dz;
            if (x >= 0 && x < nx && y >= 0 && y < ny && z >= 0 && z < nz &&
                E[((long long)z * ny + y) * nx + x] > 0.0)
              any = true;
          }
      if (!any) m = 0;
    }
    mask[n] = m;
  }
}

// Coarse free mask by injection from the collocated fine node
// (core/multigrid.py:264-278).
__global__ void k_inject_mask(int Lcx, int Lcy, int Lcz, const uint8_t* __restrict__ fmask,
                              uint8_t* __restrict__ cmask) {
  const int Lfx = 2 * Lcx - 1, Lfy = 2 * Lcy - 1;
  const long long Nc = (long long)Lcx * Lcy * Lcz;
  for (long long J = blockIdx.x * (long long)blockDim.x + threadIdx.x; J < Nc;
       J += (long long)gridDim.x * blockDim.x) {
    const int jx = (int)(J % Lcx);
    const int jy = (int)((J / Lcx) % Lcy);
    const int jz = (int)(J / ((long long)Lcx * Lcy));
    cmask[J] = fmask[((long long)(2 * jz) * Lfy + 2 * jy) * Lfx + 2 * jx];
  }
}

// ---------------------------------------------------------------------------
// PCG vector kernels.  Each writes one partial per work-group, in a fixed
// grid, so the reductions are deterministic.
// ---------------------------------------------------------------------------


enum {
  F_SET_TMP = 0,    // s.tmp = sum
  F_FNORM = 1,      // s.fnorm = sqrt(sum)
  F_RZ0 = 2,        // s.rz = sum; beta = alpha = 0, frozen = 0
  F_ALPHA = 3,      // s.pAp = sum; s.alpha = sum == 0 ? 0 : rz / sum
  F_RATIO = 4,      // s.rr = sum; s.ratio = sqrt(sum) / fnorm
  F_BETA = 5,       // s.rz_new = sum; s.beta = rz == 0 ? 0 : sum / rz; s.rz = sum
  F_SET_CS = 6,     // s.cs = sum
  F_RZNEW = 7,      // s.rz_new = sum
  F_BETA_FLEX = 8,  // sum = z.Ap: s.beta = -sum / pAp (flexible CG); s.rz = s.rz_new
  F_RZ0_KEEP = 9,   // s.rz = sum; alpha = 0, frozen = 0; beta continues the previous
                    // direction: sum / (rz_a * pscale) (0 without one or when the
                    // refresh moved the residual by more than drift_max)
  F_DRIFT = 10      // s.drift = sqrt(sum)
};

constexpr int FIN_THREADS = 1024;
// the n partials of one right-hand side summed in a fixed order, then op
__device__ __forceinline__ void finalize_run(const double* __restrict__ partial, int n, int op,
                                             Scalars* s) {
  __shared__ double sred[FIN_THREADS / 64];
  // four independent chains per thread keep four loads in flight
  double v0 = 0.0, v1 = 0.0, v2 = 0.0, v3 = 0.0;
  const int bd = blockDim.x;
  int i = threadIdx.x;
  for (; i + 3 * bd < n; i += 4 * bd) {
    v0 += partial[i];
    v1 += partial[i + bd];
    v2 += partial[i + 2 * bd];
    v3 += partial[i + 3 * bd];
  }
  for (; i < n; i += bd) v0 += partial[i];
  double v = block_sum((v0 + v1) + (v2 + v3), sred);
  if (threadIdx.x == 0) {
    switch (op) {
      case F_SET_TMP: s->tmp = v; break;
      case F_FNORM: s->fnorm = sqrt(v); break;
      case F_RZ0:
        s->rz = v;
        s->beta = 0.0;
        s->alpha = 0.0;   // no deferred solution update before the first step
        s->frozen = 0;
        break;
      case F_ALPHA:
        s->pAp = v;
        s->alpha = (v == 0.0 || s->frozen) ? 0.0 : s->rz / v;
        s->rz_a = s->rz;
        break;
      case F_RZ0_KEEP:
        s->rz = v;
        s->beta = (s->rz_a > 0.0 && s->pscale > 0.0 && (s->drift_max <= 0.0 || s->drift <= s->drift_max))
                      ? v / (s->rz_a * s->pscale)
                      : 0.0;
        s->alpha = 0.0;
        s->frozen = 0;
        break;
      case F_RATIO:
        s->rr = v;
        s->ratio = sqrt(v) / s->fnorm;
        break;
      case F_BETA:
        s->rz_new = v;
        s->beta = (s->rz == 0.0) ? 0.0 : v / s->rz;
        s->rz = v;
        break;
      case F_SET_CS: s->cs = v; break;
      case F_DRIFT: s->drift = sqrt(v); break;
      case F_RZNEW: s->rz_new = v; break;
      case F_BETA_FLEX:
        s->beta = (s->pAp == 0.0) ? 0.0 : -v / s->pAp;
        s->rz = s->rz_new;
        break;
    }
  }
}

__global__ __launch_bounds__(FIN_THREADS) void k_finalize(const double* __restrict__ partial, int n,
                                                          int op, Scalars* s) {
  finalize_run(partial + (long long)blockIdx.x * n, n, op, s + blockIdx.x);
}

__global__ void k_set_fnorm(Scalars* s, double v) { s[blockIdx.x].fnorm = v; }

// frozen flags of all right-hand sides from a device int array
__global__ void k_set_frozen(Scalars* s, const int* __restrict__ fr, int nb) {
  for (int i = threadIdx.x; i < nb; i += blockDim.x) s[i].frozen = fr[i];
}

// 4 consecutive values through one (float4) or two (double2) vector accesses
__device__ __forceinline__ void ld4(const float* p, float v[4]) {
  const float4 q = *reinterpret_cast<const float4*>(p);
  v[0] = q.x; v[1] = q.y; v[2] = q.z; v[3] = q.w;
}
__device__ __forceinline__ void st4(float* p, const float v[4]) {
  *reinterpret_cast<float4*>(p) = make_float4(v[0], v[1], v[2], v[3]);
}
__device__ __forceinline__ void ld4(const double* p, double v[4]) {
  const double2 a = *reinterpret_cast<const double2*>(p);
  const double2 b = *reinterpret_cast<const double2*>(p + 2);
  v[0] = a.x; v[1] = a.y; v[2] = b.x; v[3] = b.y;
}
__device__ __forceinline__ void st4(double* p, const double v[4]) {
  *reinterpret_cast<double2*>(p) = make_double2(v[0], v[1]);
  *reinterpret_cast<double2*>(p + 2) = make_double2(v[2], v[3]);
}

// u += alpha p ; r -= alpha Ap ; partial sum r.r  (arrays 16-byte aligned)
template <typename T>
__device__ __forceinline__ double cg_update_run(long long n, T a, T* __restrict__ u,
                                                T* __restrict__ r, const T* __restrict__ p,
                                                const T* __restrict__ Ap) {
  double acc = 0.0;
  const long long n4 = n >> 2;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n4;
       i += (long long)gridDim.x * blockDim.x) {
    T uu[4], rr[4], pp[4], aa[4];
    ld4(u + 4 * i, uu);
    ld4(r + 4 * i, rr);
    ld4(p + 4 * i, pp);
    ld4(Ap + 4 * i, aa);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      uu[k] = uu[k] + a * pp[k];
      rr[k] = rr[k] - a * aa[k];
      acc += (double)rr[k] * (double)rr[k];
    }
    st4(u + 4 * i, uu);
    st4(r + 4 * i, rr);
  }
  if (blockIdx.x == 0 && threadIdx.x < (n & 3)) {
    const long long i = 4 * n4 + threadIdx.x;
    u[i] = u[i] + a * p[i];
    const T rn = r[i] - a * Ap[i];
    r[i] = rn;
    acc += (double)rn * (double)rn;
  }
  return acc;
}

template <typename T>
__global__ void k_cg_update(long long n, const Scalars* __restrict__ s, T* __restrict__ u,
                            T* __restrict__ r, const T* __restrict__ p, const T* __restrict__ Ap,
                            double* __restrict__ partial) {
  __shared__ double sred[4];
  const long long off = (long long)blockIdx.y * n;
  double acc = cg_update_run<T>(n, (T)s[blockIdx.y].alpha, u + off, r + off, p + off, Ap + off);
  acc = block_sum(acc, sred);
  if (threadIdx.x == 0) partial[blockIdx.y * gridDim.x + blockIdx.x] = acc;
}

// The residual half of k_cg_update fused with the first smoothing step from
// zero of the next V-cycle on the same level: r -= alpha Ap ; partial r.r ;
// x1 = mask ? c * minv * r : 0 (the output k_zero_step would write from the
// new r).  The solution update u += alpha p is deferred to the next
// direction update (k_p_update_u), which streams p anyway.  FP32, four nodes
// per thread through float4 / uint32 accesses (every vector 16-byte aligned
// per right-hand side), the < 4 tail nodes one by one.  Grid (EW blocks,
// rhs); partial[rhs * gridDim.x + block].
__device__ __forceinline__ double cg_update_x1_run(long long Nn, float a, float* __restrict__ r,
                                                   const float* __restrict__ Ap,
                                                   const uint8_t* __restrict__ mask,
                                                   const float* __restrict__ minv, float c,
                                                   float* __restrict__ x1) {
  double acc = 0.0;
  const long long ng = Nn >> 2;
  for (long long g = blockIdx.x * (long long)blockDim.x + threadIdx.x; g < ng;
       g += (long long)gridDim.x * blockDim.x) {
    const unsigned mk4 = reinterpret_cast<const unsigned*>(mask)[g];
    const float4 mv = reinterpret_cast<const float4*>(minv)[g];
    float rr[12], aa[12];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      ld4(r + 12 * g + 4 * k, rr + 4 * k);
      ld4(Ap + 12 * g + 4 * k, aa + 4 * k);
    }
    const float w4[4] = {c * mv.x, c * mv.y, c * mv.z, c * mv.w};
    float xx[12];
#pragma unroll
    for (int i = 0; i < 12; ++i) {
      rr[i] = rr[i] - a * aa[i];
      acc += (double)rr[i] * (double)rr[i];
      const int nd = i / 3, k = i - 3 * (i / 3);
      const unsigned mk = (mk4 >> (8 * nd)) & 0xFFu;
      xx[i] = ((mk >> k) & 1u) ? w4[nd] * rr[i] : 0.f;
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      st4(r + 12 * g + 4 * k, rr + 4 * k);
      st4(x1 + 12 * g + 4 * k, xx + 4 * k);
    }
  }
  for (long long n = 4 * ng + blockIdx.x * (long long)blockDim.x + threadIdx.x; n < Nn;
       n += (long long)gridDim.x * blockDim.x) {
    const unsigned mk = mask[n];
    const float w = c * minv[n];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const long long i = 3 * n + k;
      const float rn = r[i] - a * Ap[i];
      r[i] = rn;
      acc += (double)rn * (double)rn;
      x1[i] = ((mk >> k) & 1u) ? w * rn : 0.f;
    }
  }
  return acc;
}

__global__ void k_cg_update_x1(long long Nn, const Scalars* __restrict__ s, float* __restrict__ r,
                               const float* __restrict__ Ap, double* __restrict__ partial,
                               const uint8_t* __restrict__ mask, const float* __restrict__ minv,
                               float c, float* __restrict__ x1) {
  __shared__ double sred[4];
  const long long off = (long long)blockIdx.y * 3 * Nn;
  double acc = cg_update_x1_run(Nn, (float)s[blockIdx.y].alpha, r + off, Ap + off, mask, minv, c,
                                x1 + off);
  acc = block_sum(acc, sred);
  if (threadIdx.x == 0) partial[blockIdx.y * gridDim.x + blockIdx.x] = acc;
}

// r -= alpha Ap and the r.r partials (FP32, 16-byte aligned rows): the PCG
// update when the next V-cycle's level-0 pre-smoothing runs as M_CHEB_Z2
// (u += alpha p is deferred, as with k_cg_update_x1)
__global__ void k_cg_update_r(long long n, const Scalars* __restrict__ s, float* __restrict__ r,
                              const float* __restrict__ Ap, double* __restrict__ partial) {
  __shared__ double sred[4];
  const long long off = (long long)blockIdx.y * n;
  const float a = (float)s[blockIdx.y].alpha;
  float* rr = r + off;
  const float* aa = Ap + off;
  double acc = 0.0;
  const long long n4 = n >> 2;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n4;
       i += (long long)gridDim.x * blockDim.x) {
    float v[4], w[4];
    ld4(rr + 4 * i, v);
    ld4(aa + 4 * i, w);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      v[k] = v[k] - a * w[k];
      acc += (double)v[k] * (double)v[k];
    }
    st4(rr + 4 * i, v);
  }
  if (blockIdx.x == 0 && threadIdx.x < (n & 3)) {
    const long long i = 4 * n4 + threadIdx.x;
    const float v = rr[i] - a * aa[i];
    rr[i] = v;
    acc += (double)v * (double)v;
  }
  acc = block_sum(acc, sred);
  if (threadIdx.x == 0) partial[blockIdx.y * gridDim.x + blockIdx.x] = acc;
}

// p = z + beta p per right-hand side (blockIdx.y = rhs, n values each)
template <typename T>
__device__ __forceinline__ void p_update_b_run(long long n, T beta, const T* __restrict__ z,
                                               T* __restrict__ p) {
  const long long n4 = n >> 2;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n4;
       i += (long long)gridDim.x * blockDim.x) {
    T zz[4], pp[4];
    ld4(z + 4 * i, zz);
    ld4(p + 4 * i, pp);
#pragma unroll
    for (int k = 0; k < 4; ++k) pp[k] = zz[k] + beta * pp[k];
    st4(p + 4 * i, pp);
  }
  if (blockIdx.x == 0 && threadIdx.x < (n & 3)) {
    const long long i = 4 * n4 + threadIdx.x;
    p[i] = z[i] + beta * p[i];
  }
}

template <typename T>
__global__ void k_p_update_b(long long n, const Scalars* __restrict__ s, const T* __restrict__ z,
                             T* __restrict__ p) {
  const long long off = (long long)blockIdx.y * n;
  p_update_b_run<T>(n, (T)s[blockIdx.y].beta, z + off, p + off);
}

// The deferred solution update of FP32 PCG with the direction update: u +=
// alpha p with the previous step's alpha (this step's F_ALPHA has not run
// yet; 0 before the first step), then p = z + beta p; the arithmetic of
// k_cg_update's u update followed by k_p_update_b.  Per right-hand side
// (blockIdx.y, n values each).
template <typename T>
__device__ __forceinline__ void p_update_u_run(long long n, T alpha, T beta,
                                               const T* __restrict__ z, T* __restrict__ p,
                                               T* __restrict__ u) {
  const long long n4 = n >> 2;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n4;
       i += (long long)gridDim.x * blockDim.x) {
    T zz[4], pp[4], uu[4];
    ld4(z + 4 * i, zz);
    ld4(p + 4 * i, pp);
    ld4(u + 4 * i, uu);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      uu[k] = uu[k] + alpha * pp[k];
      pp[k] = zz[k] + beta * pp[k];
    }
    st4(u + 4 * i, uu);
    st4(p + 4 * i, pp);
  }
  if (blockIdx.x == 0 && threadIdx.x < (n & 3)) {
    const long long i = 4 * n4 + threadIdx.x;
    u[i] = u[i] + alpha * p[i];
    p[i] = z[i] + beta * p[i];
  }
}

template <typename T>
__device__ __forceinline__ void u_flush_run(long long n, T alpha, const T* __restrict__ p,
                                            T* __restrict__ u) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    u[i] = u[i] + alpha * p[i];
}

// a frozen right-hand side takes only its pending u += alpha p, and nothing
// once its alpha is 0 (its p is no longer used)
template <typename T>
__global__ void k_p_update_u(long long n, const Scalars* __restrict__ s, const T* __restrict__ z,
                             T* __restrict__ p, T* __restrict__ u) {
  const Scalars& q = s[blockIdx.y];
  if (q.frozen && q.alpha == 0.0) return;
  const long long off = (long long)blockIdx.y * n;
  if (q.frozen)
    u_flush_run<T>(n, (T)q.alpha, p + off, u + off);
  else
    p_update_u_run<T>(n, (T)q.alpha, (T)q.beta, z + off, p + off, u + off);
}


// the last deferred update: u += alpha p per right-hand side (blockIdx.y)
template <typename T>
__global__ void k_u_flush(long long n, const Scalars* __restrict__ s, const T* __restrict__ p,
                          T* __restrict__ u) {
  const long long off = (long long)blockIdx.y * n;
  u_flush_run<T>(n, (T)s[blockIdx.y].alpha, p + off, u + off);
}

// p = z + beta p
template <typename T>
__global__ void k_p_update(long long n, const Scalars* __restrict__ s, const T* __restrict__ z,
                           T* __restrict__ p) {
  const T beta = (T)s->beta;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    p[i] = z[i] + beta * p[i];
}

// out = mask * in (per DOF), partial sum out.out
template <typename Tin, typename Tout>
__device__ __forceinline__ double mask_copy_run(long long nn, const uint8_t* __restrict__ mask,
                                                const Tin* in, Tout* out) {
  double acc = 0.0;
  for (long long n = blockIdx.x * (long long)blockDim.x + threadIdx.x; n < nn;
       n += (long long)gridDim.x * blockDim.x) {
    const unsigned mk = mask[n];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const Tout v = ((mk >> c) & 1u) ? (Tout)in[3 * n + c] : Tout(0);
      out[3 * n + c] = v;
      acc += (double)v * (double)v;
    }
  }
  return acc;
}

template <typename Tin, typename Tout>
__global__ void k_mask_copy(long long nn, const uint8_t* __restrict__ mask, const Tin* in,
                            Tout* out, double* __restrict__ partial) {
  __shared__ double sred[4];
  const long long off = blockIdx.y * 3 * nn;
  double acc = mask_copy_run<Tin, Tout>(nn, mask, in + off, out + off);
  if (partial) {
    acc = block_sum(acc, sred);
    if (threadIdx.x == 0) partial[blockIdx.y * gridDim.x + blockIdx.x] = acc;
  }
}

// partial sum of a.b
template <typename T>
__global__ void k_dot(long long n, const T* __restrict__ a, const T* __restrict__ b,
                      double* __restrict__ partial) {
  __shared__ double sred[4];
  a += blockIdx.y * n;
  b += blockIdx.y * n;
  double acc = 0.0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    acc += (double)a[i] * (double)b[i];
  acc = block_sum(acc, sred);
  if (threadIdx.x == 0) partial[blockIdx.y * gridDim.x + blockIdx.x] = acc;
}

// r32 = (float)(r64 * scale) with scale = 1/||r64|| from s->tmp = ||r64||^2
__global__ void k_scale_to_f32(long long n, const Scalars* __restrict__ s,
                               const double* __restrict__ r64, float* __restrict__ r32) {
  // tmp2 != 0 marks a right-hand side whose refinement is done: zero input
  const double nr = sqrt(s[blockIdx.y].tmp);
  const double sc = (nr > 0.0 && s[blockIdx.y].tmp2 == 0.0) ? 1.0 / nr : 0.0;
  r64 += blockIdx.y * n;
  r32 += blockIdx.y * n;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    r32[i] = (float)(r64[i] * sc);
}

// k_scale_to_f32 for a refresh that may continue the previous inner solve's
// direction: before r32 (the previous inner recurrence residual, in that
// solve's units) is overwritten, the partial sums of ||r64 / ||r64|| -
// pscale * r32||^2, the refresh's move in units of the new residual
__global__ void k_scale_to_f32_drift(long long n, const Scalars* __restrict__ s,
                                     const double* __restrict__ r64, float* __restrict__ r32,
                                     double* __restrict__ partial) {
  __shared__ double sred[4];
  const double nr = sqrt(s[blockIdx.y].tmp);
  const double sc = (nr > 0.0 && s[blockIdx.y].tmp2 == 0.0) ? 1.0 / nr : 0.0;
  const double c = s[blockIdx.y].pscale;
  r64 += blockIdx.y * n;
  r32 += blockIdx.y * n;
  double acc = 0.0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const double v = r64[i] * sc;
    const double d = v - c * (double)r32[i];
    acc += d * d;
    r32[i] = (float)v;
  }
  acc = block_sum(acc, sred);
  if (threadIdx.x == 0) partial[blockIdx.y * gridDim.x + blockIdx.x] = acc;
}

// u64 += d32 * ||r64|| (undo the inner scaling)
__global__ void k_add_scaled(long long n, const Scalars* __restrict__ s,
                             const float* __restrict__ d32, double* __restrict__ u64) {
  const double sc = sqrt(s[blockIdx.y].tmp);
  d32 += blockIdx.y * n;
  u64 += blockIdx.y * n;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    u64[i] += (double)d32[i] * sc;
}

// ---------------------------------------------------------------------------
// Von Mises stress at the element centroid (core/fea.py:604-626).  The
// centroid B has entries +-1/4, so the strains are the linear Hadamard modes
// divided by 4.  Reference element order in and out.
// ---------------------------------------------------------------------------
__global__ void k_von_mises(int nx, int ny, int nz, double nu, const double* __restrict__ u,
                            const double* __restrict__ E, const uint8_t* __restrict__ active,
                            double* __restrict__ vm) {
  const long long Ne = (long long)nx * ny * nz;
  const int Lx = nx + 1, Ly = ny + 1;
  const double c = 1.0 / ((1.0 + nu) * (1.0 - 2.0 * nu));
  const double c11 = c * (1.0 - nu), c12 = c * nu, c44 = c * (1.0 - 2.0 * nu) / 2.0;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < Ne;
       e += (long long)gridDim.x * blockDim.x) {
    if (active && !active[e]) {
      vm[e] = 0.0;
      continue;
    }
    const int iz = (int)(e % nz);
    const int iy = (int)((e / nz) % ny);
    const int ix = (int)(e / ((long long)nz * ny));
    double h[3][3];   // h[mode x,y,z][component]
    for (int m = 0; m < 3; ++m)
      for (int k = 0; k < 3; ++k) h[m][k] = 0.0;
    for (int q = 0; q < 8; ++q) {
      const int sx = q & 1, sy = (q >> 1) & 1, sz = q >> 2;
      const long long n = ((long long)(iz + sz) * Ly + (iy + sy)) * Lx + (ix + sx);
      for (int k = 0; k < 3; ++k) {
        const double v = u[3 * n + k];
        h[0][k] += sx ? v : -v;
        h[1][k] += sy ? v : -v;
        h[2][k] += sz ? v : -v;
      }
    }
    const double exx = 0.25 * h[0][0], eyy = 0.25 * h[1][1], ezz = 0.25 * h[2][2];
    const double gxy = 0.25 * (h[1][0] + h[0][1]);
    const double gyz = 0.25 * (h[2][1] + h[1][2]);
    const double gzx = 0.25 * (h[2][0] + h[0][2]);
    const double Ee = E[e];
    const double sx_ = Ee * (c11 * exx + c12 * eyy + c12 * ezz);
    const double sy_ = Ee * (c12 * exx + c11 * eyy + c12 * ezz);
    const double sz_ = Ee * (c12 * exx + c12 * eyy + c11 * ezz);
    const double txy = Ee * c44 * gxy, tyz = Ee * c44 * gyz, tzx = Ee * c44 * gzx;
    const double q2 = 0.5 * ((sx_ - sy_) * (sx_ - sy_) + (sy_ - sz_) * (sy_ - sz_) +
                             (sz_ - sx_) * (sz_ - sx_)) +
                      3.0 * (txy * txy + tyz * tyz + tzx * tzx);
    vm[e] = sqrt(q2 > 0.0 ? q2 : 0.0);
  }
}

// ---------------------------------------------------------------------------
// Jacobi-PCG fallback helpers (core/fea.py:374-398, 497-593).
// ---------------------------------------------------------------------------
// Reduced-system mask (free and touched by an active element) and inverse
// diagonal per node from the active-only E (x-fastest layout).
__global__ void k_jacobi_setup(int nx, int ny, int nz, const uint8_t* __restrict__ active,
                               const uint8_t* __restrict__ mask0, const double* __restrict__ Ea,
                               double kd, uint8_t* __restrict__ jmask,
                               double* __restrict__ minv) {
  const int Lx = nx + 1, Ly = ny + 1, Lz = nz + 1;
  const long long Nn = (long long)Lx * Ly * Lz;
  for (long long n = blockIdx.x * (long long)blockDim.x + threadIdx.x; n < Nn;
       n += (long long)gridDim.x * blockDim.x) {
    const int ix = (int)(n % Lx);
    const int iy = (int)((n / Lx) % Ly);
    const int iz = (int)(n / ((long long)Lx * Ly));
    bool used = false;
    double es = 0.0;
    for (int dz = -1; dz <= 0; ++dz)
      for (int dy = -1; dy <= 0; ++dy)
        for (int dx = -1; dx <= 0; ++dx) {
          const int x = ix + dx, y = iy + dy, z = iz + dz;
          if (x >= 0 && x < nx && y >= 0 && y < ny && z >= 0 && z < nz) {
            const long long r = (long long)z + (long long)nz * (y + (long long)ny * x);
            if (!active || active[r]) used = true;
            es += Ea[((long long)z * ny + y) * nx + x];
          }
        }
    jmask[n] = used ? mask0[n] : 0;
    const double d = kd * es;
    minv[n] = 1.0 / (d == 0.0 ? 1.0 : d);
  }
}

// z = mask * minv * r ; partial r.z
__global__ void k_diag_precond(long long Nn, const uint8_t* __restrict__ mask,
                               const double* __restrict__ minv, const double* __restrict__ r,
                               double* __restrict__ z, double* __restrict__ partial) {
  __shared__ double sred[4];
  double acc = 0.0;
  for (long long n = blockIdx.x * (long long)blockDim.x + threadIdx.x; n < Nn;
       n += (long long)gridDim.x * blockDim.x) {
    const unsigned mk = mask[n];
    const double mi = minv[n];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const double rv = r[3 * n + c];
      const double zv = ((mk >> c) & 1u) ? mi * rv : 0.0;
      z[3 * n + c] = zv;
      acc += rv * zv;
    }
  }
  acc = block_sum(acc, sred);
  if (threadIdx.x == 0) partial[blockIdx.x] = acc;
}

// partial count of free DOFs
__global__ void k_count_free(long long Nn, const uint8_t* __restrict__ mask,
                             double* __restrict__ partial) {
  __shared__ double sred[4];
  double acc = 0.0;
  for (long long n = blockIdx.x * (long long)blockDim.x + threadIdx.x; n < Nn;
       n += (long long)gridDim.x * blockDim.x)
    acc += (double)__popc((unsigned)mask[n] & 7u);
  acc = block_sum(acc, sred);
  if (threadIdx.x == 0) partial[blockIdx.x] = acc;
}

// u_e^T KE u_e per active element, unit E (core/fea.py:595-602).
__global__ void k_strain_energy(int nx, int ny, int nz, ElemConst<double> K,
                                const double* __restrict__ u, const uint8_t* __restrict__ active,
                                double* __restrict__ ce) {
  const long long Ne = (long long)nx * ny * nz;
  const int Lx = nx + 1, Ly = ny + 1;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < Ne;
       e += (long long)gridDim.x * blockDim.x) {
    if (active && !active[e]) {
      ce[e] = 0.0;
      continue;
    }
    const int iz = (int)(e % nz);
    const int iy = (int)((e / nz) % ny);
    const int ix = (int)(e / ((long long)nz * ny));
    double ue[8][3], ye[8][3];
    for (int q = 0; q < 8; ++q) {
      const int sx = q & 1, sy = (q >> 1) & 1, sz = q >> 2;
      const long long n = ((long long)(iz + sz) * Ly + (iy + sy)) * Lx + (ix + sx);
      for (int c = 0; c < 3; ++c) ue[q][c] = u[3 * n + c];
    }
    elem_apply<double>(ue, 1.0, K, ye);
    double s = 0.0;
    for (int q = 0; q < 8; ++q)
      for (int c = 0; c < 3; ++c) s += ue[q][c] * ye[q][c];
    ce[e] = s;
  }
}

// ---------------------------------------------------------------------------
// Island removal (core/fea.py:165-227).  Reference element order.
// ---------------------------------------------------------------------------
__global__ void k_island_seed(int nx, int ny, int nz, const uint8_t* __restrict__ active,
                              const uint8_t* __restrict__ node_fixed, uint8_t* __restrict__ reached) {
  const long long Ne = (long long)nx * ny * nz;
  const int Lx = nx + 1, Ly = ny + 1;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < Ne;
       e += (long long)gridDim.x * blockDim.x) {
    const int iz = (int)(e % nz);
    const int iy = (int)((e / nz) % ny);
    const int ix = (int)(e / ((long long)nz * ny));
    uint8_t s = 0;
    for (int q = 0; q < 8; ++q) {
      const int sx = q & 1, sy = (q >> 1) & 1, sz = q >> 2;
      s |= node_fixed[((long long)(iz + sz) * Ly + (iy + sy)) * Lx + (ix + sx)];
    }
    reached[e] = (s && active[e]) ? 1 : 0;
  }
}

// One monotone dilation pass along the lines of one axis, in place: an
// active element becomes reached when a face neighbour is.  Each thread sweeps
// one line forward and back, carrying the state of the element behind it in a
// register, so the only serial dependence is that register.  The line's
// elements are loaded ISL_U at a time before any of them is stored: none of
// those loads reads an element this thread stores earlier in the group (an
// element's own value is loaded before its store; the one ahead is read as it
// was, which a monotone fill allows).  Neighbours outside the grid, and
// positions past the line's end, read an element of the line and are masked,
// so the loads carry no branches.  Consecutive threads take consecutive z (y
// for lines along z).  The fixed point, the 6-connected components holding a
// seed, does not depend on the order of the passes; a pass along any axis
// checks all six neighbours, so a pass that changes nothing ends the fill.
constexpr int ISL_U = 8;

__global__ void k_island_sweep(int nx, int ny, int nz, int axis, const uint8_t* __restrict__ active,
                               uint8_t* reached, int* changed) {
  // element e = iz + nz (iy + ny ix); a = swept axis, b and c the others
  const long long sX = (long long)ny * nz, sY = nz;
  long long sa, sb, sc;
  int na, nb, nc;
  if (axis == 0) {
    sa = sX; na = nx; sb = sY; nb = ny; sc = 1; nc = nz;
  } else if (axis == 1) {
    sa = sY; na = ny; sb = sX; nb = nx; sc = 1; nc = nz;
  } else {
    sa = 1; na = nz; sb = sX; nb = nx; sc = sY; nc = ny;
  }
  const long long lines = (long long)nb * nc;
  for (long long l = blockIdx.x * (long long)blockDim.x + threadIdx.x; l < lines;
       l += (long long)gridDim.x * blockDim.x) {
    const int ic = (int)(l % nc), ib = (int)(l / nc);
    const long long base = ib * sb + ic * sc;
    const bool hb0 = ib > 0, hb1 = ib + 1 < nb, hc0 = ic > 0, hc1 = ic + 1 < nc;
    const long long ob0 = hb0 ? -sb : 0, ob1 = hb1 ? sb : 0, oc0 = hc0 ? -sc : 0,
                    oc1 = hc1 ? sc : 0;
    int any = 0;
    for (int dir = 0; dir < 2; ++dir) {
      const long long step = dir == 0 ? sa : -sa;
      const long long e0 = dir == 0 ? base : base + (long long)(na - 1) * sa;
      unsigned prev = 0u;   // reached state of the element behind (none before the line)
      for (int t0 = 0; t0 < na; t0 += ISL_U) {
        unsigned act[ISL_U], cur[ISL_U], nbr[ISL_U];
#pragma unroll
        for (int j = 0; j < ISL_U; ++j) {
          const int t = t0 + j;
          const bool in = t < na, ahead = t + 1 < na;
          const long long e = e0 + (long long)(in ? t : na - 1) * step;
          const unsigned va = active[e], vc = reached[e], vf = reached[e + (ahead ? step : 0)];
          const unsigned x0 = reached[e + ob0], x1 = reached[e + ob1];
          const unsigned y0 = reached[e + oc0], y1 = reached[e + oc1];
          act[j] = (in && va) ? 1u : 0u;
          cur[j] = vc ? 1u : 0u;
          nbr[j] = ((ahead && vf) | (hb0 && x0) | (hb1 && x1) | (hc0 && y0) | (hc1 && y1)) ? 1u : 0u;
        }
#pragma unroll
        for (int j = 0; j < ISL_U; ++j) {
          if (t0 + j >= na) break;
          const unsigned nv = cur[j] | (act[j] & (prev | nbr[j]));
          if (nv && !cur[j]) {
            reached[e0 + (long long)(t0 + j) * step] = 1;
            any = 1;
          }
          prev = nv;
        }
      }
    }
    if (any) atomicOr(changed, 1);
  }
}

__global__ void k_count_u8(long long n, const uint8_t* __restrict__ a, unsigned long long* cnt) {
  unsigned long long c = 0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    c += a[i] ? 1 : 0;
  for (int o = 32; o >= 1; o >>= 1) c += __shfl_xor(c, o, 64);
  if ((threadIdx.x & 63) == 0 && c) atomicAdd(cnt, c);
}

// ---------------------------------------------------------------------------
// Learned coarse-grid corrector.  At the corrector level the V-cycle forms
// the residual t, the plain coarse-grid correction v (restrict, coarser
// levels, prolong) and the network's correction y, and adds the combination
// alpha v + beta y with the smallest energy error ||A^-1 t - alpha v - beta y||_A
// (a 2x2 Galerkin system per right-hand side).  Per node the network sees 16
// BF16 channels:
//   0-2  t / rms(t),
//   3    mean E of the adjacent elements / max E of the level,
//   4    1 + log2(max(that ratio, 2^-20)) / 20  (0 where no element is stiff),
//   5    fraction of free components (0, 1/3, 2/3, 1),
//   6-8  v / (gain * rms(t) * minv[n])   (v in the units of the output),
//   9-15 zero,
// and its first 3 output channels give y[3n+c] = gain * rms * minv[n] * out[n][c]
// on free components.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint16_t f2bf16(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

// max of non-negative floats: their bit patterns order like the values
__global__ void k_max_nonneg(long long n, const float* __restrict__ a, unsigned int* out) {
  __shared__ unsigned int red[256];
  unsigned int m = 0u;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    m = max(m, __float_as_uint(fmaxf(a[i], 0.f)));
  red[threadIdx.x] = m;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] = max(red[threadIdx.x], red[threadIdx.x + o]);
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicMax(out, red[0]);
}

// per-node features [Nn][4] of a level (E internal order, x fastest)
__global__ void k_net_feat(int nx, int ny, int nz, const float* __restrict__ E,
                           const uint8_t* __restrict__ mask, const unsigned int* __restrict__ emax,
                           float* __restrict__ feat) {
  const int Lx = nx + 1, Ly = ny + 1, Lz = nz + 1;
  const long long Nn = (long long)Lx * Ly * Lz;
  const float em = __uint_as_float(*emax);
  for (long long n = blockIdx.x * (long long)blockDim.x + threadIdx.x; n < Nn;
       n += (long long)gridDim.x * blockDim.x) {
    const int ix = (int)(n % Lx), iy = (int)((n / Lx) % Ly), iz = (int)(n / ((long long)Lx * Ly));
    float sum = 0.f;
    int cnt = 0;
    for (int dz = -1; dz <= 0; ++dz)
      for (int dy = -1; dy <= 0; ++dy)
        for (int dx = -1; dx <= 0; ++dx) {
          const int x = ix + dx, y = iy + dy, z = iz + dz;
          if (x < 0 || x >= nx || y < 0 || y >= ny || z < 0 || z >= nz) continue;
          sum += E[((long long)z * ny + y) * nx + x];
          ++cnt;
        }
    const float r = (cnt > 0 && em > 0.f) ? sum / (float)cnt / em : 0.f;
    feat[4 * n + 0] = r;
    feat[4 * n + 1] = r > 0.f ? 1.f + log2f(fmaxf(r, 9.5367431640625e-07f)) / 20.f : 0.f;
    feat[4 * n + 2] = (float)__popc((unsigned)mask[n] & 7u) / 3.f;
    feat[4 * n + 3] = 0.f;
  }
}

// rms[r] = sqrt(cs / ndof) per right-hand side
__global__ void k_net_rms(const Scalars* __restrict__ sc, int nb, double ndof,
                          float* __restrict__ rms) {
  for (int r = threadIdx.x; r < nb; r += blockDim.x) rms[r] = (float)sqrt(sc[r].cs / ndof);
}

// rms[r] = sqrt(sum of partials / ndof); one work-group per sample
__global__ void k_rms_from_partials(const double* __restrict__ part, int np, double ndof,
                                    float* __restrict__ rms) {
  __shared__ double sred[4];
  double v = 0.0;
  for (int i = threadIdx.x; i < np; i += blockDim.x) v += part[(long long)blockIdx.x * np + i];
  v = block_sum(v, sred);
  if (threadIdx.x == 0) rms[blockIdx.x] = (float)sqrt(v / ndof);
}

// network input [nb][Nn][16] bf16 from t and v [nb][3 Nn] (blockIdx.y = rhs)
__global__ void k_net_input(long long Nn, const float* __restrict__ t, const float* __restrict__ v,
                            const float* __restrict__ feat, const float* __restrict__ minv,
                            const float* __restrict__ rms, float gain, uint16_t* __restrict__ in) {
  const int r = blockIdx.y;
  const float inv = rms[r] > 0.f ? 1.f / rms[r] : 0.f;
  t += (long long)r * 3 * Nn;
  v += (long long)r * 3 * Nn;
  in += (long long)r * Nn * 16;
  for (long long n = blockIdx.x * (long long)blockDim.x + threadIdx.x; n < Nn;
       n += (long long)gridDim.x * blockDim.x) {
    const float den = gain * rms[r] * minv[n];
    const float iv = den > 0.f ? 1.f / den : 0.f;
    const uint32_t c0 = f2bf16(t[3 * n] * inv), c1 = f2bf16(t[3 * n + 1] * inv);
    const uint32_t c2 = f2bf16(t[3 * n + 2] * inv);
    const uint32_t c3 = f2bf16(feat[4 * n]), c4 = f2bf16(feat[4 * n + 1]);
    const uint32_t c5 = f2bf16(feat[4 * n + 2]);
    const uint32_t c6 = f2bf16(v[3 * n] * iv), c7 = f2bf16(v[3 * n + 1] * iv);
    const uint32_t c8 = f2bf16(v[3 * n + 2] * iv);
    uint4* o = reinterpret_cast<uint4*>(in + n * 16);
    o[0] = make_uint4(c0 | (c1 << 16), c2 | (c3 << 16), c4 | (c5 << 16), c6 | (c7 << 16));
    o[1] = make_uint4(c8, 0u, 0u, 0u);
  }
}

// partials of the 5 products v.Av, v.Ay, y.Ay, v.t, y.t per right-hand side:
// part[(r * 5 + k) * gridDim.x + block]
__global__ void k_dot5(long long n3, const float* __restrict__ v, const float* __restrict__ y,
                       const float* __restrict__ Av, const float* __restrict__ Ay,
                       const float* __restrict__ t, double* __restrict__ part) {
  __shared__ double sred[4];
  const int r = blockIdx.y;
  const long long o = (long long)r * n3;
  double a[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n3;
       i += (long long)gridDim.x * blockDim.x) {
    const double vi = v[o + i], yi = y[o + i], avi = Av[o + i], ayi = Ay[o + i], ti = t[o + i];
    a[0] += vi * avi;
    a[1] += vi * ayi;
    a[2] += yi * ayi;
    a[3] += vi * ti;
    a[4] += yi * ti;
  }
  for (int k = 0; k < 5; ++k) {
    const double sum = block_sum(a[k], sred);
    if (threadIdx.x == 0) part[((long long)r * 5 + k) * gridDim.x + blockIdx.x] = sum;
    __syncthreads();
  }
}

// per right-hand side (one work-group each): the energy-optimal (alpha, beta)
// of [v.Av v.Ay; v.Ay y.Ay] (alpha, beta) = (v.t, y.t); beta = 0 when y adds
// no independent direction
__global__ void k_combine_coef(const double* __restrict__ part, int np, double* __restrict__ coef) {
  __shared__ double sred[4];
  const int r = blockIdx.x;
  double m[5];
  for (int k = 0; k < 5; ++k) {
    double v = 0.0;
    for (int i = threadIdx.x; i < np; i += blockDim.x) v += part[((long long)r * 5 + k) * np + i];
    m[k] = block_sum(v, sred);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double a = m[0], b = m[1], c = m[2], d = m[3], e = m[4];
    const double det = a * c - b * b;
    double al = a > 0.0 ? d / a : 0.0, be = 0.0;
    if (a > 0.0 && c > 0.0 && det > 1e-10 * a * c) {
      al = (c * d - b * e) / det;
      be = (a * e - b * d) / det;
    } else if (!(a > 0.0) && c > 0.0) {
      al = 0.0;
      be = e / c;
    }
    coef[2 * r] = al;
    coef[2 * r + 1] = be;
  }
}

// x += alpha v + beta y per right-hand side (blockIdx.y)
__global__ void k_axpy2(long long n3, const double* __restrict__ coef, const float* __restrict__ v,
                        const float* __restrict__ y, float* __restrict__ x) {
  const int r = blockIdx.y;
  const float al = (float)coef[2 * r], be = (float)coef[2 * r + 1];
  const long long o = (long long)r * n3;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n3;
       i += (long long)gridDim.x * blockDim.x)
    x[o + i] += al * v[o + i] + be * y[o + i];
}

// r = t - Av
__global__ void k_sub(long long n, const float* __restrict__ t, const float* __restrict__ Av,
                      float* __restrict__ r) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    r[i] = t[i] - Av[i];
}

// x[3n+c] (+)= gain * rms * minv[n] * out[n][c] on free components (blockIdx.y = rhs)
__global__ void k_net_apply(long long Nn, const float* __restrict__ out,
                            const float* __restrict__ minv, const uint8_t* __restrict__ mask,
                            const float* __restrict__ rms, float gain, float* __restrict__ x,
                            int accumulate) {
  const int r = blockIdx.y;
  const float sc = gain * rms[r];
  out += (long long)r * Nn * 16;
  x += (long long)r * 3 * Nn;
  for (long long n = blockIdx.x * (long long)blockDim.x + threadIdx.x; n < Nn;
       n += (long long)gridDim.x * blockDim.x) {
    const unsigned mk = mask[n];
    const float w = sc * minv[n];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const float v = ((mk >> c) & 1u) ? w * out[n * 16 + c] : 0.f;
      if (accumulate) x[3 * n + c] += v;
      else x[3 * n + c] = v;
    }
  }
}

// chain rule through k_net_apply: dout[n][c] = gain * rms * minv[n] * gy[3n+c]
// on free components, 0 elsewhere and on channels 3-15
__global__ void k_net_apply_bwd(long long Nn, const float* __restrict__ gy,
                                const float* __restrict__ minv, const uint8_t* __restrict__ mask,
                                const float* __restrict__ rms, float gain,
                                float* __restrict__ dout) {
  const int r = blockIdx.y;
  const float sc = gain * rms[r];
  gy += (long long)r * 3 * Nn;
  dout += (long long)r * Nn * 16;
  for (long long n = blockIdx.x * (long long)blockDim.x + threadIdx.x; n < Nn;
       n += (long long)gridDim.x * blockDim.x) {
    const unsigned mk = mask[n];
    const float w = sc * minv[n];
    float4* o = reinterpret_cast<float4*>(dout + n * 16);
    o[0] = make_float4(((mk >> 0) & 1u) ? w * gy[3 * n] : 0.f,
                       ((mk >> 1) & 1u) ? w * gy[3 * n + 1] : 0.f,
                       ((mk >> 2) & 1u) ? w * gy[3 * n + 2] : 0.f, 0.f);
    o[1] = make_float4(0.f, 0.f, 0.f, 0.f);
    o[2] = o[1];
    o[3] = o[1];
  }
}

// Training loss per sample: the energy of the error left by the correction
// the V-cycle applies, the best combination alpha v + beta y (k_combine_coef),
// relative to the plain correction v alone and to v.t:
//   l = (E(alpha v + beta y) - E(v)) / (v.t),  E(c) = ||A^-1 t - c||_A^2
// (<= 0; from the five products of k_dot5, without A^-1 t).  By the envelope
// theorem its gradient w.r.t. y is -2 beta (t - alpha A v - beta A y) / (v.t).
__global__ void k_train_gy(long long n3, const float* __restrict__ t, const float* __restrict__ Av,
                           const float* __restrict__ Ay, const double* __restrict__ coef,
                           const double* __restrict__ vt, float* __restrict__ gy) {
  const int r = blockIdx.y;
  const float al = (float)coef[2 * r], be = (float)coef[2 * r + 1];
  const float sc = (float)(-2.0 * coef[2 * r + 1] / vt[r]);
  const long long o = (long long)r * n3;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n3;
       i += (long long)gridDim.x * blockDim.x)
    gy[o + i] = sc * (t[o + i] - al * Av[o + i] - be * Ay[o + i]);
}

// mean over the B samples of l (one work-group); part as written by k_dot5
__global__ void k_train_loss(const double* __restrict__ part, int np, int B,
                             const double* __restrict__ coef, double* __restrict__ loss) {
  __shared__ double sred[4];
  double tot = 0.0;
  for (int r = 0; r < B; ++r) {
    double m[5];
    for (int k = 0; k < 5; ++k) {
      double v = 0.0;
      for (int i = threadIdx.x; i < np; i += blockDim.x) v += part[((long long)r * 5 + k) * np + i];
      m[k] = block_sum(v, sred);
      __syncthreads();
    }
    if (threadIdx.x == 0 && m[3] > 0.0) {
      const double al = coef[2 * r], be = coef[2 * r + 1];
      const double ec = -2.0 * (al * m[3] + be * m[4]) + al * al * m[0] +
                        2.0 * al * be * m[1] + be * be * m[2];
      const double ev = -2.0 * m[3] + m[0];
      tot += (ec - ev) / m[3];
    }
  }
  if (threadIdx.x == 0) *loss = tot / B;
}

// ---------------------------------------------------------------------------
// Parts of uneven size in one launch per operation (vx_solve_parts).
//
// Every part has its own grid and hierarchy.  The parts present at a level
// are packed one after another in that level's buffers, each part's block
// starting on a 64-element boundary (16-byte aligned for the vector paths),
// and a launch covers a range [q0, q1) of them.  Per part and level, a
// PartGrid holds the dims and the offsets of its vectors (3 Nn values), its
// node arrays (mask, inverse diagonal) and its element array (E), plus, on
// level 0, the offsets of its values and elements in the caller's packed
// arrays.  A PartTiles / PartTr holds its tile / transfer grid and its first
// work-group in the launch (a prefix over the level's parts; entry np is the
// total).  A work-group finds its part by binary search over that prefix and
// then runs the single-grid kernel body on the part's own buffers, so each
// part computes exactly what its own solve would.  Given a scalars pointer,
// the work-groups of a frozen (converged) part return at once: a converged
// part costs no bandwidth while the others iterate.
// ---------------------------------------------------------------------------
struct PartGrid {
  int nx, ny, nz;
  long long voff, noff, eoff;   // this level's packed vectors / node arrays / element arrays
  long long uoff, ueoff;        // level 0: offsets in the caller's packed values / elements
};

struct PartTiles {
  int zc, gx, gy, gz;
  int begin;
};

struct PartTr {
  int Lcx, Lcy, Lcz;
  TrGrid g;
  int begin;
  long long src, dst, node;     // input and output vector offsets; node offset of the mask
                                // (and the inverse diagonal) of the output level
};

__device__ __forceinline__ long long part_nodes(const PartGrid& g) {
  return (long long)(g.nx + 1) * (g.ny + 1) * (g.nz + 1);
}

// the part of global work-group vg: the last q in [q0, q1) with t[q].begin <= vg
template <typename D>
__device__ __forceinline__ int part_of(const D* __restrict__ t, int q0, int q1, int vg) {
  int lo = q0, hi = q1 - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (t[mid].begin <= vg) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

template <typename T, int MODE, bool DOT>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 4 : 1)))
void k_tile_parts(TileArgs<T> a, const PartGrid* __restrict__ pg,
                  const PartTiles* __restrict__ pt, int q0, int q1,
                  const Scalars* __restrict__ fz) {
  const int base = pt[q0].begin;
  int v;
  if (!tr_block(pt[q1].begin - base, v)) return;
  const int vg = v + base;
  const int q = part_of(pt, q0, q1, vg);
  if (fz && fz[q].frozen) return;
  const PartGrid g = pg[q];
  const PartTiles t = pt[q];
  TileArgs<T> b = a;
  b.nx = g.nx; b.ny = g.ny; b.nz = g.nz;
  b.Lx = g.nx + 1; b.Ly = g.ny + 1; b.Lz = g.nz + 1;
  b.zc = t.zc; b.gx = t.gx; b.gy = t.gy; b.gz = t.gz;
  b.nb = 1;
  b.E = a.E + g.eoff;
  b.mask = a.mask + g.noff;
  tile_run<T, MODE, DOT>(b, vg - t.begin, vg, a.u + g.voff, a.b ? a.b + g.voff : nullptr,
                         a.up ? a.up + g.voff : nullptr, a.out + g.voff);
}

// restriction of parts [q0, q1) of the fine level into the coarse level; the
// coarse first smoothing step x1 is written for parts q < x1_lim
template <typename T>
__global__ __launch_bounds__(TR_THREADS) void k_restrict_parts(
    const PartTr* __restrict__ pr, int q0, int q1, const T* __restrict__ rf,
    const uint8_t* __restrict__ cmask, T* __restrict__ rc, const T* __restrict__ cminv, T w0,
    T* __restrict__ x1, int x1_lim, const Scalars* __restrict__ fz) {
  const int base = pr[q0].begin;
  int v;
  if (!tr_block(pr[q1].begin - base, v)) return;
  const int vg = v + base;
  const int q = part_of(pr, q0, q1, vg);
  if (fz && fz[q].frozen) return;
  const PartTr t = pr[q];
  restrict_run<T>(t.Lcx, t.Lcy, t.Lcz, t.g, vg - t.begin, rf + t.src, cmask + t.node,
                  rc + t.dst, cminv + t.node, w0, (x1 && q < x1_lim) ? x1 + t.dst : nullptr);
}

template <typename T>
__global__ __launch_bounds__(TR_THREADS) void k_prolong_parts(
    const PartTr* __restrict__ pp, int q0, int q1, const T* __restrict__ ec,
    const uint8_t* __restrict__ fmask, T* __restrict__ x, const Scalars* __restrict__ fz) {
  const int base = pp[q0].begin;
  int v;
  if (!tr_block(pp[q1].begin - base, v)) return;
  const int vg = v + base;
  const int q = part_of(pp, q0, q1, vg);
  if (fz && fz[q].frozen) return;
  const PartTr t = pp[q];
  prolong_run<T>(t.Lcx, t.Lcy, t.Lcz, t.g, vg - t.begin, ec + t.src, fmask + t.node, x + t.dst);
}

// all coarsest-level steps of parts q0 + blockIdx.x, one work-group each
template <typename T>
__global__ __launch_bounds__(1024) void k_coarse_smooth_parts(
    const PartGrid* __restrict__ pg, int q0, const T* __restrict__ E,
    const uint8_t* __restrict__ mask, const T* __restrict__ minv, const T* __restrict__ b,
    T* __restrict__ out, ElemConst<T> K, CoarseCoef cc, const Scalars* __restrict__ fz) {
  const int q = q0 + blockIdx.x;
  if (fz && fz[q].frozen) return;
  const PartGrid g = pg[q];
  coarse_smooth_run<T>(g.nx, g.ny, g.nz, E + g.eoff, mask + g.noff, minv + g.noff, b + g.voff,
                       out + g.voff, K, cc);
}

// The elementwise kernels below run on parts q0 + blockIdx.y.
template <typename T>
__global__ void k_zero_step_parts(const PartGrid* __restrict__ pg, int q0,
                                  const uint8_t* __restrict__ mask, const T* __restrict__ minv,
                                  const T* __restrict__ b, T* __restrict__ out, T c, int vec4,
                                  const Scalars* __restrict__ fz) {
  const int q = q0 + blockIdx.y;
  if (fz && fz[q].frozen) return;
  const PartGrid g = pg[q];
  zero_step_run<T>(part_nodes(g), mask + g.noff, minv + g.noff, b + g.voff, out + g.voff, c, vec4);
}

// r -= alpha Ap, partial r.r, next V-cycle's first step (FP32, level 0, all parts)
__global__ void k_cg_update_x1_parts(const PartGrid* __restrict__ pg,
                                     const Scalars* __restrict__ s, float* __restrict__ r,
                                     const float* __restrict__ Ap, double* __restrict__ partial,
                                     const uint8_t* __restrict__ mask,
                                     const float* __restrict__ minv, float c,
                                     float* __restrict__ x1) {
  __shared__ double sred[4];
  const int q = blockIdx.y;
  if (s[q].frozen) return;
  const PartGrid g = pg[q];
  double acc = cg_update_x1_run(part_nodes(g), (float)s[q].alpha, r + g.voff, Ap + g.voff,
                                mask + g.noff, minv + g.noff, c, x1 + g.voff);
  acc = block_sum(acc, sred);
  if (threadIdx.x == 0) partial[q * gridDim.x + blockIdx.x] = acc;
}

// u += alpha p (the deferred update), then p = z + beta p; a frozen part
// takes only its pending update u += alpha p, and nothing once its alpha is 0
// (every step after the one that froze it)
template <typename T>
__global__ void k_p_update_u_parts(const PartGrid* __restrict__ pg,
                                   const Scalars* __restrict__ s, const T* __restrict__ z,
                                   T* __restrict__ p, T* __restrict__ u) {
  const int q = blockIdx.y;
  if (s[q].frozen && s[q].alpha == 0.0) return;
  const PartGrid g = pg[q];
  const long long n = 3 * part_nodes(g);
  if (s[q].frozen)
    u_flush_run<T>(n, (T)s[q].alpha, p + g.voff, u + g.voff);
  else
    p_update_u_run<T>(n, (T)s[q].alpha, (T)s[q].beta, z + g.voff, p + g.voff, u + g.voff);
}

template <typename T>
__global__ void k_u_flush_parts(const PartGrid* __restrict__ pg, const Scalars* __restrict__ s,
                                const T* __restrict__ p, T* __restrict__ u) {
  const int q = blockIdx.y;
  const PartGrid g = pg[q];
  u_flush_run<T>(3 * part_nodes(g), (T)s[q].alpha, p + g.voff, u + g.voff);
}

template <typename T>
__global__ void k_cg_update_parts(const PartGrid* __restrict__ pg, const Scalars* __restrict__ s,
                                  T* __restrict__ u, T* __restrict__ r, const T* __restrict__ p,
                                  const T* __restrict__ Ap, double* __restrict__ partial) {
  __shared__ double sred[4];
  const int q = blockIdx.y;
  if (s[q].frozen) return;
  const PartGrid g = pg[q];
  double acc = cg_update_run<T>(3 * part_nodes(g), (T)s[q].alpha, u + g.voff, r + g.voff,
                                p + g.voff, Ap + g.voff);
  acc = block_sum(acc, sred);
  if (threadIdx.x == 0) partial[q * gridDim.x + blockIdx.x] = acc;
}

template <typename T>
__global__ void k_p_update_b_parts(const PartGrid* __restrict__ pg,
                                   const Scalars* __restrict__ s, const T* __restrict__ z,
                                   T* __restrict__ p) {
  const int q = blockIdx.y;
  if (s[q].frozen) return;
  const PartGrid g = pg[q];
  p_update_b_run<T>(3 * part_nodes(g), (T)s[q].beta, z + g.voff, p + g.voff);
}

// out = mask * in on level 0 of all parts (and partials of out.out when
// partial is set).  dir 0: in from the caller's packing, out to this level's;
// dir 1: in from this level's packing, out to the caller's; dir 2: both here.
template <typename Tin, typename Tout>
__global__ void k_mask_copy_parts(const PartGrid* __restrict__ pg, int dir,
                                  const uint8_t* __restrict__ mask, const Tin* in, Tout* out,
                                  double* __restrict__ partial) {
  __shared__ double sred[4];
  const int q = blockIdx.y;
  const PartGrid g = pg[q];
  const long long io = dir == 0 ? g.uoff : g.voff;
  const long long oo = dir == 1 ? g.uoff : g.voff;
  double acc = mask_copy_run<Tin, Tout>(part_nodes(g), mask + g.noff, in + io, out + oo);
  if (partial) {
    acc = block_sum(acc, sred);
    if (threadIdx.x == 0) partial[q * gridDim.x + blockIdx.x] = acc;
  }
}

__global__ void k_scale_to_f32_parts(const PartGrid* __restrict__ pg,
                                     const Scalars* __restrict__ s,
                                     const double* __restrict__ r64, float* __restrict__ r32) {
  const int q = blockIdx.y;
  const PartGrid g = pg[q];
  const double nr = sqrt(s[q].tmp);
  const double sc = (nr > 0.0 && s[q].tmp2 == 0.0) ? 1.0 / nr : 0.0;
  const long long n = 3 * part_nodes(g);
  r64 += g.voff;
  r32 += g.voff;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    r32[i] = (float)(r64[i] * sc);
}

__global__ void k_add_scaled_parts(const PartGrid* __restrict__ pg,
                                   const Scalars* __restrict__ s,
                                   const float* __restrict__ d32, double* __restrict__ u64) {
  const int q = blockIdx.y;
  const PartGrid g = pg[q];
  const double sc = sqrt(s[q].tmp);
  const long long n = 3 * part_nodes(g);
  d32 += g.voff;
  u64 += g.voff;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    u64[i] += (double)d32[i] * sc;
}

// per part blockIdx.x: its partials are [pt[q].begin, pt[q + 1].begin) of a
// tile launch, or the n_ew of an elementwise launch (pt == nullptr)
__global__ __launch_bounds__(FIN_THREADS) void k_finalize_parts(
    const double* __restrict__ partial, const PartTiles* __restrict__ pt, int n_ew, int op,
    Scalars* s) {
  const int q = blockIdx.x;
  const int b0 = pt ? pt[q].begin : q * n_ew;
  const int n = pt ? pt[q + 1].begin - b0 : n_ew;
  finalize_run(partial + b0, n, op, s + q);
}

}  // namespace vx
