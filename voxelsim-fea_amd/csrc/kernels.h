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

#include "elem.h"

namespace vx {

constexpr int TILE = 16;              // element columns per tile edge
constexpr int OWN = TILE - 1;         // owned node columns per tile edge
constexpr int FOOT = TILE + 1;        // node footprint per tile edge
constexpr int NT = TILE * TILE;       // threads per tile work-group
constexpr int FOOTVALS = FOOT * FOOT * 3;
constexpr int PRE = (FOOTVALS + NT - 1) / NT;

enum { M_MATVEC = 0, M_RESID = 1, M_JACOBI = 2, M_CHEB = 3 };

template <typename T>
struct TileArgs {
  int nx, ny, nz;        // elements
  int Lx, Ly, Lz;        // nodes
  int zc;                // owned node planes per work-group
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
};

template <typename T>
__device__ inline T shfl_up1(T v) {
  return __shfl_up(v, 1, 64);
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
// stages one node plane of the 17x17 footprint in LDS (coalesced rows, the
// next plane prefetched in registers), every thread applies its element of
// the current layer in the Hadamard basis, the 2D modes of the forces of the
// bottom plane are completed with the top-plane half kept from the previous
// layer, and the node sums over the four element columns are formed by a
// lane shuffle in x and one LDS exchange in y.  No atomics: every node's sum
// is formed in a fixed order, so the result is deterministic.
//
// MODE  M_MATVEC: out = mask ? K u : 0                (core/multigrid.py:318-341)
//       M_RESID : out = b - (mask ? K u : 0)          (:431)
//       M_JACOBI: out = u + omega*(mask ? (b-Ku)/d : 0), d = kd*sum E
//                                                     (:367-402, :300-304)
//       M_CHEB  : out = u + mask ? cu (u - up) + cr (b-Ku)/d : 0, one step of
//                 the Chebyshev iteration on D^-1 A (three-term recurrence)
// DOT   adds the work-group partial of  sum out*u (MATVEC) or sum out*b
//       (JACOBI) to partial[block].
// ---------------------------------------------------------------------------
template <typename T, int MODE, bool DOT>
__global__ __launch_bounds__(NT) void k_tile(TileArgs<T> a) {
  __shared__ T sP[2][3][FOOT][FOOT];
  __shared__ T sQ[4][TILE][TILE + 1];
  __shared__ double sred[NT / 64];

  const int tid = threadIdx.x;
  const int tx = tid & (TILE - 1), ty = tid / TILE;
  const int X0 = blockIdx.x * OWN, Y0 = blockIdx.y * OWN;
  const int Z0 = blockIdx.z * a.zc;
  const int Z1 = min(Z0 + a.zc, a.Lz);
  const int ex = X0 - 1 + tx, ey = Y0 - 1 + ty;
  const bool col_in = ex >= 0 && ex < a.nx && ey >= 0 && ey < a.ny;
  const bool owner = tx >= 1 && ty >= 1 && ex < a.Lx && ey < a.Ly;
  const long long plane = (long long)a.Lx * a.Ly;

  T pre[PRE];
  auto fetch = [&](int pz) {
#pragma unroll
    for (int k = 0; k < PRE; ++k) {
      const int i = tid + k * NT;
      T v = T(0);
      if (i < FOOTVALS) {
        const int row = i / (FOOT * 3);
        const int col = i - row * (FOOT * 3);
        const int x = X0 - 1 + col / 3;
        const int y = Y0 - 1 + row;
        if (pz >= 0 && pz < a.Lz && x >= 0 && x < a.Lx && y >= 0 && y < a.Ly)
          v = a.u[((long long)pz * plane + (long long)y * a.Lx + x) * 3 + (col - (col / 3) * 3)];
      }
      pre[k] = v;
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int k = 0; k < PRE; ++k) {
      const int i = tid + k * NT;
      if (i < FOOTVALS) {
        const int row = i / (FOOT * 3);
        const int col = i - row * (FOOT * 3);
        sP[buf][col - (col / 3) * 3][row][col / 3] = pre[k];
      }
    }
  };

  T Pb[3][4], acc[3][4];
  T ub0[3], ut0[3];
  T Eprev = T(0);
  double dacc = 0.0;

  // prologue: bottom plane of the first layer (Z0-1)
  fetch(Z0 - 1);
  store(0);
  __syncthreads();
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    had2(sP[0][c][ty][tx], sP[0][c][ty][tx + 1], sP[0][c][ty + 1][tx], sP[0][c][ty + 1][tx + 1],
         Pb[c]);
    ut0[c] = sP[0][c][ty][tx];
#pragma unroll
    for (int m = 0; m < 4; ++m) acc[c][m] = T(0);
  }
  fetch(Z0);
  int buf = 1;

  for (int iz = Z0 - 1; iz < Z1; ++iz) {
    store(buf);
    __syncthreads();
    if (iz + 2 <= Z1) fetch(iz + 2);

    T Pt[3][4];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      ub0[c] = ut0[c];
      const T u00 = sP[buf][c][ty][tx];
      had2(u00, sP[buf][c][ty][tx + 1], sP[buf][c][ty + 1][tx], sP[buf][c][ty + 1][tx + 1], Pt[c]);
      ut0[c] = u00;
    }
    const T Ee = (col_in && iz >= 0 && iz < a.nz)
                     ? a.E[((long long)iz * a.ny + ey) * a.nx + ex]
                     : T(0);
    T Yb[3][4], Yt[3][4];
    elem_modes(Pb, Pt, Ee, a.K, Yb, Yt);
    const T Es = Eprev + Ee;
    Eprev = Ee;

    T A0[3], A1[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      T Y[4];
#pragma unroll
      for (int m = 0; m < 4; ++m) {
        Y[m] = acc[c][m] + Yb[c][m];
        acc[c][m] = Yt[c][m];
        Pb[c][m] = Pt[c][m];
      }
      T y00, y10, y01, y11;
      ihad2(Y, y00, y10, y01, y11);
      A0[c] = y00 + shfl_up1(y10);
      A1[c] = y01 + shfl_up1(y11);
      sQ[c][ty][tx] = A1[c];
    }
    const T EX = Es + shfl_up1(Es);
    sQ[3][ty][tx] = EX;
    __syncthreads();

    if (iz >= Z0 && owner) {
      const long long n = ((long long)iz * a.Ly + ey) * a.Lx + ex;
      const unsigned mk = a.mask[n];
      T y[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) y[c] = A0[c] + sQ[c][ty - 1][tx];
      if (MODE == M_MATVEC) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const T o = ((mk >> c) & 1u) ? y[c] : T(0);
          a.out[3 * n + c] = o;
          if (DOT) dacc += (double)o * (double)ub0[c];
        }
      } else if (MODE == M_RESID) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const T o = a.b[3 * n + c] - (((mk >> c) & 1u) ? y[c] : T(0));
          a.out[3 * n + c] = o;
          if (DOT) dacc += (double)o * (double)o;
        }
      } else if (MODE == M_JACOBI) {
        const T d = a.kd * (EX + sQ[3][ty - 1][tx]);
        const T minv = T(1) / (d == T(0) ? T(1) : d);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const T bb = a.b[3 * n + c];
          const T o = ub0[c] + a.omega * (((mk >> c) & 1u) ? minv * (bb - y[c]) : T(0));
          a.out[3 * n + c] = o;
          if (DOT) dacc += (double)o * (double)bb;
        }
      } else {
        const T d = a.kd * (EX + sQ[3][ty - 1][tx]);
        const T minv = T(1) / (d == T(0) ? T(1) : d);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const T bb = a.b[3 * n + c];
          const T du = a.up ? ub0[c] - a.up[3 * n + c] : ub0[c];
          const T o = ub0[c] + (((mk >> c) & 1u) ? a.cu * du + a.cr * (minv * (bb - y[c])) : T(0));
          a.out[3 * n + c] = o;
          if (DOT) dacc += (double)o * (double)bb;
        }
      }
    }
    buf ^= 1;
  }

  if (DOT) {
    const double s = block_sum(dacc, sred);
    if (tid == 0)
      a.partial[(blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x] = s;
  }
}

// First damped-Jacobi sweep from a zero initial guess (core/multigrid.py:430
// starts from zeros_like): A*0 = 0, so out = omega * mask * b / d.
template <typename T>
__global__ void k_jacobi_zero(int nx, int ny, int nz, const T* __restrict__ E,
                              const uint8_t* __restrict__ mask, const T* __restrict__ b,
                              T* __restrict__ out, T omega, T kd) {
  const int Lx = nx + 1, Ly = ny + 1, Lz = nz + 1;
  const long long Nn = (long long)Lx * Ly * Lz;
  for (long long n = blockIdx.x * (long long)blockDim.x + threadIdx.x; n < Nn;
       n += (long long)gridDim.x * blockDim.x) {
    const int ix = (int)(n % Lx);
    const int iy = (int)((n / Lx) % Ly);
    const int iz = (int)(n / ((long long)Lx * Ly));
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
    const T d = kd * es;
    const T minv = T(1) / (d == T(0) ? T(1) : d);
    const unsigned mk = mask[n];
#pragma unroll
    for (int c = 0; c < 3; ++c)
      out[3 * n + c] = omega * (((mk >> c) & 1u) ? minv * b[3 * n + c] : T(0));
  }
}

// Inverse Jacobi diagonal per DOF (diagnostic / parity only).
template <typename T>
__global__ void k_inv_diag(int nx, int ny, int nz, const T* __restrict__ E, T kd,
                           double* __restrict__ minv_out) {
  const int Lx = nx + 1, Ly = ny + 1, Lz = nz + 1;
  const long long Nn = (long long)Lx * Ly * Lz;
  for (long long n = blockIdx.x * (long long)blockDim.x + threadIdx.x; n < Nn;
       n += (long long)gridDim.x * blockDim.x) {
    const int ix = (int)(n % Lx);
    const int iy = (int)((n / Lx) % Ly);
    const int iz = (int)(n / ((long long)Lx * Ly));
    T es = T(0);
    for (int dz = -1; dz <= 0; ++dz)
      for (int dy = -1; dy <= 0; ++dy)
        for (int dx = -1; dx <= 0; ++dx) {
          const int x = ix + dx, y = iy + dy, z = iz + dz;
          if (x >= 0 && x < nx && y >= 0 && y < ny && z >= 0 && z < nz)
            es += E[((long long)z * ny + y) * nx + x];
        }
    const T d = kd * es;
    const double mi = 1.0 / (double)(d == T(0) ? T(1) : d);
    for (int c = 0; c < 3; ++c) minv_out[3 * n + c] = mi;
  }
}

// R = P^T (27-point full weighting, unscaled) followed by the coarse mask
// (core/multigrid.py:404-416).  One thread per coarse node.
template <typename T>
__global__ void k_restrict(int Lcx, int Lcy, int Lcz, const T* __restrict__ rf,
                           const uint8_t* __restrict__ cmask, T* __restrict__ rc) {
  const int Lfx = 2 * Lcx - 1, Lfy = 2 * Lcy - 1, Lfz = 2 * Lcz - 1;
  const long long Nc = (long long)Lcx * Lcy * Lcz;
  for (long long J = blockIdx.x * (long long)blockDim.x + threadIdx.x; J < Nc;
       J += (long long)gridDim.x * blockDim.x) {
    const int jx = (int)(J % Lcx);
    const int jy = (int)((J / Lcx) % Lcy);
    const int jz = (int)(J / ((long long)Lcx * Lcy));
    T s[3] = {T(0), T(0), T(0)};
    for (int dz = -1; dz <= 1; ++dz) {
      const int z = 2 * jz + dz;
      if (z < 0 || z >= Lfz) continue;
      const T wz = dz == 0 ? T(1) : T(0.5);
      for (int dy = -1; dy <= 1; ++dy) {
        const int y = 2 * jy + dy;
        if (y < 0 || y >= Lfy) continue;
        const T wzy = wz * (dy == 0 ? T(1) : T(0.5));
        const long long rowb = ((long long)z * Lfy + y) * Lfx;
#pragma unroll
        for (int dx = -1; dx <= 1; ++dx) {
          const int x = 2 * jx + dx;
          if (x < 0 || x >= Lfx) continue;
          const T w = wzy * (dx == 0 ? T(1) : T(0.5));
          const long long i = (rowb + x) * 3;
          s[0] += w * rf[i];
          s[1] += w * rf[i + 1];
          s[2] += w * rf[i + 2];
        }
      }
    }
    const unsigned mk = cmask[J];
#pragma unroll
    for (int c = 0; c < 3; ++c) rc[3 * J + c] = ((mk >> c) & 1u) ? s[c] : T(0);
  }
}

// out = x + mask * (P e) with trilinear P (core/multigrid.py:418-424, 434).
// x may alias out.
template <typename T>
__global__ void k_prolong_add(int Lfx, int Lfy, int Lfz, const T* __restrict__ ec,
                              const uint8_t* __restrict__ fmask, const T* x, T* out) {
  const int Lcx = (Lfx + 1) / 2, Lcy = (Lfy + 1) / 2;
  const long long Nf = (long long)Lfx * Lfy * Lfz;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < Nf;
       i += (long long)gridDim.x * blockDim.x) {
    const int ix = (int)(i % Lfx);
    const int iy = (int)((i / Lfx) % Lfy);
    const int iz = (int)(i / ((long long)Lfx * Lfy));
    const int cx = ix >> 1, cy = iy >> 1, cz = iz >> 1;
    const int nxs = (ix & 1) ? 2 : 1, nys = (iy & 1) ? 2 : 1, nzs = (iz & 1) ? 2 : 1;
    const T wx = (ix & 1) ? T(0.5) : T(1), wy = (iy & 1) ? T(0.5) : T(1),
            wz = (iz & 1) ? T(0.5) : T(1);
    T s[3] = {T(0), T(0), T(0)};
    for (int kz = 0; kz < nzs; ++kz)
      for (int ky = 0; ky < nys; ++ky)
        for (int kx = 0; kx < nxs; ++kx) {
          const long long J = ((long long)(cz + kz) * Lcy + (cy + ky)) * Lcx + (cx + kx);
          const T w = wx * wy * wz;
          s[0] += w * ec[3 * J];
          s[1] += w * ec[3 * J + 1];
          s[2] += w * ec[3 * J + 2];
        }
    const unsigned mk = fmask[i];
#pragma unroll
    for (int c = 0; c < 3; ++c) out[3 * i + c] = x[3 * i + c] + (((mk >> c) & 1u) ? s[c] : T(0));
  }
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
};

enum {
  F_SET_TMP = 0,    // s.tmp = sum
  F_FNORM = 1,      // s.fnorm = sqrt(sum)
  F_RZ0 = 2,        // s.rz = sum
  F_ALPHA = 3,      // s.pAp = sum; s.alpha = sum == 0 ? 0 : rz / sum
  F_RATIO = 4,      // s.rr = sum; s.ratio = sqrt(sum) / fnorm
  F_BETA = 5        // s.rz_new = sum; s.beta = rz == 0 ? 0 : sum / rz; s.rz = sum
};

__global__ void k_finalize(const double* __restrict__ partial, int n, int op, Scalars* s) {
  __shared__ double sred[4];
  double v = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) v += partial[i];
  v = block_sum(v, sred);
  if (threadIdx.x == 0) {
    switch (op) {
      case F_SET_TMP: s->tmp = v; break;
      case F_FNORM: s->fnorm = sqrt(v); break;
      case F_RZ0: s->rz = v; break;
      case F_ALPHA:
        s->pAp = v;
        s->alpha = (v == 0.0) ? 0.0 : s->rz / v;
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
    }
  }
}

__global__ void k_set_fnorm(Scalars* s, double v) { s->fnorm = v; }

// u += alpha p ; r -= alpha Ap ; partial sum r.r
template <typename T>
__global__ void k_cg_update(long long n, const Scalars* __restrict__ s, T* __restrict__ u,
                            T* __restrict__ r, const T* __restrict__ p, const T* __restrict__ Ap,
                            double* __restrict__ partial) {
  __shared__ double sred[4];
  const double alpha = s->alpha;
  const T a = (T)alpha;
  double acc = 0.0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    u[i] = u[i] + a * p[i];
    const T rn = r[i] - a * Ap[i];
    r[i] = rn;
    acc += (double)rn * (double)rn;
  }
  acc = block_sum(acc, sred);
  if (threadIdx.x == 0) partial[blockIdx.x] = acc;
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
__global__ void k_mask_copy(long long nn, const uint8_t* __restrict__ mask, const Tin* __restrict__ in,
                            Tout* __restrict__ out, double* __restrict__ partial) {
  __shared__ double sred[4];
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
  if (partial) {
    acc = block_sum(acc, sred);
    if (threadIdx.x == 0) partial[blockIdx.x] = acc;
  }
}

// partial sum of a.b
template <typename T>
__global__ void k_dot(long long n, const T* __restrict__ a, const T* __restrict__ b,
                      double* __restrict__ partial) {
  __shared__ double sred[4];
  double acc = 0.0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    acc += (double)a[i] * (double)b[i];
  acc = block_sum(acc, sred);
  if (threadIdx.x == 0) partial[blockIdx.x] = acc;
}

// r32 = (float)(r64 * scale) with scale = 1/||r64|| from s->tmp = ||r64||^2
__global__ void k_scale_to_f32(long long n, const Scalars* __restrict__ s,
                               const double* __restrict__ r64, float* __restrict__ r32) {
  const double sc = 1.0 / sqrt(s->tmp);
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    r32[i] = (float)(r64[i] * sc);
}

// u64 += d32 * ||r64|| (undo the inner scaling)
__global__ void k_add_scaled(long long n, const Scalars* __restrict__ s,
                             const float* __restrict__ d32, double* __restrict__ u64) {
  const double sc = sqrt(s->tmp);
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

// One monotone dilation pass, in place: an active element becomes reached
// when a face neighbour is.  Lines along z are swept forward and back so a
// pass propagates far; the fixed point (the 6-connected components holding a
// seed) is independent of the order.
__global__ void k_island_grow(int nx, int ny, int nz, const uint8_t* __restrict__ active,
                              uint8_t* reached, int* changed) {
  const long long lines = (long long)nx * ny;
  for (long long l = blockIdx.x * (long long)blockDim.x + threadIdx.x; l < lines;
       l += (long long)gridDim.x * blockDim.x) {
    const int iy = (int)(l % ny);
    const int ix = (int)(l / ny);
    const long long base = ((long long)ix * ny + iy) * nz;
    int any = 0;
    for (int pass = 0; pass < 2; ++pass) {
      for (int k = 0; k < nz; ++k) {
        const int iz = pass == 0 ? k : nz - 1 - k;
        const long long e = base + iz;
        if (!active[e] || reached[e]) continue;
        bool r = false;
        if (iz > 0 && reached[e - 1]) r = true;
        if (iz + 1 < nz && reached[e + 1]) r = true;
        if (iy > 0 && reached[e - nz]) r = true;
        if (iy + 1 < ny && reached[e + nz]) r = true;
        if (ix > 0 && reached[e - (long long)ny * nz]) r = true;
        if (ix + 1 < nx && reached[e + (long long)ny * nz]) r = true;
        if (r) {
          reached[e] = 1;
          any = 1;
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

}  // namespace vx
