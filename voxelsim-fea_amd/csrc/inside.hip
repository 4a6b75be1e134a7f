// inside.hip -- voxelisation inside test on the device: single-direction
// crossing-parity ray cast (Moeller-Trumbore) of _inside_mask_numpy,
// core/inside_worker.py:118-162, with the same tolerances and the same
// floating-point operation order (no contraction into FMA), so that the mask
// equals the reference's bit for bit.
//
// Broad phase: every ray has the same direction d, so a ray from p can only
// hit triangles whose projection along d covers p's projection.  Triangles
// are binned into a uniform 2-D grid over the plane orthogonal to d (their
// projected bounding boxes grown by a margin that covers the barycentric
// tolerance), and each point tests only the triangles of its cell.  The cell
// lists are a counting sort (count, exclusive scan, fill); the fill order
// inside a cell varies but a parity count does not depend on it.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

#include "vxfea.h"

#pragma clang fp contract(off)

namespace vxin {

struct Basis {
  double d[3], a[3], b[3];   // ray direction and two unit vectors orthogonal to it
};

struct Tri {
  double v0[3], e1[3], e2[3], pvec[3];
  double inv_det;
  int parallel;
};

struct Binning {
  double lo_a, lo_b, inv_a, inv_b;   // cell = floor((x - lo) * inv)
  int ga, gb;
};

__device__ __forceinline__ void cross(const double* a, const double* b, double* c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}
__device__ __forceinline__ double dot3(const double* a, const double* b) {
  return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2];
}

// per triangle: Moeller-Trumbore set-up terms and the grown projected bbox
__global__ void k_tri_setup(const double* __restrict__ V, const int64_t* __restrict__ F,
                            long long nt, Basis B, Tri* __restrict__ tri,
                            double* __restrict__ box) {
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < nt;
       t += (long long)gridDim.x * blockDim.x) {
    const double* p0 = V + 3 * F[3 * t];
    const double* p1 = V + 3 * F[3 * t + 1];
    const double* p2 = V + 3 * F[3 * t + 2];
    Tri T;
    for (int c = 0; c < 3; ++c) {
      T.v0[c] = p0[c];
      T.e1[c] = p1[c] - p0[c];
      T.e2[c] = p2[c] - p0[c];
    }
    cross(B.d, T.e2, T.pvec);
    const double det = dot3(T.e1, T.pvec);
    T.parallel = fabs(det) < 1e-9;
    T.inv_det = T.parallel ? 0.0 : 1.0 / det;
    tri[t] = T;
    double lo_a = INFINITY, hi_a = -INFINITY, lo_b = INFINITY, hi_b = -INFINITY;
    const double* pts[3] = {p0, p1, p2};
    for (int k = 0; k < 3; ++k) {
      const double pa = dot3(pts[k], B.a), pb = dot3(pts[k], B.b);
      lo_a = fmin(lo_a, pa);
      hi_a = fmax(hi_a, pa);
      lo_b = fmin(lo_b, pb);
      hi_b = fmax(hi_b, pb);
    }
    // barycentric tolerance 1e-7 of the triangle's extent (x10), plus
    // rounding of the projections themselves
    const double ext = fmax(hi_a - lo_a, hi_b - lo_b);
    const double mag = fmax(fmax(fabs(lo_a), fabs(hi_a)), fmax(fabs(lo_b), fabs(hi_b)));
    const double m = 1e-6 * ext + 1e-12 * mag;
    // parallel triangles never count: an empty box keeps them out of the lists
    box[4 * t + 0] = T.parallel ? INFINITY : lo_a - m;
    box[4 * t + 1] = T.parallel ? -INFINITY : hi_a + m;
    box[4 * t + 2] = T.parallel ? INFINITY : lo_b - m;
    box[4 * t + 3] = T.parallel ? -INFINITY : hi_b + m;
  }
}

__device__ __forceinline__ bool cell_range(const double* bx, const Binning& G, int* a0, int* a1,
                                           int* b0, int* b1) {
  if (!(bx[0] <= bx[1]) || !(bx[2] <= bx[3])) return false;
  *a0 = max(0, (int)floor((bx[0] - G.lo_a) * G.inv_a));
  *a1 = min(G.ga - 1, (int)floor((bx[1] - G.lo_a) * G.inv_a));
  *b0 = max(0, (int)floor((bx[2] - G.lo_b) * G.inv_b));
  *b1 = min(G.gb - 1, (int)floor((bx[3] - G.lo_b) * G.inv_b));
  return *a0 <= *a1 && *b0 <= *b1;
}

__global__ void k_count(const double* __restrict__ box, long long nt, Binning G,
                        unsigned int* __restrict__ cnt) {
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < nt;
       t += (long long)gridDim.x * blockDim.x) {
    int a0, a1, b0, b1;
    if (!cell_range(box + 4 * t, G, &a0, &a1, &b0, &b1)) continue;
    for (int j = b0; j <= b1; ++j)
      for (int i = a0; i <= a1; ++i) atomicAdd(&cnt[(long long)j * G.ga + i], 1u);
  }
}

__global__ void k_fill(const double* __restrict__ box, long long nt, Binning G,
                       const unsigned int* __restrict__ off, unsigned int* __restrict__ cur,
                       unsigned int* __restrict__ list) {
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < nt;
       t += (long long)gridDim.x * blockDim.x) {
    int a0, a1, b0, b1;
    if (!cell_range(box + 4 * t, G, &a0, &a1, &b0, &b1)) continue;
    for (int j = b0; j <= b1; ++j)
      for (int i = a0; i <= a1; ++i) {
        const long long c = (long long)j * G.ga + i;
        list[off[c] + atomicAdd(&cur[c], 1u)] = (unsigned int)t;
      }
  }
}

// point generation of core/inside_worker.py:87-111: target 0 voxel centres
// (element order, x slowest), 1 nodes (x fastest); -1 explicit points
__device__ __forceinline__ void grid_point(long long i, int target, int nx, int ny, int nz,
                                           const double* org, double vs, const double* P,
                                           double* p) {
  if (target == 0) {
    const long long iz = i % nz, iy = (i / nz) % ny, ix = i / ((long long)nz * ny);
    p[0] = ((double)ix + 0.5) * vs + org[0];
    p[1] = ((double)iy + 0.5) * vs + org[1];
    p[2] = ((double)iz + 0.5) * vs + org[2];
  } else if (target == 1) {
    const long long lx = nx + 1, ly = ny + 1;
    const long long ix = i % lx, iy = (i / lx) % ly, iz = i / (lx * ly);
    p[0] = (double)ix * vs + org[0];
    p[1] = (double)iy * vs + org[1];
    p[2] = (double)iz * vs + org[2];
  } else {
    p[0] = P[3 * i];
    p[1] = P[3 * i + 1];
    p[2] = P[3 * i + 2];
  }
}

struct GridPts {
  int target, nx, ny, nz;
  double org[3], vs;
};

__global__ void k_query(long long np, GridPts gp, const double* __restrict__ P, Basis B,
                        Binning G, const Tri* __restrict__ tri,
                        const unsigned int* __restrict__ off,
                        const unsigned int* __restrict__ list, uint8_t* __restrict__ out) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < np;
       i += (long long)gridDim.x * blockDim.x) {
    double p[3];
    grid_point(i, gp.target, gp.nx, gp.ny, gp.nz, gp.org, gp.vs, P, p);
    const double pa = dot3(p, B.a), pb = dot3(p, B.b);
    const double fa = floor((pa - G.lo_a) * G.inv_a), fb = floor((pb - G.lo_b) * G.inv_b);
    unsigned int hits = 0;
    if (fa >= 0.0 && fa < (double)G.ga && fb >= 0.0 && fb < (double)G.gb) {
      const long long c = (long long)fb * G.ga + (long long)fa;
      for (unsigned int k = off[c]; k < off[c + 1]; ++k) {
        const Tri& T = tri[list[k]];
        double tv[3], q[3];
        tv[0] = p[0] - T.v0[0];
        tv[1] = p[1] - T.v0[1];
        tv[2] = p[2] - T.v0[2];
        const double u = dot3(tv, T.pvec) * T.inv_det;
        cross(tv, T.e1, q);
        const double v = dot3(B.d, q) * T.inv_det;
        const double t = dot3(T.e2, q) * T.inv_det;
        hits += (!T.parallel && u >= -1e-7 && v >= -1e-7 && u + v <= 1.0 + 1e-7 && t > 1e-7);
      }
    }
    out[i] = (uint8_t)(hits & 1u);
  }
}

__global__ void k_bbox_red(const double* __restrict__ box, long long nt, double* __restrict__ out) {
  // one work-group: [min a, max a, min b, max b] over non-empty boxes
  __shared__ double r[4][256];
  double lo_a = INFINITY, hi_a = -INFINITY, lo_b = INFINITY, hi_b = -INFINITY;
  for (long long t = threadIdx.x; t < nt; t += blockDim.x) {
    const double* bx = box + 4 * t;
    if (!(bx[0] <= bx[1])) continue;
    lo_a = fmin(lo_a, bx[0]);
    hi_a = fmax(hi_a, bx[1]);
    lo_b = fmin(lo_b, bx[2]);
    hi_b = fmax(hi_b, bx[3]);
  }
  r[0][threadIdx.x] = lo_a;
  r[1][threadIdx.x] = hi_a;
  r[2][threadIdx.x] = lo_b;
  r[3][threadIdx.x] = hi_b;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      r[0][threadIdx.x] = fmin(r[0][threadIdx.x], r[0][threadIdx.x + o]);
      r[1][threadIdx.x] = fmax(r[1][threadIdx.x], r[1][threadIdx.x + o]);
      r[2][threadIdx.x] = fmin(r[2][threadIdx.x], r[2][threadIdx.x + o]);
      r[3][threadIdx.x] = fmax(r[3][threadIdx.x], r[3][threadIdx.x + o]);
    }
    __syncthreads();
  }
  if (threadIdx.x < 4) out[threadIdx.x] = r[threadIdx.x][0];
}

int grid1(long long n) {
  return (int)std::max<long long>(1, std::min<long long>((n + 255) / 256, 8192));
}

thread_local std::string g_err;

int fail(const std::string& m) {
  g_err = m;
  return VX_ERR_HIP;
}

}  // namespace vxin

using namespace vxin;

static int inside_impl(void* stream, const double* V, int64_t nv, const int64_t* F, int64_t nt,
                       GridPts gp, const double* P, int64_t np, const double* dir,
                       uint8_t* out) {
  hipStream_t st = (hipStream_t)stream;
  if ((nt > 0 && (!V || !F)) || (np > 0 && !out) || nv < 0 || nt < 0 || np < 0 || !dir) {
    g_err = "bad inside-test arguments";
    return VX_ERR_ARG;
  }
  if (np == 0) return VX_OK;
  std::vector<void*> mine;
  auto release = [&]() {
    (void)hipStreamSynchronize(st);
    for (void* q : mine) (void)hipFree(q);
  };
  auto get = [&](size_t bytes) -> void* {
    void* q = nullptr;
    if (hipMalloc(&q, std::max<size_t>(bytes, 16)) != hipSuccess) return nullptr;
    mine.push_back(q);
    return q;
  };
  if (nt == 0) {
    if (hipMemsetAsync(out, 0, (size_t)np, st) != hipSuccess) return fail("memset");
    return hipStreamSynchronize(st) == hipSuccess ? VX_OK : fail("sync");
  }
  // basis: d and two orthonormal vectors across it
  Basis B;
  for (int c = 0; c < 3; ++c) B.d[c] = dir[c];
  const double ref[3] = {0.0, 0.0, 1.0};
  double a[3] = {B.d[1] * ref[2] - B.d[2] * ref[1], B.d[2] * ref[0] - B.d[0] * ref[2],
                 B.d[0] * ref[1] - B.d[1] * ref[0]};
  double na = std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
  if (na < 1e-3) {   // d nearly along z: use x instead
    const double rx[3] = {1.0, 0.0, 0.0};
    a[0] = B.d[1] * rx[2] - B.d[2] * rx[1];
    a[1] = B.d[2] * rx[0] - B.d[0] * rx[2];
    a[2] = B.d[0] * rx[1] - B.d[1] * rx[0];
    na = std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
  }
  for (int c = 0; c < 3; ++c) B.a[c] = a[c] / na;
  B.b[0] = B.d[1] * B.a[2] - B.d[2] * B.a[1];
  B.b[1] = B.d[2] * B.a[0] - B.d[0] * B.a[2];
  B.b[2] = B.d[0] * B.a[1] - B.d[1] * B.a[0];

  Tri* tri = (Tri*)get(sizeof(Tri) * nt);
  double* box = (double*)get(sizeof(double) * 4 * nt);
  double* bb = (double*)get(sizeof(double) * 4);
  if (!tri || !box || !bb) { release(); g_err = "allocation"; return VX_ERR_NOMEM; }
  hipLaunchKernelGGL(k_tri_setup, dim3(grid1(nt)), dim3(256), 0, st, V, F, (long long)nt, B, tri,
                     box);
  hipLaunchKernelGGL(k_bbox_red, dim3(1), dim3(256), 0, st, box, (long long)nt, bb);
  double hb[4];
  if (hipGetLastError() != hipSuccess ||
      hipMemcpyAsync(hb, bb, sizeof hb, hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess) {
    release();
    return fail("triangle set-up");
  }
  Binning G;
  if (!(hb[0] <= hb[1])) {   // every triangle parallel to d: nothing is hit
    G.ga = G.gb = 1;
    G.lo_a = G.lo_b = 0.0;
    G.inv_a = G.inv_b = 0.0;
  } else {
    const double ea = std::max(hb[1] - hb[0], 1e-300), eb = std::max(hb[3] - hb[2], 1e-300);
    const double side = std::sqrt((double)nt) * 2.0;
    const double mx = std::max(ea, eb);
    G.ga = (int)std::min(2048.0, std::max(1.0, std::ceil(side * ea / mx)));
    G.gb = (int)std::min(2048.0, std::max(1.0, std::ceil(side * eb / mx)));
    G.lo_a = hb[0];
    G.lo_b = hb[2];
    G.inv_a = G.ga / ea;
    G.inv_b = G.gb / eb;
  }
  const long long ncell = (long long)G.ga * G.gb;
  unsigned int* cnt = (unsigned int*)get(sizeof(unsigned int) * (ncell + 1));
  unsigned int* off = (unsigned int*)get(sizeof(unsigned int) * (ncell + 1));
  unsigned int* cur = (unsigned int*)get(sizeof(unsigned int) * (ncell + 1));
  if (!cnt || !off || !cur) { release(); g_err = "allocation"; return VX_ERR_NOMEM; }
  if (hipMemsetAsync(cnt, 0, sizeof(unsigned int) * (ncell + 1), st) != hipSuccess ||
      hipMemsetAsync(cur, 0, sizeof(unsigned int) * (ncell + 1), st) != hipSuccess) {
    release();
    return fail("memset");
  }
  hipLaunchKernelGGL(k_count, dim3(grid1(nt)), dim3(256), 0, st, box, (long long)nt, G, cnt);
  size_t tmpb = 0;
  if (hipcub::DeviceScan::ExclusiveSum(nullptr, tmpb, cnt, off, (int)(ncell + 1), st) !=
      hipSuccess) {
    release();
    return fail("scan size");
  }
  void* tmp = get(tmpb);
  if (!tmp || hipcub::DeviceScan::ExclusiveSum(tmp, tmpb, cnt, off, (int)(ncell + 1), st) !=
                  hipSuccess) {
    release();
    return fail("scan");
  }
  unsigned int total = 0;
  if (hipMemcpyAsync(&total, off + ncell, sizeof total, hipMemcpyDeviceToHost, st) !=
          hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess) {
    release();
    return fail("scan readback");
  }
  unsigned int* list = (unsigned int*)get(sizeof(unsigned int) * std::max(total, 1u));
  if (!list) { release(); g_err = "allocation"; return VX_ERR_NOMEM; }
  hipLaunchKernelGGL(k_fill, dim3(grid1(nt)), dim3(256), 0, st, box, (long long)nt, G, off, cur,
                     list);
  hipLaunchKernelGGL(k_query, dim3(grid1(np)), dim3(256), 0, st, (long long)np, gp, P, B, G, tri,
                     off, list, out);
  const bool ok = hipGetLastError() == hipSuccess && hipStreamSynchronize(st) == hipSuccess;
  release();
  return ok ? VX_OK : fail("query");
}

extern "C" {

int vx_inside_mask(void* stream, const double* verts, int64_t nv, const int64_t* faces,
                   int64_t nt, const double* points, int64_t np, const double* dir_host,
                   uint8_t* out) {
  if (!points && np > 0) {
    g_err = "null points";
    return VX_ERR_ARG;
  }
  GridPts gp{-1, 0, 0, 0, {0.0, 0.0, 0.0}, 0.0};
  return inside_impl(stream, verts, nv, faces, nt, gp, points, np, dir_host, out);
}

int vx_inside_grid(void* stream, const double* verts, int64_t nv, const int64_t* faces,
                   int64_t nt, int32_t nx, int32_t ny, int32_t nz, const double* origin_host,
                   double vsize, int32_t target, const double* dir_host, uint8_t* out) {
  if (nx < 1 || ny < 1 || nz < 1 || (target != 0 && target != 1) || !origin_host) {
    g_err = "bad grid";
    return VX_ERR_ARG;
  }
  GridPts gp{target, nx, ny, nz, {origin_host[0], origin_host[1], origin_host[2]}, vsize};
  const int64_t np = target == 0 ? (int64_t)nx * ny * nz
                                 : (int64_t)(nx + 1) * (ny + 1) * (nz + 1);
  return inside_impl(stream, verts, nv, faces, nt, gp, nullptr, np, dir_host, out);
}

const char* vx_inside_last_error(void) { return g_err.c_str(); }

}  // extern "C"
