// post.hip -- post-solve field kernels: the von Mises / safety-factor /
// displacement fields the reference's solve worker derives from u
// (core/solve_worker.py:358-431), on the device.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>

#include "vxfea.h"

namespace vxpost {

// max / min of non-negative doubles through their bit patterns (monotone for
// x >= 0), so the reductions are exact and order independent
__device__ inline void atomic_max_nonneg(unsigned long long* a, double v) {
  atomicMax(a, (unsigned long long)__double_as_longlong(v));
}
__device__ inline void atomic_min_nonneg(unsigned long long* a, double v) {
  atomicMin(a, (unsigned long long)__double_as_longlong(v));
}

// Per element (reference order): stress in Pa (vm / vsize), safety factor
// clip(yield / max(vm, eps), 0, cap) on active elements, displacement as the
// mean of the 8 corner-node |u| (core/solve_worker.py:363-397).  Also the
// sort keys for the stress percentile (inactive -> -1) and the maxima.
__global__ void k_post_elem(int nx, int ny, int nz, const double* __restrict__ u,
                            const double* __restrict__ vm, const uint8_t* __restrict__ active,
                            double vsize, double yield, double cap, float* __restrict__ stress3d,
                            float* __restrict__ safety3d, float* __restrict__ disp3d,
                            double* __restrict__ keys, unsigned long long* red) {
  const long long Ne = (long long)nx * ny * nz;
  const int Lx = nx + 1, Ly = ny + 1;
  const double eps = fmax(1e-9, 1e-9 * yield);
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < Ne;
       e += (long long)gridDim.x * blockDim.x) {
    const int iz = (int)(e % nz);
    const int iy = (int)((e / nz) % ny);
    const int ix = (int)(e / ((long long)nz * ny));
    const bool act = active ? active[e] != 0 : true;
    const double s = vm[e] / vsize;
    stress3d[e] = (float)s;
    double dsum = 0.0;
    for (int q = 0; q < 8; ++q) {
      const long long n = ((long long)(iz + (q >> 2)) * Ly + (iy + ((q >> 1) & 1))) * Lx +
                          (ix + (q & 1));
      const double a = u[3 * n], b = u[3 * n + 1], c = u[3 * n + 2];
      dsum += sqrt(a * a + b * b + c * c);
    }
    const double d3 = dsum / 8.0;
    disp3d[e] = (float)d3;
    float sf = 0.f;
    if (act) {
      atomic_max_nonneg(&red[0], s);
      atomic_max_nonneg(&red[2], d3);
      if (yield > 0.0) {
        const double ratio = yield / fmax(s, eps);
        atomic_min_nonneg(&red[3], ratio);
        sf = (float)fmin(fmax(ratio, 0.0), cap);
      }
    }
    safety3d[e] = sf;
    keys[e] = act ? s : -1.0;
  }
}

// max nodal |u| over ALL nodes (void-only ones included, core/solve_worker.py:378-380)
__global__ void k_max_disp(long long Nn, const double* __restrict__ u, unsigned long long* red) {
  double m = 0.0;
  for (long long n = blockIdx.x * (long long)blockDim.x + threadIdx.x; n < Nn;
       n += (long long)gridDim.x * blockDim.x) {
    const double a = u[3 * n], b = u[3 * n + 1], c = u[3 * n + 2];
    m = fmax(m, sqrt(a * a + b * b + c * c));
  }
  atomic_max_nonneg(&red[1], m);
}

}  // namespace vxpost

extern "C" int vx_post_fields_impl(int nx, int ny, int nz, void* stream, const double* u,
                                   const double* vm, const uint8_t* active, double vsize,
                                   double yield, float* stress3d, float* safety3d, float* disp3d,
                                   double* out5, char* err, int errlen) {
  using namespace vxpost;
  hipStream_t st = (hipStream_t)stream;
  const long long Ne = (long long)nx * ny * nz;
  const long long Nn = (long long)(nx + 1) * (ny + 1) * (nz + 1);
  double *keys = nullptr, *sorted = nullptr;
  unsigned long long* red = nullptr;
  void* tmp = nullptr;
  size_t tmp_bytes = 0;
  auto bail = [&](const char* what, hipError_t e) {
    if (err) snprintf(err, errlen, "%s: %s", what, hipGetErrorString(e));
    (void)hipFree(keys);
    (void)hipFree(sorted);
    (void)hipFree(red);
    (void)hipFree(tmp);
    return VX_ERR_HIP;
  };
  hipError_t e;
  if ((e = hipMalloc(&keys, Ne * sizeof(double))) != hipSuccess) return bail("alloc", e);
  if ((e = hipMalloc(&sorted, Ne * sizeof(double))) != hipSuccess) return bail("alloc", e);
  if ((e = hipMalloc(&red, 4 * sizeof(unsigned long long))) != hipSuccess) return bail("alloc", e);
  // red: [0] max stress, [1] max |u|, [2] max disp3d (active), [3] min safety
  unsigned long long init[4] = {0ull, 0ull, 0ull, 0x7FF0000000000000ull};
  if ((e = hipMemcpyAsync(red, init, sizeof init, hipMemcpyHostToDevice, st)) != hipSuccess)
    return bail("init", e);
  const int grid = (int)std::min<long long>((Ne + 255) / 256, 4096);
  hipLaunchKernelGGL(k_post_elem, dim3(grid), dim3(256), 0, st, nx, ny, nz, u, vm, active, vsize,
                     yield, 10.0, stress3d, safety3d, disp3d, keys, red);
  hipLaunchKernelGGL(k_max_disp, dim3((int)std::min<long long>((Nn + 255) / 256, 4096)), dim3(256),
                     0, st, Nn, u, red);
  if ((e = hipGetLastError()) != hipSuccess) return bail("launch", e);
  if ((e = hipcub::DeviceRadixSort::SortKeys(nullptr, tmp_bytes, keys, sorted, (int)Ne, 0, 64,
                                             st)) != hipSuccess)
    return bail("sort size", e);
  if ((e = hipMalloc(&tmp, tmp_bytes)) != hipSuccess) return bail("alloc", e);
  if ((e = hipcub::DeviceRadixSort::SortKeys(tmp, tmp_bytes, keys, sorted, (int)Ne, 0, 64, st)) !=
      hipSuccess)
    return bail("sort", e);
  unsigned long long hred[4];
  if ((e = hipMemcpyAsync(hred, red, sizeof hred, hipMemcpyDeviceToHost, st)) != hipSuccess)
    return bail("readback", e);
  if ((e = hipStreamSynchronize(st)) != hipSuccess) return bail("sync", e);
  // number of active elements = keys >= 0: find the first non-negative by bisection
  long long lo = 0, hi = Ne;   // first index with sorted >= 0
  while (lo < hi) {
    const long long mid = (lo + hi) / 2;
    double v;
    if ((e = hipMemcpy(&v, sorted + mid, sizeof v, hipMemcpyDeviceToHost)) != hipSuccess)
      return bail("bisect", e);
    if (v >= 0.0) hi = mid;
    else lo = mid + 1;
  }
  const long long na = Ne - lo;
  double pct = 0.0;
  if (na > 0) {
    // numpy.percentile(..., 99) with linear interpolation between ranks
    const double pos = 0.99 * (double)(na - 1);
    const long long i0 = (long long)std::floor(pos);
    const long long i1 = std::min(i0 + 1, na - 1);
    double v0, v1;
    if ((e = hipMemcpy(&v0, sorted + lo + i0, sizeof v0, hipMemcpyDeviceToHost)) != hipSuccess ||
        (e = hipMemcpy(&v1, sorted + lo + i1, sizeof v1, hipMemcpyDeviceToHost)) != hipSuccess)
      return bail("percentile", e);
    const double g = pos - (double)i0;
    pct = v0 + (v1 - v0) * g;
  }
  auto asd = [](unsigned long long b) {
    double d;
    memcpy(&d, &b, sizeof d);
    return d;
  };
  out5[0] = na > 0 ? asd(hred[0]) : 0.0;                 // max_stress_pa
  out5[1] = std::max(pct, 1e-6);                          // stress_vmax
  out5[2] = asd(hred[1]);                                 // max_displacement_m
  out5[3] = std::max(na > 0 ? asd(hred[2]) : 0.0, 1e-9);  // displacement_vmax_m
  out5[4] = (na > 0 && yield > 0.0) ? asd(hred[3]) : NAN; // min_safety_factor
  (void)hipFree(keys);
  (void)hipFree(sorted);
  (void)hipFree(red);
  (void)hipFree(tmp);
  return VX_OK;
}
