// net.hip -- the learned coarse-grid corrector: a 4-level 3D U-Net on a node
// grid, BF16 activations and weights, FP32 accumulation on the matrix cores
// (v_mfma_f32_16x16x16_bf16), with its own training kernels (data and weight
// gradients, Adam).  Layer table, prolongation and rounding points are those
// of the CPU restatement in oracle/convnet_ref.py.
//
// Tensors are [B][voxel][C] with voxel = x + X*(y + Y*z) and C a multiple of
// 16; a 3x3x3 convolution is an implicit GEMM with M = voxels, N = output
// channels, K = 27 taps x input channels.  One wave computes 64 voxels x all
// output channels; each lane loads 4 consecutive input channels of one voxel
// (8 bytes, the A-operand slice of the MFMA) and the weights come pre-packed
// in the B-operand lane order, so every operand load is one aligned 8-byte
// access.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "net_internal.h"
#include "vxfea.h"

namespace vxnet {

typedef short s4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);   // round to nearest even
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

enum { EPI_RELU_BF = 0, EPI_F32 = 1, EPI_ACC_F32 = 2 };

// Per-axis source index of output coordinate o for tap offset dd in {0,1,2}.
// Forward convolution: s = stride*o + dd - 1.  Transposed (data gradient of a
// forward convolution whose output grid is the source here): the forward
// outputs q with stride*q + dd - 1 = o, i.e. q = (o - dd + 1) / stride.
template <bool TR>
__device__ __forceinline__ int src_axis(int o, int dd, int stride, int n, bool& ok) {
  if (!TR) {
    const int s = stride * o + dd - 1;
    ok = ok && s >= 0 && s < n;
    return s;
  }
  const int t = o - dd + 1;
  const int q = stride == 1 ? t : (t >> 1);
  ok = ok && t >= 0 && (stride == 1 || (t & 1) == 0) && q < n;
  return q;
}

// 3x3x3 convolution (TR = false) or its data gradient (TR = true) on the
// matrix cores.  src [B][Vs][KB*16] bf16, Wp packed [27][KB][NB][64] of s4,
// output [B][Vo][NB*16].  Grid: ceil(B*Vo / 256) work-groups of 4 waves.
template <int KB, int NB, bool TR, int EPI>
__global__ __launch_bounds__(256) void k_conv(const uint16_t* __restrict__ src, int Xs, int Ys,
                                              int Zs, int Xo, int Yo, int Zo, int stride, int B,
                                              const s4* __restrict__ Wp,
                                              const float* __restrict__ bias,
                                              uint16_t* __restrict__ out_bf,
                                              float* __restrict__ out_f) {
  constexpr int CIN = KB * 16, COUT = NB * 16;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const long long Vo = (long long)Xo * Yo * Zo;
  const long long Vs = (long long)Xs * Ys * Zs;
  const long long total = Vo * B;
  const long long base = ((long long)blockIdx.x * 4 + wave) * 64;
  if (base >= total) return;   // no barriers below: a finished wave may leave
  int ox[4], oy[4], oz[4];
  long long sb[4];
  bool act[4];
#pragma unroll
  for (int mb = 0; mb < 4; ++mb) {
    const long long v = base + mb * 16 + (lane & 15);
    act[mb] = v < total;
    const long long vv = act[mb] ? v : 0;
    const long long b = vv / Vo;
    const long long r = vv - b * Vo;
    ox[mb] = (int)(r % Xo);
    oy[mb] = (int)((r / Xo) % Yo);
    oz[mb] = (int)(r / ((long long)Xo * Yo));
    sb[mb] = b * Vs;
  }
  f4 acc[4][NB];
#pragma unroll
  for (int mb = 0; mb < 4; ++mb)
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) acc[mb][nb] = f4{0.f, 0.f, 0.f, 0.f};
  const int kofs = 4 * (lane >> 4);
  const s4 zero = s4{0, 0, 0, 0};
  for (int dz = 0; dz < 3; ++dz) {
    for (int dy = 0; dy < 3; ++dy) {
      for (int dx = 0; dx < 3; ++dx) {
        const int d = dx + 3 * (dy + 3 * dz);
        long long sidx[4];
        bool ok[4];
#pragma unroll
        for (int mb = 0; mb < 4; ++mb) {
          ok[mb] = act[mb];
          const int sx = src_axis<TR>(ox[mb], dx, stride, Xs, ok[mb]);
          const int sy = src_axis<TR>(oy[mb], dy, stride, Ys, ok[mb]);
          const int sz = src_axis<TR>(oz[mb], dz, stride, Zs, ok[mb]);
          sidx[mb] = ok[mb] ? (sb[mb] + sx + (long long)Xs * (sy + (long long)Ys * sz)) : 0;
        }
#pragma unroll
        for (int kb = 0; kb < KB; ++kb) {
          s4 a[4];
#pragma unroll
          for (int mb = 0; mb < 4; ++mb)
            a[mb] = ok[mb] ? *reinterpret_cast<const s4*>(src + sidx[mb] * CIN + kb * 16 + kofs)
                           : zero;
#pragma unroll
          for (int nb = 0; nb < NB; ++nb) {
            const s4 w = Wp[((d * KB + kb) * NB + nb) * 64 + lane];
#pragma unroll
            for (int mb = 0; mb < 4; ++mb)
              acc[mb][nb] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[mb], w, acc[mb][nb], 0, 0, 0);
          }
        }
      }
    }
  }
  // D[m][n]: lane holds rows m = 4*(lane/16) + i (voxels), column n = lane%16
#pragma unroll
  for (int mb = 0; mb < 4; ++mb) {
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) {
      const int c = nb * 16 + (lane & 15);
      const float bc = bias ? bias[c] : 0.f;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const long long v = base + mb * 16 + 4 * (lane >> 4) + i;
        if (v >= total) continue;
        const float val = acc[mb][nb][i] + bc;
        if (EPI == EPI_RELU_BF) out_bf[v * COUT + c] = f2bf(fmaxf(val, 0.f));
        else if (EPI == EPI_F32) out_f[v * COUT + c] = val;
        else out_f[v * COUT + c] += val;
      }
    }
  }
}

// Weight gradient dW[d][ci][co] = sum_o X[src(o, d)][ci] * G[o][co] on the
// matrix cores (M = input channels, N = output channels, K = voxels).  Grid
// (n_chunks, 27): wave w of the grid sums k-blocks [w*kpw, (w+1)*kpw) of 16
// output voxels and writes its own partial [w][d][CIN][COUT]; k_reduce_wgrad
// adds the partials in a fixed order (deterministic).
template <int KB, int NB>
__global__ __launch_bounds__(256) void k_conv_wgrad(const uint16_t* __restrict__ X, int Xs, int Ys,
                                                    int Zs, const uint16_t* __restrict__ G, int Xo,
                                                    int Yo, int Zo, int stride, int B, int kpw,
                                                    float* __restrict__ partial) {
  constexpr int CIN = KB * 16, COUT = NB * 16;
  const int lane = threadIdx.x & 63;
  const int w = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int d = blockIdx.y;
  const int dx = d % 3, dy = (d / 3) % 3, dz = d / 9;
  const long long Vo = (long long)Xo * Yo * Zo;
  const long long Vs = (long long)Xs * Ys * Zs;
  const long long total = Vo * B;
  const long long nkb = (total + 15) / 16;
  f4 acc[KB][NB];
#pragma unroll
  for (int mb = 0; mb < KB; ++mb)
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) acc[mb][nb] = f4{0.f, 0.f, 0.f, 0.f};
  const long long k0 = (long long)w * kpw;
  const long long k1 = std::min<long long>(k0 + kpw, nkb);
  for (long long kk = k0; kk < k1; ++kk) {
    long long xi[4], gi[4];
    bool ok[4], okg[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const long long v = kk * 16 + 4 * (lane >> 4) + j;
      okg[j] = v < total;
      const long long vv = okg[j] ? v : 0;
      const long long b = vv / Vo;
      const long long r = vv - b * Vo;
      const int ox = (int)(r % Xo), oy = (int)((r / Xo) % Yo), oz = (int)(r / ((long long)Xo * Yo));
      ok[j] = okg[j];
      const int sx = src_axis<false>(ox, dx, stride, Xs, ok[j]);
      const int sy = src_axis<false>(oy, dy, stride, Ys, ok[j]);
      const int sz = src_axis<false>(oz, dz, stride, Zs, ok[j]);
      xi[j] = ok[j] ? (b * Vs + sx + (long long)Xs * (sy + (long long)Ys * sz)) * CIN : 0;
      gi[j] = vv * COUT;
    }
    s4 a[KB], bb[NB];
#pragma unroll
    for (int mb = 0; mb < KB; ++mb) {
      const int ci = mb * 16 + (lane & 15);
#pragma unroll
      for (int j = 0; j < 4; ++j) a[mb][j] = ok[j] ? (short)X[xi[j] + ci] : (short)0;
    }
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) {
      const int co = nb * 16 + (lane & 15);
#pragma unroll
      for (int j = 0; j < 4; ++j) bb[nb][j] = okg[j] ? (short)G[gi[j] + co] : (short)0;
    }
#pragma unroll
    for (int mb = 0; mb < KB; ++mb)
#pragma unroll
      for (int nb = 0; nb < NB; ++nb)
        acc[mb][nb] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[mb], bb[nb], acc[mb][nb], 0, 0, 0);
  }
  float* P = partial + ((long long)w * 27 + d) * CIN * COUT;
#pragma unroll
  for (int mb = 0; mb < KB; ++mb)
#pragma unroll
    for (int nb = 0; nb < NB; ++nb)
#pragma unroll
      for (int i = 0; i < 4; ++i)
        P[(mb * 16 + 4 * (lane >> 4) + i) * COUT + nb * 16 + (lane & 15)] = acc[mb][nb][i];
}

// dW[i] = sum_w partial[w][i] for i < n (n = 27*CIN*COUT)
__global__ void k_reduce_wgrad(const float* __restrict__ partial, int nw, long long n,
                               float* __restrict__ dW) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    float s = 0.f;
    for (int w = 0; w < nw; ++w) s += partial[(long long)w * n + i];
    dW[i] = s;
  }
}

// db[c] = sum_v G[v][c] in two fixed-order stages.  Stage 1: work-group b
// sums rows [b R, (b+1) R) of G[v][C]; thread t takes channel t % C and rows
// t / C, t / C + 256 / C, ... (consecutive threads read consecutive values),
// then the 256 / C row lanes of each channel are added in order.  C divides
// 256 (16, 32 or 64 channels).
__global__ __launch_bounds__(256) void k_bias_part(const uint16_t* __restrict__ G, long long nv,
                                                   int C, long long R, float* __restrict__ part) {
  __shared__ float red[256];
  const int t = threadIdx.x, c = t % C, rl = t / C, nrl = 256 / C;
  const long long r0 = blockIdx.x * R, r1 = min(nv, r0 + R);
  float s = 0.f;
  for (long long v = r0 + rl; v < r1; v += nrl) s += bf2f(G[v * C + c]);
  red[t] = s;
  __syncthreads();
  if (t < C) {
    float a = 0.f;
    for (int k = 0; k < nrl; ++k) a += red[k * C + t];
    part[(long long)blockIdx.x * C + t] = a;
  }
}

// Stage 2: db[c] = sum over the stage-1 blocks b of part[b][c], in order.
__global__ void k_bias_sum(const float* __restrict__ part, int nb, int C, float* __restrict__ db) {
  const int c = threadIdx.x;
  if (c >= C) return;
  float a = 0.f;
  for (int b = 0; b < nb; ++b) a += part[(long long)b * C + c];
  db[c] = a;
}

// g = bf16(relu ? (h > 0 ? d : 0) : d)
__global__ void k_relu_bwd(const float* __restrict__ d, const uint16_t* __restrict__ h, long long n,
                           uint16_t* __restrict__ g) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const float v = (h == nullptr || bf2f(h[i]) > 0.f) ? d[i] : 0.f;
    g[i] = f2bf(v);
  }
}

__global__ void k_f32_to_bf16(const float* __restrict__ a, long long n, uint16_t* __restrict__ b) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    b[i] = f2bf(a[i]);
}

// 1-D weights of the channel-wise prolongation: fine i <- coarse i/2 (even i)
// or (i-1)/2 and (i+1)/2 with 0.5 each (odd i; indices >= nc dropped)
__device__ __forceinline__ int up_taps(int i, int nc, int* j, float* w) {
  if ((i & 1) == 0) {
    j[0] = i >> 1;
    w[0] = 1.f;
    return 1;
  }
  j[0] = (i - 1) >> 1;
  w[0] = 0.5f;
  if (((i + 1) >> 1) < nc) {
    j[1] = (i + 1) >> 1;
    w[1] = 0.5f;
    return 2;
  }
  return 1;
}

// out[f][c] = bf16(up(h)[f][c] + skip[f][c]); h on the coarse grid
__global__ void k_up_add(const uint16_t* __restrict__ h, int Xc, int Yc, int Zc,
                         const uint16_t* __restrict__ skip, int Xf, int Yf, int Zf, int C, int B,
                         uint16_t* __restrict__ out) {
  const long long Vf = (long long)Xf * Yf * Zf, Vc = (long long)Xc * Yc * Zc;
  const long long n = Vf * B * C;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const int c = (int)(i % C);
    const long long v = i / C;
    const long long b = v / Vf;
    const long long r = v - b * Vf;
    const int x = (int)(r % Xf), y = (int)((r / Xf) % Yf), z = (int)(r / ((long long)Xf * Yf));
    int jx[2], jy[2], jz[2];
    float wx[2], wy[2], wz[2];
    const int nx = up_taps(x, Xc, jx, wx), ny = up_taps(y, Yc, jy, wy), nz = up_taps(z, Zc, jz, wz);
    float s = 0.f;
    for (int a = 0; a < nz; ++a)
      for (int e = 0; e < ny; ++e)
        for (int q = 0; q < nx; ++q)
          s += wz[a] * wy[e] * wx[q] *
               bf2f(h[(b * Vc + jx[q] + (long long)Xc * (jy[e] + (long long)Yc * jz[a])) * C + c]);
    out[i] = f2bf(s + bf2f(skip[i]));
  }
}

// adjoint of the prolongation: dc[j][c] = sum over fine i of w(i -> j) df[i][c]
__global__ void k_up_adj(const float* __restrict__ df, int Xf, int Yf, int Zf, int Xc, int Yc,
                         int Zc, int C, int B, float* __restrict__ dc) {
  const long long Vf = (long long)Xf * Yf * Zf, Vc = (long long)Xc * Yc * Zc;
  const long long n = Vc * B * C;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const int c = (int)(i % C);
    const long long v = i / C;
    const long long b = v / Vc;
    const long long r = v - b * Vc;
    const int x = (int)(r % Xc), y = (int)((r / Xc) % Yc), z = (int)(r / ((long long)Xc * Yc));
    float s = 0.f;
    for (int oz = -1; oz <= 1; ++oz) {
      const int fz = 2 * z + oz;
      if (fz < 0 || fz >= Zf) continue;
      const float wz = oz == 0 ? 1.f : 0.5f;
      for (int oy = -1; oy <= 1; ++oy) {
        const int fy = 2 * y + oy;
        if (fy < 0 || fy >= Yf) continue;
        const float wy = oy == 0 ? 1.f : 0.5f;
        for (int ox = -1; ox <= 1; ++ox) {
          const int fx = 2 * x + ox;
          if (fx < 0 || fx >= Xf) continue;
          const float wx = ox == 0 ? 1.f : 0.5f;
          s += wz * wy * wx * df[(b * Vf + fx + (long long)Xf * (fy + (long long)Yf * fz)) * C + c];
        }
      }
    }
    dc[i] = s;
  }
}

// pack fp32 W[27][CIN][COUT] into the MFMA B-operand lane order, rounded to bf16.
// forward:  B[k = ci][n = co]; transposed: B[k = co][n = ci]
__global__ void k_pack(const float* __restrict__ W, int CIN, int COUT, int tr,
                       uint16_t* __restrict__ Wp) {
  const int KB = (tr ? COUT : CIN) / 16, NB = (tr ? CIN : COUT) / 16;
  const long long n = 27LL * KB * NB * 64 * 4;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const int j = (int)(i & 3);
    const int lane = (int)((i >> 2) & 63);
    long long t = i >> 8;
    const int nb = (int)(t % NB);
    t /= NB;
    const int kb = (int)(t % KB);
    const int d = (int)(t / KB);
    const int k = kb * 16 + 4 * (lane >> 4) + j;
    const int nn = nb * 16 + (lane & 15);
    const int ci = tr ? nn : k, co = tr ? k : nn;
    Wp[i] = f2bf(W[((long long)d * CIN + ci) * COUT + co]);
  }
}

__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, long long n, float lr, float b1, float b2, float eps,
                       float c1, float c2) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  }
}

int grid1(long long n) { return (int)std::max<long long>(1, std::min<long long>((n + 255) / 256, 4096)); }

}  // namespace vxnet

// ===========================================================================
// host side
// ===========================================================================
using namespace vxnet;

struct NetLayer {
  int lev, cin, cout, stride, relu;
  long long woff, boff;   // offsets of W and b in the flat parameter vector
  uint16_t* wp = nullptr;   // packed forward weights
  uint16_t* wt = nullptr;   // packed transposed weights
};

struct vx_net {
  int c[4];
  std::vector<NetLayer> L;
  long long nparams = 0;
  float* params = nullptr;
  float* am = nullptr;
  float* av = nullptr;
  bool packed = false;
  hipStream_t stream = nullptr;
  std::string err;
  // activation workspace for the current (B, dims)
  int B = 0, dims[4][3] = {};
  long long V[4] = {};
  uint16_t* in = nullptr;
  uint16_t* h[14] = {};
  uint16_t *a0 = nullptr, *a1 = nullptr, *a2 = nullptr;
  float* dtmp[4] = {};
  float* ds[3] = {};
  uint16_t* g[4] = {};
  float* partial = nullptr;
  long long partial_n = 0;
  std::vector<void*> act_allocs;
  bool have_act = false;
};

namespace {

int nfail(vx_net* n, int code, const std::string& m) {
  n->err = m;
  return code;
}
#define NHIP(x)                                                              \
  do {                                                                       \
    hipError_t e_ = (x);                                                     \
    if (e_ != hipSuccess) return nfail(n, VX_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)
#define NLAUNCH()                                                            \
  do {                                                                       \
    hipError_t e_ = hipGetLastError();                                       \
    if (e_ != hipSuccess) return nfail(n, VX_ERR_HIP, std::string("launch: ") + hipGetErrorString(e_)); \
  } while (0)
#define NCHK(x)                  \
  do {                           \
    int r_ = (x);                \
    if (r_ != VX_OK) return r_;  \
  } while (0)

template <typename T>
int nalloc(vx_net* n, T** p, long long count) {
  if (hipMalloc((void**)p, sizeof(T) * std::max<long long>(count, 1)) != hipSuccess)
    return nfail(n, VX_ERR_NOMEM, "device allocation failed");
  n->act_allocs.push_back(*p);
  return VX_OK;
}

void free_act(vx_net* n) {
  for (void* p : n->act_allocs) (void)hipFree(p);
  n->act_allocs.clear();
  n->have_act = false;
  n->B = 0;
  n->partial = nullptr;
  n->partial_n = 0;
}

int ensure_act(vx_net* n, int B, int X, int Y, int Z) {
  if (n->have_act && n->B == B && n->dims[0][0] == X && n->dims[0][1] == Y && n->dims[0][2] == Z)
    return VX_OK;
  free_act(n);
  n->B = B;
  n->dims[0][0] = X;
  n->dims[0][1] = Y;
  n->dims[0][2] = Z;
  for (int l = 1; l < 4; ++l)
    for (int a = 0; a < 3; ++a) n->dims[l][a] = (n->dims[l - 1][a] + 1) / 2;
  for (int l = 0; l < 4; ++l) n->V[l] = (long long)n->dims[l][0] * n->dims[l][1] * n->dims[l][2];
  auto sz = [&](int lev, int C) { return n->V[lev] * B * C; };
  NCHK(nalloc(n, &n->in, sz(0, 16)));
  for (int i = 0; i < 13; ++i) NCHK(nalloc(n, &n->h[i], sz(n->L[i].lev, n->L[i].cout)));
  NCHK(nalloc(n, &n->a2, sz(2, n->c[2])));
  NCHK(nalloc(n, &n->a1, sz(1, n->c[1])));
  NCHK(nalloc(n, &n->a0, sz(0, n->c[0])));
  // widest tensor per level: input 16 / c0 (0), c0 / c1 (1), c1 / c2 (2), c2 (3)
  const int wide[4] = {std::max(16, n->c[0]), std::max(n->c[0], n->c[1]),
                       std::max(n->c[1], n->c[2]), n->c[2]};
  for (int l = 0; l < 4; ++l) {
    NCHK(nalloc(n, &n->dtmp[l], sz(l, wide[l])));
    NCHK(nalloc(n, &n->g[l], sz(l, wide[l])));
  }
  for (int l = 0; l < 3; ++l) NCHK(nalloc(n, &n->ds[l], sz(l, n->c[l])));
  n->have_act = true;
  return VX_OK;
}

int pack_all(vx_net* n) {
  for (auto& l : n->L) {
    const long long np = 27LL * l.cin * l.cout;
    hipLaunchKernelGGL(k_pack, dim3(grid1(np)), dim3(256), 0, n->stream, n->params + l.woff, l.cin,
                       l.cout, 0, l.wp);
    hipLaunchKernelGGL(k_pack, dim3(grid1(np)), dim3(256), 0, n->stream, n->params + l.woff, l.cin,
                       l.cout, 1, l.wt);
  }
  NLAUNCH();
  n->packed = true;
  return VX_OK;
}

template <bool TR, int EPI>
int conv_launch(vx_net* n, int KB, int NB, const uint16_t* src, const int* sd, const int* od,
                int stride, const uint16_t* wp, const float* bias, uint16_t* obf, float* of) {
  const long long total = (long long)od[0] * od[1] * od[2] * n->B;
  const dim3 grid((unsigned)((total + 255) / 256));
  const s4* W = reinterpret_cast<const s4*>(wp);
#define VXN_CONV(kb, nb)                                                                        \
  hipLaunchKernelGGL((k_conv<kb, nb, TR, EPI>), grid, dim3(256), 0, n->stream, src, sd[0], sd[1], \
                     sd[2], od[0], od[1], od[2], stride, n->B, W, bias, obf, of)
  switch (KB * 8 + NB) {
    case 9: VXN_CONV(1, 1); break;
    case 10: VXN_CONV(1, 2); break;
    case 12: VXN_CONV(1, 4); break;
    case 17: VXN_CONV(2, 1); break;
    case 18: VXN_CONV(2, 2); break;
    case 20: VXN_CONV(2, 4); break;
    case 33: VXN_CONV(4, 1); break;
    case 34: VXN_CONV(4, 2); break;
    case 36: VXN_CONV(4, 4); break;
    default: return nfail(n, VX_ERR_ARG, "unsupported channel counts");
  }
#undef VXN_CONV
  NLAUNCH();
  return VX_OK;
}

// forward of layer i from src (grid level ls) into h[i] (or the fp32 output)
int layer_fwd(vx_net* n, int i, const uint16_t* src, float* out_f) {
  const NetLayer& l = n->L[i];
  const int ls = l.stride == 2 ? l.lev - 1 : l.lev;
  const float* bias = n->params + l.boff;
  if (l.relu)
    return conv_launch<false, EPI_RELU_BF>(n, l.cin / 16, l.cout / 16, src, n->dims[ls],
                                          n->dims[l.lev], l.stride, l.wp, bias, n->h[i], nullptr);
  return conv_launch<false, EPI_F32>(n, l.cin / 16, l.cout / 16, src, n->dims[ls], n->dims[l.lev],
                                     l.stride, l.wp, bias, nullptr, out_f);
}

int up_add(vx_net* n, const uint16_t* h, int lc, const uint16_t* skip, int C, uint16_t* out) {
  const int lf = lc - 1;
  const long long tot = n->V[lf] * n->B * C;
  hipLaunchKernelGGL(k_up_add, dim3(grid1(tot)), dim3(256), 0, n->stream, h, n->dims[lc][0],
                     n->dims[lc][1], n->dims[lc][2], skip, n->dims[lf][0], n->dims[lf][1],
                     n->dims[lf][2], C, n->B, out);
  NLAUNCH();
  return VX_OK;
}

int forward_from_input(vx_net* n, float* out_f) {
  if (!n->packed) NCHK(pack_all(n));
  NCHK(layer_fwd(n, 0, n->in, nullptr));
  for (int i = 1; i < 8; ++i) NCHK(layer_fwd(n, i, n->h[i - 1], nullptr));
  NCHK(up_add(n, n->h[7], 3, n->h[5], n->c[2], n->a2));
  NCHK(layer_fwd(n, 8, n->a2, nullptr));
  NCHK(layer_fwd(n, 9, n->h[8], nullptr));
  NCHK(up_add(n, n->h[9], 2, n->h[3], n->c[1], n->a1));
  NCHK(layer_fwd(n, 10, n->a1, nullptr));
  NCHK(layer_fwd(n, 11, n->h[10], nullptr));
  NCHK(up_add(n, n->h[11], 1, n->h[1], n->c[0], n->a0));
  NCHK(layer_fwd(n, 12, n->a0, nullptr));
  return layer_fwd(n, 13, n->h[12], out_f);
}

int forward(vx_net* n, const float* in_f, float* out_f) {
  const long long nin = n->V[0] * n->B * 16;
  hipLaunchKernelGGL(k_f32_to_bf16, dim3(grid1(nin)), dim3(256), 0, n->stream, in_f, nin, n->in);
  NLAUNCH();
  return forward_from_input(n, out_f);
}

// backward of layer i: d = gradient w.r.t. its output (fp32), x = its input;
// parameter gradients into grad; data gradient written (acc=false) or added
// (acc=true) into dx (nullptr: not needed)
int layer_bwd(vx_net* n, int i, const float* d, const uint16_t* x, float* grad, float* dx,
              bool acc) {
  const NetLayer& l = n->L[i];
  const int lo = l.lev, ls = l.stride == 2 ? l.lev - 1 : l.lev;
  const long long no = n->V[lo] * n->B * l.cout;
  uint16_t* g = n->g[lo];
  hipLaunchKernelGGL(k_relu_bwd, dim3(grid1(no)), dim3(256), 0, n->stream, d,
                     l.relu ? n->h[i] : nullptr, no, g);
  NLAUNCH();
  // weight gradient
  const int KB = l.cin / 16, NB = l.cout / 16;
  const long long nkb = (n->V[lo] * n->B + 15) / 16;
  const int kpw = (int)std::max<long long>(1, (nkb + 255) / 256);
  const int nw = (int)((nkb + kpw - 1) / kpw);
  const int nchunk = (nw + 3) / 4;
  const long long wsz = 27LL * l.cin * l.cout;
  const long long need = (long long)nchunk * 4 * wsz;
  if (need > n->partial_n) {
    NCHK(nalloc(n, &n->partial, need));
    n->partial_n = need;
  }
  const dim3 wg((unsigned)nchunk, 27);
#define VXN_WG(kb, nb)                                                                        \
  hipLaunchKernelGGL((k_conv_wgrad<kb, nb>), wg, dim3(256), 0, n->stream, x, n->dims[ls][0],    \
                     n->dims[ls][1], n->dims[ls][2], g, n->dims[lo][0], n->dims[lo][1],       \
                     n->dims[lo][2], l.stride, n->B, kpw, n->partial)
  switch (KB * 8 + NB) {
    case 9: VXN_WG(1, 1); break;
    case 10: VXN_WG(1, 2); break;
    case 12: VXN_WG(1, 4); break;
    case 17: VXN_WG(2, 1); break;
    case 18: VXN_WG(2, 2); break;
    case 20: VXN_WG(2, 4); break;
    case 33: VXN_WG(4, 1); break;
    case 34: VXN_WG(4, 2); break;
    case 36: VXN_WG(4, 4); break;
    default: return nfail(n, VX_ERR_ARG, "unsupported channel counts");
  }
#undef VXN_WG
  NLAUNCH();
  // waves past the last k-block wrote zero partials: reduce over all launched waves
  hipLaunchKernelGGL(k_reduce_wgrad, dim3(grid1(wsz)), dim3(256), 0, n->stream, n->partial,
                     nchunk * 4, wsz, grad + l.woff);
  // bias gradient; its stage-1 partials (at most 256 x cout floats) reuse
  // n->partial, which holds at least 4 x 27 x cin x cout floats and which
  // k_reduce_wgrad, ordered before on the stream, has consumed
  {
    const long long nv = n->V[lo] * n->B;
    const long long R = std::max<long long>(1024, (nv + 255) / 256);
    const int nbb = (int)((nv + R - 1) / R);
    hipLaunchKernelGGL(k_bias_part, dim3(nbb), dim3(256), 0, n->stream, g, nv, l.cout, R,
                       n->partial);
    hipLaunchKernelGGL(k_bias_sum, dim3(1), dim3(64), 0, n->stream, n->partial, nbb, l.cout,
                       grad + l.boff);
  }
  NLAUNCH();
  if (!dx) return VX_OK;
  if (acc)
    return conv_launch<true, EPI_ACC_F32>(n, l.cout / 16, l.cin / 16, g, n->dims[lo], n->dims[ls],
                                          l.stride, l.wt, nullptr, nullptr, dx);
  return conv_launch<true, EPI_F32>(n, l.cout / 16, l.cin / 16, g, n->dims[lo], n->dims[ls],
                                    l.stride, l.wt, nullptr, nullptr, dx);
}

int up_adj(vx_net* n, const float* df, int lf, int C, float* dc) {
  const int lc = lf + 1;
  const long long tot = n->V[lc] * n->B * C;
  hipLaunchKernelGGL(k_up_adj, dim3(grid1(tot)), dim3(256), 0, n->stream, df, n->dims[lf][0],
                     n->dims[lf][1], n->dims[lf][2], n->dims[lc][0], n->dims[lc][1],
                     n->dims[lc][2], C, n->B, dc);
  NLAUNCH();
  return VX_OK;
}

int backward(vx_net* n, const float* dout, float* grad) {
  float** T = n->dtmp;
  float** S = n->ds;
  NCHK(layer_bwd(n, 13, dout, n->h[12], grad, T[0], false));
  NCHK(layer_bwd(n, 12, T[0], n->a0, grad, S[0], false));
  NCHK(up_adj(n, S[0], 0, n->c[0], T[1]));
  NCHK(layer_bwd(n, 11, T[1], n->h[10], grad, T[1], false));
  NCHK(layer_bwd(n, 10, T[1], n->a1, grad, S[1], false));
  NCHK(up_adj(n, S[1], 1, n->c[1], T[2]));
  NCHK(layer_bwd(n, 9, T[2], n->h[8], grad, T[2], false));
  NCHK(layer_bwd(n, 8, T[2], n->a2, grad, S[2], false));
  NCHK(up_adj(n, S[2], 2, n->c[2], T[3]));
  NCHK(layer_bwd(n, 7, T[3], n->h[6], grad, T[3], false));
  NCHK(layer_bwd(n, 6, T[3], n->h[5], grad, S[2], true));
  NCHK(layer_bwd(n, 5, S[2], n->h[4], grad, T[2], false));
  NCHK(layer_bwd(n, 4, T[2], n->h[3], grad, S[1], true));
  NCHK(layer_bwd(n, 3, S[1], n->h[2], grad, T[1], false));
  NCHK(layer_bwd(n, 2, T[1], n->h[1], grad, S[0], true));
  NCHK(layer_bwd(n, 1, S[0], n->h[0], grad, T[0], false));
  return layer_bwd(n, 0, T[0], n->in, grad, nullptr, false);
}

}  // namespace

namespace vxnet_api {
int prepare(vx_net* n, hipStream_t st, int B, int X, int Y, int Z) {
  n->stream = st;
  return ensure_act(n, B, X, Y, Z);
}
uint16_t* input(vx_net* n) { return n->in; }
int forward_bf16(vx_net* n, float* out) { return forward_from_input(n, out); }
int backward(vx_net* n, const float* dout, float* grad) { return ::backward(n, dout, grad); }
int adam(vx_net* n, const float* grad, float lr, float b1, float b2, float eps, int step) {
  return vxn_adam_step(n, grad, lr, b1, b2, eps, step);
}
long long num_params(const vx_net* n) { return n->nparams; }
const char* error(const vx_net* n) { return n->err.c_str(); }
}  // namespace vxnet_api

extern "C" {

int vxn_create(int32_t c0, int32_t c1, int32_t c2, vxn_handle* out) {
  if (!out) return VX_ERR_ARG;
  *out = nullptr;
  for (int c : {c0, c1, c2})
    if (c < 16 || c > 64 || c % 16 || (c & (c - 1))) return VX_ERR_ARG;
  vx_net* n = new vx_net();
  n->c[0] = c0;
  n->c[1] = c1;
  n->c[2] = c2;
  n->c[3] = c2;
  // (level of the output grid, cin, cout, stride, relu); oracle/convnet_ref.py layer_table
  const int T[14][5] = {{0, 16, c0, 1, 1}, {0, c0, c0, 1, 1}, {1, c0, c1, 2, 1}, {1, c1, c1, 1, 1},
                        {2, c1, c2, 2, 1}, {2, c2, c2, 1, 1}, {3, c2, c2, 2, 1}, {3, c2, c2, 1, 1},
                        {2, c2, c2, 1, 1}, {2, c2, c1, 1, 1}, {1, c1, c1, 1, 1}, {1, c1, c0, 1, 1},
                        {0, c0, c0, 1, 1}, {0, c0, 16, 1, 0}};
  long long off = 0;
  for (auto& t : T) {
    NetLayer l;
    l.lev = t[0];
    l.cin = t[1];
    l.cout = t[2];
    l.stride = t[3];
    l.relu = t[4];
    l.woff = off;
    off += 27LL * l.cin * l.cout;
    l.boff = off;
    off += l.cout;
    n->L.push_back(l);
  }
  n->nparams = off;
  bool ok = hipMalloc(&n->params, sizeof(float) * off) == hipSuccess &&
            hipMalloc(&n->am, sizeof(float) * off) == hipSuccess &&
            hipMalloc(&n->av, sizeof(float) * off) == hipSuccess;
  for (auto& l : n->L) {
    const size_t nb = sizeof(uint16_t) * 27 * l.cin * l.cout;
    ok = ok && hipMalloc(&l.wp, nb) == hipSuccess && hipMalloc(&l.wt, nb) == hipSuccess;
  }
  ok = ok && hipMemset(n->params, 0, sizeof(float) * off) == hipSuccess &&
       hipMemset(n->am, 0, sizeof(float) * off) == hipSuccess &&
       hipMemset(n->av, 0, sizeof(float) * off) == hipSuccess;
  if (!ok) {
    vxn_destroy(n);
    return VX_ERR_NOMEM;
  }
  *out = n;
  return VX_OK;
}

void vxn_destroy(vxn_handle n) {
  if (!n) return;
  free_act(n);
  for (auto& l : n->L) {
    (void)hipFree(l.wp);
    (void)hipFree(l.wt);
  }
  (void)hipFree(n->params);
  (void)hipFree(n->am);
  (void)hipFree(n->av);
  delete n;
}

const char* vxn_last_error(vxn_handle n) { return n ? n->err.c_str() : "null handle"; }

int vxn_set_stream(vxn_handle n, void* stream) {
  if (!n) return VX_ERR_ARG;
  n->stream = (hipStream_t)stream;
  return VX_OK;
}

int64_t vxn_num_params(vxn_handle n) { return n ? n->nparams : VX_ERR_ARG; }

int vxn_set_params(vxn_handle n, const float* params) {
  if (!n || !params) return VX_ERR_ARG;
  NHIP(hipMemcpyAsync(n->params, params, sizeof(float) * n->nparams, hipMemcpyDeviceToDevice,
                      n->stream));
  NHIP(hipMemsetAsync(n->am, 0, sizeof(float) * n->nparams, n->stream));
  NHIP(hipMemsetAsync(n->av, 0, sizeof(float) * n->nparams, n->stream));
  n->packed = false;
  return VX_OK;
}

int vxn_get_params(vxn_handle n, float* params) {
  if (!n || !params) return VX_ERR_ARG;
  NHIP(hipMemcpyAsync(params, n->params, sizeof(float) * n->nparams, hipMemcpyDeviceToDevice,
                      n->stream));
  return VX_OK;
}

int vxn_forward(vxn_handle n, int32_t B, int32_t X, int32_t Y, int32_t Z, const float* in,
                float* out) {
  if (!n || !in || !out || B < 1 || X < 2 || Y < 2 || Z < 2) return VX_ERR_ARG;
  NCHK(ensure_act(n, B, X, Y, Z));
  return forward(n, in, out);
}

int vxn_backward(vxn_handle n, const float* dout, float* grad) {
  if (!n || !dout || !grad) return VX_ERR_ARG;
  if (!n->have_act) return nfail(n, VX_ERR_ARG, "vxn_backward needs a vxn_forward first");
  return backward(n, dout, grad);
}

int vxn_adam_step(vxn_handle n, const float* grad, float lr, float beta1, float beta2, float eps,
                  int32_t step) {
  if (!n || !grad || step < 1) return VX_ERR_ARG;
  const float c1 = 1.f - powf(beta1, (float)step), c2 = 1.f - powf(beta2, (float)step);
  hipLaunchKernelGGL(k_adam, dim3(grid1(n->nparams)), dim3(256), 0, n->stream, n->params, grad,
                     n->am, n->av, n->nparams, lr, beta1, beta2, eps, c1, c2);
  NLAUNCH();
  n->packed = false;
  return VX_OK;
}

}  // extern "C"
