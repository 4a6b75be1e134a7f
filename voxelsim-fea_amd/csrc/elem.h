// Hex8 unit-cube element operator in the Hadamard (bilinear-mode) basis.
//
// The reference builds the dense 24x24 KE by 2x2x2 Gauss quadrature
// (core/fea.py:31-76) and applies it as a GEMM per element
// (core/multigrid.py:334-337).  Here the same operator is applied exactly,
// in ~150 flops instead of 1152:
//
//  * Corner values of each displacement component are Hadamard-transformed
//    to the trilinear monomial coefficients h_m, m in {x,y,z,xy,yz,zx,xyz}
//    (h_m = sum_a chi_m(a) u_a, chi = product of the corner's +-1 signs).
//  * Under 2x2x2 Gauss quadrature the gradient terms of different monomial
//    type are orthogonal, so the element energy splits into a constant-strain
//    block (x,y,z modes), a bilinear block (xy,yz,zx modes) and a diagonal
//    trilinear block (xyz mode).  The mode-space "stresses" below are the
//    closed forms of those blocks with lambda = nu/((1+nu)(1-2nu)),
//    mu = 1/(2(1+nu)) (unit E):
//        linear   : k1 = lambda/16, k2 = mu/16
//        bilinear : b1 = lambda/48, b2 = mu/48
//        trilinear: t1 = (lambda + 4 mu)/144
//  * The inverse Hadamard transform maps mode forces back to corner forces.
//
// The result equals KE*u to rounding (the tests compare it with the
// oracle's dense KE); the diagonal of KE is the same for all 24 DOFs,
// kd = 3*(k1+2*k2) + ... and is computed here by applying the operator to a
// unit vector.
//
// Corner indexing inside this file is binary, q = sx + 2*sy + 4*sz.
#pragma once

#include <cmath>

#if defined(__HIPCC__)
#define VX_HD __host__ __device__
#else
#define VX_HD
#endif

namespace vx {

template <typename T>
struct ElemConst {
  T k1, k2, b1, b2, t1;
};

inline ElemConst<double> elem_const(double nu) {
  const double lam = nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
  const double mu = 1.0 / (2.0 * (1.0 + nu));
  ElemConst<double> c;
  c.k1 = lam / 16.0;
  c.k2 = mu / 16.0;
  c.b1 = lam / 48.0;
  c.b2 = mu / 48.0;
  c.t1 = (lam + 4.0 * mu) / 144.0;
  return c;
}

// 2D (x,y) Hadamard modes of the four corners of one node plane:
// m[0] = sum, m[1] = x-difference, m[2] = y-difference, m[3] = xy.
// u00,u10,u01,u11 are the corner values at (sx,sy).
template <typename T>
VX_HD inline void had2(T u00, T u10, T u01, T u11, T m[4]) {
  const T s0 = u10 + u00, d0 = u10 - u00;   // row sy=0
  const T s1 = u11 + u01, d1 = u11 - u01;   // row sy=1
  m[0] = s1 + s0;
  m[1] = d1 + d0;
  m[2] = s1 - s0;
  m[3] = d1 - d0;
}

// Inverse of had2 (transpose, no 1/4): corner values from 2D modes.
template <typename T>
VX_HD inline void ihad2(const T m[4], T& y00, T& y10, T& y01, T& y11) {
  const T a = m[0] - m[2], b = m[1] - m[3];   // sy = 0 row
  const T c = m[0] + m[2], d = m[1] + m[3];   // sy = 1 row
  y00 = a - b;
  y10 = a + b;
  y01 = c - d;
  y11 = c + d;
}

// Mode-space element response.  Inputs: the 2D modes of the bottom (Pb) and
// top (Pt) node planes of the element for the three components, and the
// element's stiffness scale E.  Outputs: the 2D modes of the corner forces
// on the bottom (Yb) and top (Yt) planes.  Yb/Yt are OVERWRITTEN.
template <typename T>
VX_HD inline void elem_modes(const T Pb[3][4], const T Pt[3][4], T E, const ElemConst<T>& K,
                             T Yb[3][4], T Yt[3][4]) {
  // z stage of the forward transform: 3D modes per component.
  T hx[3], hy[3], hz[3], hxy[3], hyz[3], hzx[3], hxyz[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    hz[c] = Pt[c][0] - Pb[c][0];
    hx[c] = Pt[c][1] + Pb[c][1];
    hzx[c] = Pt[c][1] - Pb[c][1];
    hy[c] = Pt[c][2] + Pb[c][2];
    hyz[c] = Pt[c][2] - Pb[c][2];
    hxy[c] = Pt[c][3] + Pb[c][3];
    hxyz[c] = Pt[c][3] - Pb[c][3];
  }
  const T k1 = K.k1 * E, k2 = K.k2 * E, b1 = K.b1 * E, b2 = K.b2 * E, t1 = K.t1 * E;
  const T k22 = k2 + k2, b23 = b2 * T(3);
  // constant-strain block
  const T tr = hx[0] + hy[1] + hz[2];
  const T ltr = k1 * tr;
  T yx[3], yy[3], yz[3];
  yx[0] = ltr + k22 * hx[0];
  yy[1] = ltr + k22 * hy[1];
  yz[2] = ltr + k22 * hz[2];
  yy[0] = yx[1] = k2 * (hy[0] + hx[1]);
  yz[0] = yx[2] = k2 * (hz[0] + hx[2]);
  yz[1] = yy[2] = k2 * (hz[1] + hy[2]);
  // bilinear block
  const T trxi = b1 * (hxy[1] + hzx[2]);
  const T treta = b1 * (hxy[0] + hyz[2]);
  const T trze = b1 * (hzx[0] + hyz[1]);
  const T s = hxy[2] + hzx[1] + hyz[0];
  T yxy[3], yyz[3], yzx[3];
  yxy[0] = treta + b23 * hxy[0];
  yxy[1] = trxi + b23 * hxy[1];
  yxy[2] = b2 * (hxy[2] + s);
  yyz[0] = b2 * (hyz[0] + s);
  yyz[1] = trze + b23 * hyz[1];
  yyz[2] = treta + b23 * hyz[2];
  yzx[0] = trze + b23 * hzx[0];
  yzx[1] = b2 * (hzx[1] + s);
  yzx[2] = trxi + b23 * hzx[2];
  // inverse z stage (mode 0 carries no force)
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const T yxyz = t1 * hxyz[c];
    Yt[c][0] = yz[c];
    Yb[c][0] = -yz[c];
    Yt[c][1] = yx[c] + yzx[c];
    Yb[c][1] = yx[c] - yzx[c];
    Yt[c][2] = yy[c] + yyz[c];
    Yb[c][2] = yy[c] - yyz[c];
    Yt[c][3] = yxy[c] + yxyz;
    Yb[c][3] = yxy[c] - yxyz;
  }
}

// Full element apply on binary-ordered corners: y[q][c] = E * (KE u)[q][c].
template <typename T>
VX_HD inline void elem_apply(const T u[8][3], T E, const ElemConst<T>& K, T y[8][3]) {
  T Pb[3][4], Pt[3][4], Yb[3][4], Yt[3][4];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    had2(u[0][c], u[1][c], u[2][c], u[3][c], Pb[c]);
    had2(u[4][c], u[5][c], u[6][c], u[7][c], Pt[c]);
  }
  elem_modes(Pb, Pt, E, K, Yb, Yt);
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    ihad2(Yb[c], y[0][c], y[1][c], y[2][c], y[3][c]);
    ihad2(Yt[c], y[4][c], y[5][c], y[6][c], y[7][c]);
  }
}

// Diagonal entry of the unit-E KE (identical for all 24 DOFs of the cube).
inline double elem_diag(double nu) {
  ElemConst<double> K = elem_const(nu);
  double u[8][3] = {}, y[8][3];
  u[0][0] = 1.0;
  elem_apply<double>(u, 1.0, K, y);
  return y[0][0];
}

// Largest eigenvalue of the unit-E KE by power iteration on the element
// operator.  Divided by the diagonal kd it bounds lambda_max(D^-1 A) for any
// assembly of scaled copies of KE (Rayleigh-quotient mediant over elements),
// which is the upper end of the Chebyshev smoothing interval.
inline double elem_lambda_max(double nu) {
  ElemConst<double> K = elem_const(nu);
  double u[8][3], y[8][3];
  for (int q = 0; q < 8; ++q)
    for (int c = 0; c < 3; ++c) u[q][c] = 1.0 + 0.37 * q - 0.21 * c + 0.05 * q * c * (q & 1 ? -1 : 1);
  double lam = 0.0;
  for (int it = 0; it < 500; ++it) {
    elem_apply<double>(u, 1.0, K, y);
    double nrm = 0.0, ray = 0.0;
    for (int q = 0; q < 8; ++q)
      for (int c = 0; c < 3; ++c) {
        nrm += y[q][c] * y[q][c];
        ray += u[q][c] * y[q][c];
      }
    double un = 0.0;
    for (int q = 0; q < 8; ++q)
      for (int c = 0; c < 3; ++c) un += u[q][c] * u[q][c];
    lam = ray / un;
    nrm = std::sqrt(nrm);
    for (int q = 0; q < 8; ++q)
      for (int c = 0; c < 3; ++c) u[q][c] = y[q][c] / nrm;
  }
  return lam;
}

// Reference local node k (core/fea.py:33-36) -> binary corner index.
VX_HD inline int ref_to_bin(int k) {
  const int tab[8] = {0, 1, 3, 2, 4, 5, 7, 6};
  return tab[k];
}

}  // namespace vx
