// Shape functions at a Gauss point for the element kernels whose lane is
// (Gauss point q, node a) in an 8 x 8 layout.
//   S = 8: linhexshl (eval_shape_FEM.f90:148-196) at the truncated 8-point
//          rule of genGPandGWtetpri (genGP_GW.f90:355-389), unit weights.
//   S = 4: lintetshl (eval_shape_FEM.f90:201-225) at the 4-point rule
//          (a, b, b), (b, a, b), (b, b, a), (b, b, b), weights 1/24.
//   S = 6: linprishl (eval_shape_FEM.f90:438-477) at the 6-point rule of
//          genGPandGWtetpri (triangle points (2/3, 1/6), (1/6, 2/3),
//          (1/6, 1/6) times zeta = -+0.5773502691896257645091488, weights
//          0.1666666666666667, the reference's literals).
// Node lanes a >= S get zero value and gradient (sums over the 8 lanes keep
// their value); point lanes q >= S repeat the first points with weight 0.
// NGAUSS = NSHL for all three.
#pragma once

#include <hip/hip_runtime.h>

#include "wave_dpp.h"

namespace ofd {

template <int S>
__device__ __forceinline__ void elem_shape(int q, int a, double &Na, double &dl0, double &dl1,
                                           double &dl2, double &gwq) {
  if constexpr (S == 8) {
    const double gpt = 0.57735026918963;
    const double xi = (q & 1) ? gpt : -gpt;
    const double eta = (q & 2) ? gpt : -gpt;
    const double zeta = (q & 4) ? gpt : -gpt;
    // corner signs of node a: (-,+,-) (+,+,-) (+,+,+) (-,+,+) (-,-,-) (+,-,-)
    // (+,-,+) (-,-,+)
    const double sx = ((a + 1) & 2) ? 1.0 : -1.0;
    const double sy = (a < 4) ? 1.0 : -1.0;
    const double sz = (a & 2) ? 1.0 : -1.0;
    const double fx = (sx > 0) ? (xi + 1.0) : (1.0 - xi);
    const double fy = (sy > 0) ? (eta + 1.0) : (1.0 - eta);
    const double fz = (sz > 0) ? (zeta + 1.0) : (1.0 - zeta);
    Na = fx * fy * fz * 0.125;
    dl0 = sx * (fy * fz * 0.125);
    dl1 = sy * (fx * fz * 0.125);
    dl2 = sz * (fx * fy * 0.125);
    gwq = 1.0;
  } else if constexpr (S == 6) {
    const double ta = 0.6666666666666667, tb = 0.1666666666666667;
    const double tz = 0.5773502691896257645091488;
    const int q6 = q < 6 ? q : q - 6;
    const double zi = (q6 % 3 == 0) ? ta : tb;
    const double eta = (q6 % 3 == 1) ? ta : tb;
    const double zeta = (q6 < 3) ? -tz : tz;
    const double lo = 1.0 - zeta, hi = 1.0 + zeta, z3 = 1.0 - zi - eta;
    // N_1..N_6 and their derivatives in (zi, eta, zeta), the node order of
    // linprishl: nodes 1-3 at zeta = -1, 4-6 at zeta = +1
    Na = (a == 0) ? 0.5 * zi * lo : (a == 1) ? 0.5 * eta * lo : (a == 2) ? 0.5 * z3 * lo
       : (a == 3) ? 0.5 * zi * hi : (a == 4) ? 0.5 * eta * hi : (a == 5) ? 0.5 * z3 * hi : 0.0;
    dl0 = (a == 0) ? 0.5 * lo : (a == 2) ? -0.5 * lo : (a == 3) ? 0.5 * hi
        : (a == 5) ? -0.5 * hi : 0.0;
    dl1 = (a == 1) ? 0.5 * lo : (a == 2) ? -0.5 * lo : (a == 4) ? 0.5 * hi
        : (a == 5) ? -0.5 * hi : 0.0;
    dl2 = (a == 0) ? -0.5 * zi : (a == 1) ? -0.5 * eta : (a == 2) ? -0.5 * z3
        : (a == 3) ? 0.5 * zi : (a == 4) ? 0.5 * eta : (a == 5) ? 0.5 * z3 : 0.0;
    gwq = (q < 6) ? 0.1666666666666667 : 0.0;
  } else {
    const double ga_ = 0.5854101966249685, gb_ = 0.1381966011250105;
    const int q4 = q & 3;
    const double xi = (q4 == 0) ? ga_ : gb_;
    const double eta = (q4 == 1) ? ga_ : gb_;
    const double zeta = (q4 == 2) ? ga_ : gb_;
    Na = (a == 0) ? xi : (a == 1) ? eta : (a == 2) ? zeta : (a == 3) ? 1.0 - xi - eta - zeta : 0.0;
    dl0 = (a == 0) ? 1.0 : (a == 3) ? -1.0 : 0.0;
    dl1 = (a == 1) ? 1.0 : (a == 3) ? -1.0 : 0.0;
    dl2 = (a == 2) ? 1.0 : (a == 3) ? -1.0 : 0.0;
    gwq = (q < 4) ? 0.2500000000000000 / 6.0 : 0.0;
  }
}

// ---- sums over an element's S nodes or S Gauss points, for the kernels
// whose lane is (element, Gauss point) ----

// the sum of S products in the lane-butterfly association (S = 8:
// ((p0+p1)+(p2+p3))+((p4+p5)+(p6+p7)); S = 6: ((p0+p1)+(p2+p3))+(p4+p5);
// S = 4: (p0+p1)+(p2+p3)); the adds carry no contraction, so no product is
// fused into them
template <int S>
__device__ __forceinline__ double tree_sum(const double (&p)[S]) {
#pragma clang fp contract(off)
  if constexpr (S == 8)
    return ((p[0] + p[1]) + (p[2] + p[3])) + ((p[4] + p[5]) + (p[6] + p[7]));
  else if constexpr (S == 6)
    return ((p[0] + p[1]) + (p[2] + p[3])) + (p[4] + p[5]);
  else
    return (p[0] + p[1]) + (p[2] + p[3]);
}

// lanes per element in the kernels whose lane is (element, Gauss point):
// the S points, prisms padded to 8 (their lanes 6, 7 carry zero weight)
template <int S>
constexpr int point_lanes() { return S == 4 ? 4 : 8; }

// sum and maximum over the lanes of an element's Gauss points
template <int S>
__device__ __forceinline__ double point_sum(double v) {
  if constexpr (S == 4) return wsum4(v);
  else return wsum8(v);
}
template <int S>
__device__ __forceinline__ double point_max(double v) {
  if constexpr (S == 4) return wmax4(v);
  else return wmax8(v);
}

// the shape value of node a at Gauss point q (elem_shape<S>'s Na)
template <int S>
__device__ __forceinline__ double shape_N(int q, int a) {
  double Na, d0, d1, d2, w;
  elem_shape<S>(q, a, Na, d0, d1, d2, w);
  return Na;
}

// the sum of S products in node order, (((p0 + p1) + p2) + ...), the order of
// the reference's e3int_qr; no product is fused into the adds
template <int S>
__device__ __forceinline__ double seq_sum(const double (&p)[S]) {
#pragma clang fp contract(off)
  double s = p[0];
#pragma unroll
  for (int a = 1; a < S; ++a) s = s + p[a];
  return s;
}

}  // namespace ofd
