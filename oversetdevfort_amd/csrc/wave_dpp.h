// Sums and maxima over groups of 8 lanes of a 64-lane wavefront by DPP
// moves (the VALU's own lane network, no LDS traffic), for the element
// kernels whose lane = (Gauss point, node) layout reduces over the 8 lanes
// 8k..8k+7.
#pragma once

#include <hip/hip_runtime.h>

namespace ofd {

// A double moved across lanes by a DPP modifier: CTRL 0xB1 = quad_perm
// [1,0,3,2] (lane ^ 1), 0x4E = quad_perm [2,3,0,1] (lane ^ 2), 0x141 =
// row_half_mirror (lane i of 8 takes lane 7 - i).  Every lane has a source
// for these three, so no lane keeps the `old` operand.
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
  static_assert(CTRL == 0xB1 || CTRL == 0x4E || CTRL == 0x141, "unchecked DPP control");
  union {
    double d;
    int i[2];
  } u;
  u.d = v;
  u.i[0] = __builtin_amdgcn_update_dpp(0, u.i[0], CTRL, 0xF, 0xF, false);
  u.i[1] = __builtin_amdgcn_update_dpp(0, u.i[1], CTRL, 0xF, 0xF, false);
  return u.d;
}

// Sum over the 8 lanes 8k..8k+7, the same bits in every lane and the same
// bits as the xor butterfly over lanes ^1, ^2, ^4: lanes i and i ^ 1 add in
// either order (exact commutation), then the quad sums, then the two quads,
// i.e. the association ((v0+v1)+(v2+v3))+((v4+v5)+(v6+v7)).  The adds carry
// no contraction: were a product passed in fused into the first add, each
// lane would add its own product unrounded to its partner's rounded one,
// and the lanes would disagree.
__device__ __forceinline__ double wsum8(double v) {
#pragma clang fp contract(off)
  v += dpp_f64<0xB1>(v);
  v += dpp_f64<0x4E>(v);
  v += dpp_f64<0x141>(v);
  return v;
}

// the same over the 4 lanes 4k..4k+3: (v0+v1)+(v2+v3)
__device__ __forceinline__ double wsum4(double v) {
#pragma clang fp contract(off)
  v += dpp_f64<0xB1>(v);
  v += dpp_f64<0x4E>(v);
  return v;
}

__device__ __forceinline__ double wmax4(double v) {
  v = fmax(v, dpp_f64<0xB1>(v));
  v = fmax(v, dpp_f64<0x4E>(v));
  return v;
}

__device__ __forceinline__ double wmax8(double v) {
  v = fmax(v, dpp_f64<0xB1>(v));
  v = fmax(v, dpp_f64<0x4E>(v));
  v = fmax(v, dpp_f64<0x141>(v));
  return v;
}

}  // namespace ofd
