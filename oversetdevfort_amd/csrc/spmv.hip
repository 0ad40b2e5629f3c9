// Block-CSR SpMV and block-Jacobi preconditioner on gfx950.
//
// SparseProdUP_3D (SparseProd.f90:4-60) computes, for every node row,
// y_i = sum_k [K11 G; D1 M]_k [u; p]_{row(k)}.  Here the four reference
// arrays are one 4x4 row-major block per nonzero (128 B in FP64, 64 B in
// FP32), kept in the reference's nonzero order, and vectors are node-
// interleaved [N][4].
//
// Mapping: 4 lanes per block row, lane c computes output component c, so a
// 64-lane wave covers 16 rows.  Per nonzero the 4 lanes of a row read the 4
// rows of the block -- 128 contiguous bytes, one full cache line -- and all
// 4 read the same 32-byte x entry (broadcast).  The matrix is streamed once
// at HBM rate; x is gathered from the 27 neighbours and stays in L2.
//
// The block-Jacobi inverse (SparseGMRES_NS.f90:53-121) is kept as a 4x4
// block-diagonal [Binv 0; 0 1/M] so that applying it is a quad-lane 4x4
// product fused into the SpMV epilogue (Arnoldi step z = P^-1 A v).
#include "gmres_dev.h"

namespace ofd {

namespace {

template <typename T>
struct Row4;
template <>
struct Row4<double> {
  static __device__ __forceinline__ void load(const double *p, double &a0, double &a1,
                                              double &a2, double &a3) {
    const double2 u = reinterpret_cast<const double2 *>(p)[0];
    const double2 v = reinterpret_cast<const double2 *>(p)[1];
    a0 = u.x;
    a1 = u.y;
    a2 = v.x;
    a3 = v.y;
  }
};
template <>
struct Row4<float> {
  static __device__ __forceinline__ void load(const float *p, double &a0, double &a1,
                                              double &a2, double &a3) {
    const float4 u = reinterpret_cast<const float4 *>(p)[0];
    a0 = u.x;
    a1 = u.y;
    a2 = u.z;
    a3 = u.w;
  }
};

__device__ __forceinline__ void load_x4(const double *x, double &x0, double &x1,
                                        double &x2, double &x3) {
  const double2 u = reinterpret_cast<const double2 *>(x)[0];
  const double2 v = reinterpret_cast<const double2 *>(x)[1];
  x0 = u.x;
  x1 = u.y;
  x2 = v.x;
  x3 = v.y;
}

}  // namespace

// mode 0: y = A x; mode 1: y = b - A x; mode 2: Arnoldi product
// y = s_k P^-1 (A W_k), whose prologue first closes Arnoldi step k-1 from the
// partials the previous sweep left in `part` (k = 0: the initial residual)
template <typename T, int MODE>
__global__ void __launch_bounds__(256)
k_spmv(int N, const int32_t *__restrict__ col, const int32_t *__restrict__ row,
       const T *__restrict__ A, const double *__restrict__ x,
       double *__restrict__ y, const double *__restrict__ b,
       const T *__restrict__ Pinv, GmresState *__restrict__ gs, int kidx,
       const double *__restrict__ part, int nb, int cgs, int32_t *hflag) {
  double s = 1.0;
  if (MODE == 2) {
    if (gs->done) return;
    if (arnoldi_close(gs, part, nb, kidx - 1, cgs, hflag, &s)) return;
  }
  const int gt = blockIdx.x * blockDim.x + threadIdx.x;
  const int r = gt >> 2;
  const int c = gt & 3;
  if (r >= N) return;
  const int k0 = col[r], k1 = col[r + 1];
  double acc = 0.0;
#pragma unroll 4
  for (int k = k0; k < k1; ++k) {
    const int cn = row[k];
    double a0, a1, a2, a3, x0, x1, x2, x3;
    Row4<T>::load(A + (size_t)k * 16 + c * 4, a0, a1, a2, a3);
    load_x4(x + (size_t)cn * 4, x0, x1, x2, x3);
    acc += a0 * x0 + a1 * x1 + a2 * x2 + a3 * x3;
  }
  if (MODE == 0) {
    y[(size_t)r * 4 + c] = acc;
  } else if (MODE == 1) {
    y[(size_t)r * 4 + c] = b[(size_t)r * 4 + c] - acc;
  } else {
    const int base = threadIdx.x & ~3;
    const double z0 = __shfl(acc, base + 0);
    const double z1 = __shfl(acc, base + 1);
    const double z2 = __shfl(acc, base + 2);
    const double z3 = __shfl(acc, base + 3);
    double p0, p1, p2, p3;
    Row4<T>::load(Pinv + (size_t)r * 16 + c * 4, p0, p1, p2, p3);
    y[(size_t)r * 4 + c] = s * (p0 * z0 + p1 * z1 + p2 * z2 + p3 * z3);
  }
}

// block-Jacobi inverse: 3x3 cofactor inverse of the K11 diagonal block and
// 1/M (0 -> 1), SparseGMRES_NS.f90:73-121
__global__ void k_build_prec(int N, const int32_t *__restrict__ diag,
                             const double *__restrict__ A, double *__restrict__ P,
                             float *__restrict__ P32) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const double *blk = A + (size_t)diag[i] * 16;
  double B[9];
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) B[r * 3 + c] = blk[r * 4 + c];
  double Bi[9];
  Bi[0] = B[4] * B[8] - B[7] * B[5];
  Bi[1] = B[7] * B[2] - B[1] * B[8];
  Bi[2] = B[1] * B[5] - B[2] * B[4];
  const double tmp2 = Bi[0] * B[0] + Bi[1] * B[3] + Bi[2] * B[6];
  const double tmp = 1.0 / tmp2;
  Bi[0] *= tmp;
  Bi[1] *= tmp;
  Bi[2] *= tmp;
  Bi[3] = (B[5] * B[6] - B[3] * B[8]) * tmp;
  Bi[4] = (B[0] * B[8] - B[6] * B[2]) * tmp;
  Bi[5] = (B[3] * B[2] - B[0] * B[5]) * tmp;
  Bi[6] = (B[3] * B[7] - B[4] * B[6]) * tmp;
  Bi[7] = (B[6] * B[1] - B[0] * B[7]) * tmp;
  Bi[8] = (B[0] * B[4] - B[1] * B[3]) * tmp;
  const double m = blk[15];
  const double mi = (m == 0.0) ? 1.0 : 1.0 / m;
  double out[16];
#pragma unroll
  for (int r = 0; r < 3; ++r) {
#pragma unroll
    for (int c = 0; c < 3; ++c) out[r * 4 + c] = Bi[r * 3 + c];
    out[r * 4 + 3] = 0.0;
  }
  out[12] = out[13] = out[14] = 0.0;
  out[15] = mi;
  double2 *o = reinterpret_cast<double2 *>(P + (size_t)i * 16);
#pragma unroll
  for (int t = 0; t < 8; ++t) o[t] = make_double2(out[2 * t], out[2 * t + 1]);
  if (P32) {
    float4 *o32 = reinterpret_cast<float4 *>(P32 + (size_t)i * 16);
#pragma unroll
    for (int t = 0; t < 4; ++t)
      o32[t] = make_float4((float)out[4 * t], (float)out[4 * t + 1],
                           (float)out[4 * t + 2], (float)out[4 * t + 3]);
  }
}

__global__ void k_convert_a32(size_t n4, const double *__restrict__ A,
                              float *__restrict__ A32) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n4; i += stride) {
    const double2 u = reinterpret_cast<const double2 *>(A)[2 * i];
    const double2 v = reinterpret_cast<const double2 *>(A)[2 * i + 1];
    reinterpret_cast<float4 *>(A32)[i] =
        make_float4((float)u.x, (float)u.y, (float)v.x, (float)v.y);
  }
}

hipError_t launch_build_prec(Ctx &c, bool want32) {
  const int N = c.nnode;
  hipLaunchKernelGGL(k_build_prec, dim3((N + 255) / 256), dim3(256), 0, c.stream, N,
                     c.d_diag, c.d_A64, c.d_P64, want32 ? c.d_P32 : nullptr);
  return hipGetLastError();
}

hipError_t launch_convert_a32(Ctx &c) {
  const size_t n4 = (size_t)c.icnt * 4;
  size_t blocks = (n4 + 255) / 256;
  if (blocks > 16384) blocks = 16384;
  hipLaunchKernelGGL(k_convert_a32, dim3((unsigned)blocks), dim3(256), 0, c.stream,
                     n4, c.d_A64, c.d_A32);
  return hipGetLastError();
}

static inline unsigned spmv_blocks(const Ctx &c) {
  return (unsigned)(((size_t)c.nnode * 4 + 255) / 256);
}

hipError_t launch_spmv(Ctx &c, int precision, const double *x, double *y,
                       const double *b, int mode) {
  const unsigned nb = spmv_blocks(c);
  if (precision == OFD_FP32) {
    if (mode == 0)
      hipLaunchKernelGGL((k_spmv<float, 0>), dim3(nb), dim3(256), 0, c.stream, c.nnode,
                         c.d_col, c.d_row, c.d_A32, x, y, b, c.d_P32, c.d_gs, 0,
                         nullptr, 0, 0, nullptr);
    else
      hipLaunchKernelGGL((k_spmv<float, 1>), dim3(nb), dim3(256), 0, c.stream, c.nnode,
                         c.d_col, c.d_row, c.d_A32, x, y, b, c.d_P32, c.d_gs, 0,
                         nullptr, 0, 0, nullptr);
  } else {
    if (mode == 0)
      hipLaunchKernelGGL((k_spmv<double, 0>), dim3(nb), dim3(256), 0, c.stream,
                         c.nnode, c.d_col, c.d_row, c.d_A64, x, y, b, c.d_P64,
                         c.d_gs, 0, nullptr, 0, 0, nullptr);
    else
      hipLaunchKernelGGL((k_spmv<double, 1>), dim3(nb), dim3(256), 0, c.stream,
                         c.nnode, c.d_col, c.d_row, c.d_A64, x, y, b, c.d_P64,
                         c.d_gs, 0, nullptr, 0, 0, nullptr);
  }
  return hipGetLastError();
}

// Arnoldi product: W_{k+1} = s_k P^-1 A W_k (closes step k-1 first)
hipError_t launch_spmv_prec(Ctx &c, int precision, int k, int ortho) {
  const unsigned nb = spmv_blocks(c);
  const size_t vs = (size_t)c.nnode * 4;
  const double *x = c.d_W + vs * k;
  double *y = c.d_W + vs * (k + 1);
  const double *part =
      c.d_part + (size_t)kMaxRedBlocks * kPartStride * close_buffer(k - 1, ortho);
  const int pnb = nblocks_nodes(c);
  const int cgs = ortho == OFD_ORTHO_MGS ? 0 : 1;
  if (precision == OFD_FP32)
    hipLaunchKernelGGL((k_spmv<float, 2>), dim3(nb), dim3(256), 0, c.stream, c.nnode,
                       c.d_col, c.d_row, c.d_A32, x, y, nullptr, c.d_P32, c.d_gs, k, part,
                       pnb, cgs, c.h_flags);
  else
    hipLaunchKernelGGL((k_spmv<double, 2>), dim3(nb), dim3(256), 0, c.stream, c.nnode,
                       c.d_col, c.d_row, c.d_A64, x, y, nullptr, c.d_P64, c.d_gs, k, part,
                       pnb, cgs, c.h_flags);
  return hipGetLastError();
}

}  // namespace ofd
