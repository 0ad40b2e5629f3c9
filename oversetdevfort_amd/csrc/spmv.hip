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
// block-diagonal [Binv 0; 0 1/M].  Left preconditioning is applied once per
// assembled matrix: the Krylov loop runs on the stored product P^-1 A (same
// pattern, FP64 and FP32 copies), so its SpMV has no epilogue.
#include <hip/hip_ext.h>

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

template <typename T>
struct Pair;
template <>
struct Pair<double> {
  static __device__ __forceinline__ void load(const double *p, double &a0, double &a1) {
    const double2 u = reinterpret_cast<const double2 *>(p)[0];
    a0 = u.x;
    a1 = u.y;
  }
};
template <>
struct Pair<float> {
  static __device__ __forceinline__ void load(const float *p, double &a0, double &a1) {
    const float2 u = reinterpret_cast<const float2 *>(p)[0];
    a0 = u.x;
    a1 = u.y;
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

__device__ __forceinline__ void load_x4(const float *x, double &x0, double &x1,
                                        double &x2, double &x3) {
  const float4 u = reinterpret_cast<const float4 *>(x)[0];
  x0 = u.x;
  x1 = u.y;
  x2 = u.z;
  x3 = u.w;
}

}  // namespace

// mode 0: y = A x; mode 1: y = b - A x; mode 2: Arnoldi product y = A x that
// returns at once when the cycle has converged.  T: matrix entries, V:
// vector entries (accumulation in FP64 either way).
template <typename T, typename V, int MODE>
__global__ void __launch_bounds__(256)
k_spmv(int N, const int32_t *__restrict__ col, const int32_t *__restrict__ row,
       const T *__restrict__ A, const V *__restrict__ x, V *__restrict__ y,
       const double *__restrict__ b, const GmresState *__restrict__ gs) {
  if (MODE == 2 && gs->done) return;
  // restarted solve: the residual of a cycle queued behind the one that met
  // the tolerance is the one already computed
  if (MODE == 1 && gs && gs->outer_done) return;
  // Workgroups are dispatched round-robin over the 8 XCDs, each with its own
  // L2.  The grid is a multiple of 8 and XCD j takes the j-th contiguous
  // eighth of the rows, so neighbouring rows, which gather the same x
  // entries, share one L2 instead of every L2 fetching all of x.
  const int per = gridDim.x >> 3;
  const int lb = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
  // rows are spread evenly over the grid: rpb <= 64 per 256-thread block
  const int rpb = (N + (int)gridDim.x - 1) / (int)gridDim.x;
  const int rb = threadIdx.x >> 2;
  const int r = lb * rpb + rb;
  const int c = threadIdx.x & 3;
  if (rb >= rpb || r >= N) return;
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
  if (MODE == 1)
    y[(size_t)r * 4 + c] = (V)(b[(size_t)r * 4 + c] - acc);
  else
    y[(size_t)r * 4 + c] = (V)acc;
}

// Scalar CSR product of the temperature system (SparseProd_LS,
// SparseProd.f90:128-161): mode 0 y = A x, 1 y = b - A x, 2 the Arnoldi
// product (returns at once once the cycle has converged).  8 lanes per row:
// lane c takes nonzeros c, c+8, ... of the row, the first 32 as four
// independent (value, column) loads issued before any x is gathered, so a
// wave keeps its 8 rows' contiguous nonzeros in flight; the lanes are added
// in a fixed shuffle order.  Rows n .. npad-1 (the zero padding to whole
// 4-tuples) are written as 0.  The same XCD-aware block order as the block
// product.
template <typename T, typename V, int MODE>
__global__ void __launch_bounds__(256)
k_spmv_s(int n, int npad, const int32_t *__restrict__ col, const int32_t *__restrict__ row,
         const T *__restrict__ A, const V *__restrict__ x, V *__restrict__ y,
         const double *__restrict__ b, const GmresState *__restrict__ gs) {
  if (MODE == 2 && gs->done) return;
  if (MODE == 1 && gs && gs->outer_done) return;
  const int per = gridDim.x >> 3;
  const int lb = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
  const int gt = lb * blockDim.x + threadIdx.x;
  const int r = gt >> 3;
  const int c = gt & 7;
  if (r >= npad) return;  // uniform over each group of 8 lanes
  double acc = 0.0;
  if (r < n) {
    const int k0 = col[r] + c, k1 = col[r + 1];
    T av[4];
    int jv[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int k = k0 + 8 * u;
      const bool in = k < k1;
      av[u] = in ? A[k] : (T)0;
      jv[u] = in ? row[k] : -1;
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (jv[u] >= 0) acc += (double)av[u] * (double)x[jv[u]];
    for (int k = k0 + 32; k < k1; k += 8) acc += (double)A[k] * (double)x[row[k]];
  }
  acc += __shfl_xor(acc, 1);
  acc += __shfl_xor(acc, 2);
  acc += __shfl_xor(acc, 4);
  if (c == 0) {
    if (r >= n)
      y[r] = (V)0.0;
    else if (MODE == 1)
      y[r] = (V)(b[r] - acc);
    else
      y[r] = (V)acc;
  }
}

// inverse diagonal of the temperature matrix, 0 -> 1 (SparseGMRES_tem.f90:
// 38-59), and the preconditioned copy Ap[k] = dinv_i A[k] (FP64, FP32); 8
// lanes per row stream the row's contiguous nonzeros
__global__ void __launch_bounds__(256)
k_build_prec_s(int N, const int32_t *__restrict__ col, const int32_t *__restrict__ diag,
               const double *__restrict__ A, double *__restrict__ P, double *__restrict__ Ap,
               float *__restrict__ Ap32) {
  const int gt = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = gt >> 3, c = gt & 7;
  if (i >= N) return;
  const double d = A[diag[i]];
  const double di = (d != 0.0) ? 1.0 / d : 1.0;
  if (c == 0) P[i] = di;
  for (int k = col[i] + c; k < col[i + 1]; k += 8) {
    const double v = di * A[k];
    Ap[k] = v;
    if (Ap32) Ap32[k] = (float)v;
  }
}

__global__ void k_convert_s32(size_t n, const double *__restrict__ A, float *__restrict__ A32) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) A32[i] = (float)A[i];
}

// Product of the coupled NS + VOF system (SparseProd_NS_conv, SparseProd.f90:
// 167-242).  A nonzero k is the 4x4 block A[k][16] [K11 G; D1 M] row-major
// (as the NS system) and the phi couplings Ap[k][8] as pairs (col phi of row
// r, row phi of col r): (ULS(r), LSU(r)) for r < 3, (PLS, LS); the pair
// array follows the Z blocks in the same allocation.  Vectors are [N][4]
// (u0 u1 u2 p) followed by [N] phi, zero-padded to whole 4-tuples (npad
// entries).
//
// 8 lanes per block row (kVStreams = 2 streams of 4): lanes h*4 + c,
// h = 0, 1, take the row's nonzeros k0 + h, k0 + h + 2, ...; lane c reads
// row c of the 4x4 block (32 / 16 bytes) and its pair (16 / 8 bytes) and
// forms output c, so the 4 lanes read a nonzero's aligned 128 / 64-byte
// block and 64 / 32-byte pair group; it also forms the part of the phi row
// that multiplies x_c
// (LSU(c) x_uc, or LS x_phi on lane 3).  The two streams are added, then the
// four phi parts, in a fixed shuffle order.
// Two streams per row double the loads in flight per row and spread the
// 48^3 tank's 117,649 rows over about three rounds of resident waves
// instead of one and a half: the 48^3 tank's FP64 product went from 135.6 to
// 120.7 us, the FP32-matrix one from 96.5 to 74.9 us (four streams: 117.3
// and 73.5 us, no gain in the solver).  The 8 lanes after the last row
// write the padding zeros.  mode as k_spmv.
// Records of 24 contiguous entries (192 / 96 bytes) straddle cache lines;
// with the blocks and pair groups in separate aligned arrays the FP32
// Arnoldi product of the tank went from 81.8 to 77.5 us (FP64: 126 us
// either way).
constexpr int kVStreams = 2;  // nonzero streams per row (4 lanes each)
template <typename T, typename V, int MODE>
__global__ void __launch_bounds__(256)
k_spmv_v(int N, int npad, const int32_t *__restrict__ col, const int32_t *__restrict__ row,
         const T *__restrict__ A, const T *__restrict__ Ap, const V *__restrict__ x,
         V *__restrict__ y, const double *__restrict__ b, const GmresState *__restrict__ gs) {
  if (MODE == 2 && gs->done) return;
  if (MODE == 1 && gs && gs->outer_done) return;
  const int per = gridDim.x >> 3;
  const int lb = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
  const int gt = lb * blockDim.x + threadIdx.x;
  const int r = gt / (4 * kVStreams);
  const int h = (gt >> 2) & (kVStreams - 1);
  const int c = gt & 3;
  if (r > N) return;  // uniform over each group of 4 kVStreams lanes
  if (r == N) {
    const int t = 5 * N + c;
    if (h == 0 && t < npad) y[t] = (V)0.0;
    return;
  }
  const V *xphi = x + (size_t)4 * N;
  const int k0 = col[r] + h, k1 = col[r + 1];
  double acc = 0.0, accp = 0.0;
#pragma unroll 4
  for (int k = k0; k < k1; k += kVStreams) {
    const int cn = row[k];
    double a0, a1, a2, a3, x0, x1, x2, x3, e0, e1;
    Row4<T>::load(A + (size_t)k * 16 + c * 4, a0, a1, a2, a3);
    Pair<T>::load(Ap + (size_t)k * 8 + c * 2, e0, e1);
    load_x4(x + (size_t)cn * 4, x0, x1, x2, x3);
    const double xp = (double)xphi[cn];
    acc += a0 * x0 + a1 * x1 + a2 * x2 + a3 * x3 + e0 * xp;
    const double xc = (c == 0) ? x0 : ((c == 1) ? x1 : ((c == 2) ? x2 : xp));
    accp += e1 * xc;
  }
#pragma unroll
  for (int m = 4; m < 4 * kVStreams; m <<= 1) {
    acc += __shfl_xor(acc, m);
    accp += __shfl_xor(accp, m);
  }
  accp += __shfl_xor(accp, 1);
  accp += __shfl_xor(accp, 2);
  if (h) return;
  if (MODE == 1) {
    y[(size_t)r * 4 + c] = (V)(b[(size_t)r * 4 + c] - acc);
    if (c == 0) y[(size_t)4 * N + r] = (V)(b[(size_t)4 * N + r] - accp);
  } else {
    y[(size_t)r * 4 + c] = (V)acc;
    if (c == 0) y[(size_t)4 * N + r] = (V)accp;
  }
}

// phi part of the NSVOF block-Jacobi preconditioner (SparseGMRES_NS_conv.f:
// 53-150): 1/LS with 0 -> 0; the u-p part is the NS block inverse of the 4x4
// diagonal block (3x3 cofactor inverse of K11, 1/M with 0 -> 1)
__global__ void k_build_prec_phi(int N, const int32_t *__restrict__ diag,
                                 const double *__restrict__ Ap, double *__restrict__ pphi) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const double l = Ap[(size_t)diag[i] * 8 + 7];
  pphi[i] = (l == 0.0) ? 0.0 : 1.0 / l;
}

// the couplings of P^-1 A: thread (i, c) forms, for every nonzero k of block
// row i, entry (row c, col phi) = sum_s P_i(c, s) e[2s] and entry (row phi,
// col c) = pphi_i e[2c+1], e = the pair group Ae[k][8]; into the pair groups
// Ap (and Ap32) of the preconditioned copy
__global__ void k_precondition_phi(int N, const int32_t *__restrict__ col,
                                   const double *__restrict__ P, const double *__restrict__ pphi,
                                   const double *__restrict__ Ae, double *__restrict__ Ap,
                                   float *__restrict__ Ap32) {
  const int gt = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = gt >> 2, c = gt & 3;
  if (i >= N) return;
  const double *p = P + (size_t)i * 16 + c * 4;
  const double p0 = p[0], p1 = p[1], p2 = p[2], p3 = p[3], pf = pphi[i];
  for (int k = col[i]; k < col[i + 1]; ++k) {
    const double *a = Ae + (size_t)k * 8;
    const double o0 = p0 * a[0] + p1 * a[2] + p2 * a[4] + p3 * a[6];
    const double o1 = pf * a[2 * c + 1];
    reinterpret_cast<double2 *>(Ap + (size_t)k * 8)[c] = make_double2(o0, o1);
    if (Ap32)
      reinterpret_cast<float2 *>(Ap32 + (size_t)k * 8)[c] = make_float2((float)o0, (float)o1);
  }
}

// block-Jacobi inverse: 3x3 cofactor inverse of the K11 diagonal block and
// 1/M (0 -> 1), SparseGMRES_NS.f90:73-121
// (blocks of `stride` doubles whose first 16 are the 4x4 block)
__global__ void k_build_prec(int N, const int32_t *__restrict__ diag,
                             const double *__restrict__ A, int stride, double *__restrict__ P) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const double *blk = A + (size_t)diag[i] * stride;
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
}

// Ap[k] = P_i^-1 A[k] for every nonzero k of block row i; thread (i, c)
// forms row c of the row's 4x4 blocks (the first 16 of `stride` doubles)
__global__ void k_precondition_matrix(int N, const int32_t *__restrict__ col,
                                      const double *__restrict__ P,
                                      const double *__restrict__ A, int stride,
                                      double *__restrict__ Ap, float *__restrict__ Ap32) {
  const int gt = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = gt >> 2, c = gt & 3;
  if (i >= N) return;
  const double *p = P + (size_t)i * 16 + c * 4;
  const double p0 = p[0], p1 = p[1], p2 = p[2], p3 = p[3];
  for (int k = col[i]; k < col[i + 1]; ++k) {
    const double *a = A + (size_t)k * stride;
    double o[4];
#pragma unroll
    for (int q = 0; q < 4; ++q)
      o[q] = p0 * a[0 + q] + p1 * a[4 + q] + p2 * a[8 + q] + p3 * a[12 + q];
    double2 *d = reinterpret_cast<double2 *>(Ap + (size_t)k * stride + c * 4);
    d[0] = make_double2(o[0], o[1]);
    d[1] = make_double2(o[2], o[3]);
    if (Ap32)
      reinterpret_cast<float4 *>(Ap32 + (size_t)k * stride + c * 4)[0] =
          make_float4((float)o[0], (float)o[1], (float)o[2], (float)o[3]);
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

hipError_t launch_build_prec(Ctx &c) {
  const int N = c.nnode;
  hipLaunchKernelGGL(k_build_prec, dim3((N + 255) / 256), dim3(256), 0, c.stream, N,
                     c.d_diag, c.d_A64, 16, c.d_P64);
  return hipGetLastError();
}

hipError_t launch_precondition_matrix(Ctx &c, bool want32) {
  const unsigned nb = (unsigned)(((size_t)c.nnode * 4 + 255) / 256);
  hipLaunchKernelGGL(k_precondition_matrix, dim3(nb), dim3(256), 0, c.stream, c.nnode,
                     c.d_col, c.d_P64, c.d_A64, 16, c.d_Ap64, want32 ? c.d_Ap32 : nullptr);
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

// 4 lanes per row, 256-thread blocks: at least ceil(N/64) blocks, rounded up
// to whole rounds over the compute units, so every CU holds the same number
// of blocks and the rows are spread evenly over them (k_spmv takes
// ceil(N/grid) <= 64 rows per block).  ceil(N/64) blocks as such leave some
// CUs one block more than the others, and the product takes as long as
// those do: 46.2 -> 43.3 us for the FP64 Arnoldi product of the 64x32x32
// channel.  A multiple of 8 (one share per XCD).
static inline unsigned spmv_blocks(const Ctx &c) {
  const unsigned need = (unsigned)(((size_t)c.nnode * 4 + 255) / 256);
  const unsigned ncu = (unsigned)c.n_cu;
  const unsigned nb = ((need + ncu - 1) / ncu) * ncu;
  return (nb + 7u) & ~7u;
}

hipError_t launch_spmv(Ctx &c, int precision, const double *x, double *y,
                       const double *b, int mode) {
  const unsigned nb = spmv_blocks(c);
  if (precision == OFD_FP32) {
    if (mode == 0)
      hipLaunchKernelGGL((k_spmv<float, double, 0>), dim3(nb), dim3(256), 0, c.stream, c.nnode,
                         c.d_col, c.d_row, c.d_A32, x, y, b, c.d_gs);
    else
      hipLaunchKernelGGL((k_spmv<float, double, 1>), dim3(nb), dim3(256), 0, c.stream, c.nnode,
                         c.d_col, c.d_row, c.d_A32, x, y, b, c.d_gs);
  } else {
    if (mode == 0)
      hipLaunchKernelGGL((k_spmv<double, double, 0>), dim3(nb), dim3(256), 0, c.stream,
                         c.nnode, c.d_col, c.d_row, c.d_A64, x, y, b, c.d_gs);
    else
      hipLaunchKernelGGL((k_spmv<double, double, 1>), dim3(nb), dim3(256), 0, c.stream,
                         c.nnode, c.d_col, c.d_row, c.d_A64, x, y, b, c.d_gs);
  }
  return hipGetLastError();
}

// Arnoldi product W_{k+1} = (P^-1 A) W_k, unscaled (FP32 basis in the mixed
// CGS solver)
// scalar rows, 8 lanes each, over the padded length, rounded to 8 blocks
static inline unsigned spmv_s_blocks(const Ctx &c) {
  const unsigned nb = (unsigned)(((size_t)c.gm_n * 32 + 255) / 256);
  return (nb + 7u) & ~7u;
}

hipError_t launch_spmv_s(Ctx &c, int precision, const double *x, double *y,
                         const double *b, int mode) {
  const unsigned nb = spmv_s_blocks(c);
  const int n = c.nnode, np = c.gm_n * 4;
  if (precision == OFD_FP32) {
    if (mode == 0)
      hipLaunchKernelGGL((k_spmv_s<float, double, 0>), dim3(nb), dim3(256), 0, c.stream, n,
                         np, c.d_col, c.d_row, c.d_AT32, x, y, b, c.d_gs);
    else
      hipLaunchKernelGGL((k_spmv_s<float, double, 1>), dim3(nb), dim3(256), 0, c.stream, n,
                         np, c.d_col, c.d_row, c.d_AT32, x, y, b, c.d_gs);
  } else {
    if (mode == 0)
      hipLaunchKernelGGL((k_spmv_s<double, double, 0>), dim3(nb), dim3(256), 0, c.stream, n,
                         np, c.d_col, c.d_row, c.d_AT, x, y, b, c.d_gs);
    else
      hipLaunchKernelGGL((k_spmv_s<double, double, 1>), dim3(nb), dim3(256), 0, c.stream, n,
                         np, c.d_col, c.d_row, c.d_AT, x, y, b, c.d_gs);
  }
  return hipGetLastError();
}

hipError_t launch_build_prec_s(Ctx &c, bool want32) {
  const int N = c.nnode;
  hipLaunchKernelGGL(k_build_prec_s, dim3((unsigned)(((size_t)N * 8 + 255) / 256)), dim3(256),
                     0, c.stream, N,
                     c.d_col, c.d_diag, c.d_AT, c.d_PT, c.d_ApT, want32 ? c.d_ApT32 : nullptr);
  return hipGetLastError();
}

hipError_t launch_convert_aT32(Ctx &c) {
  hipLaunchKernelGGL(k_convert_s32, dim3(2048), dim3(256), 0, c.stream, (size_t)c.icnt,
                     c.d_AT, c.d_AT32);
  return hipGetLastError();
}

// 4 kVStreams lanes per node row plus one group for the padding, rounded to
// 8 blocks
static inline unsigned spmv_v_blocks(const Ctx &c) {
  const unsigned nb = (unsigned)(((size_t)(c.nnode + 1) * 4 * kVStreams + 255) / 256);
  return (nb + 7u) & ~7u;
}

// A: an NSVOF matrix allocation (Z blocks of 16, then Z pair groups of 8)
template <typename T, typename V, int MODE>
static void launch_v(Ctx &c, const T *A, const V *x, V *y, const double *b,
                     hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr) {
  const T *Ap = A + (size_t)16 * c.icnt;
  hipExtLaunchKernelGGL((k_spmv_v<T, V, MODE>), dim3(spmv_v_blocks(c)), dim3(256), 0,
                        c.stream, ev0, ev1, 0, c.nnode, c.gm_n * 4, (const int32_t *)c.d_col,
                        (const int32_t *)c.d_row, A, Ap, x, y, b, (const GmresState *)c.d_gs);
}

hipError_t launch_spmv_v(Ctx &c, int precision, const double *x, double *y,
                         const double *b, int mode) {
  if (precision == OFD_FP32) {
    if (mode == 0)
      launch_v<float, double, 0>(c, c.d_AV32, x, y, b);
    else
      launch_v<float, double, 1>(c, c.d_AV32, x, y, b);
  } else {
    if (mode == 0)
      launch_v<double, double, 0>(c, c.d_AV, x, y, b);
    else
      launch_v<double, double, 1>(c, c.d_AV, x, y, b);
  }
  return hipGetLastError();
}

// P (the 4x4 block inverses, [N][16]) and pphi ([N], after it in d_PV); with
// `ap` the preconditioned copy Ap = P^-1 A in the same layout
hipError_t launch_build_prec_v(Ctx &c, bool want32, bool ap) {
  const int N = c.nnode;
  double *P = c.d_PV, *pphi = c.d_PV + (size_t)N * 16;
  const size_t zp = (size_t)16 * c.icnt;  // the pair groups follow the blocks
  hipLaunchKernelGGL(k_build_prec, dim3((N + 255) / 256), dim3(256), 0, c.stream, N, c.d_diag,
                     c.d_AV, 16, P);
  hipLaunchKernelGGL(k_build_prec_phi, dim3((N + 255) / 256), dim3(256), 0, c.stream, N,
                     c.d_diag, c.d_AV + zp, pphi);
  if (!ap) return hipGetLastError();
  const unsigned nb = (unsigned)(((size_t)N * 4 + 255) / 256);
  hipLaunchKernelGGL(k_precondition_matrix, dim3(nb), dim3(256), 0, c.stream, N, c.d_col, P,
                     c.d_AV, 16, c.d_ApV, want32 ? c.d_ApV32 : nullptr);
  hipLaunchKernelGGL(k_precondition_phi, dim3(nb), dim3(256), 0, c.stream, N, c.d_col, P,
                     pphi, c.d_AV + zp, c.d_ApV + zp, want32 ? c.d_ApV32 + zp : nullptr);
  return hipGetLastError();
}

hipError_t launch_convert_aV32(Ctx &c) {
  hipLaunchKernelGGL(k_convert_s32, dim3(4096), dim3(256), 0, c.stream,
                     (size_t)c.icnt * 24, c.d_AV, c.d_AV32);
  return hipGetLastError();
}

// Arnoldi product of the NSVOF system: W_{k+1} = (P^-1 A) W_k, or A P^-1 W_k
static hipError_t launch_spmv_arnoldi_v(Ctx &c, int precision, int k, hipEvent_t ev0,
                                        hipEvent_t ev1) {
  const size_t vs = (size_t)c.gm_n * 4;
  if (c.right) {
    // right form: W_{k+1} = A T with T = P^-1 W_k from the previous sweep
    if (c.basis32)
      launch_v<float, float, 2>(c, c.d_AV32, c.d_T32, c.d_W32 + vs * (k + 1), nullptr, ev0,
                                ev1);
    else
      launch_v<double, double, 2>(c, c.d_AV, c.d_T, c.d_W + vs * (k + 1), nullptr, ev0, ev1);
    return hipGetLastError();
  }
  if (c.basis32)
    launch_v<float, float, 2>(c, c.d_ApV32, c.d_W32 + vs * k, c.d_W32 + vs * (k + 1), nullptr,
                              ev0, ev1);
  else if (precision == OFD_FP32)
    launch_v<float, double, 2>(c, c.d_ApV32, c.d_W + vs * k, c.d_W + vs * (k + 1), nullptr,
                               ev0, ev1);
  else
    launch_v<double, double, 2>(c, c.d_ApV, c.d_W + vs * k, c.d_W + vs * (k + 1), nullptr,
                                ev0, ev1);
  return hipGetLastError();
}

// Arnoldi product of the temperature system: W_{k+1} = (D^-1 A) W_k
static hipError_t launch_spmv_arnoldi_s(Ctx &c, int precision, int k, hipEvent_t ev0,
                                        hipEvent_t ev1) {
  const unsigned nb = spmv_s_blocks(c);
  const int n = c.nnode, np = c.gm_n * 4;
  const size_t vs = (size_t)c.gm_n * 4;
  if (c.basis32) {
    hipExtLaunchKernelGGL((k_spmv_s<float, float, 2>), dim3(nb), dim3(256), 0, c.stream, ev0,
                          ev1, 0, n, np, (const int32_t *)c.d_col, (const int32_t *)c.d_row,
                          (const float *)c.d_ApT32, (const float *)(c.d_W32 + vs * k),
                          c.d_W32 + vs * (k + 1), (const double *)nullptr,
                          (const GmresState *)c.d_gs);
  } else if (precision == OFD_FP32) {
    hipExtLaunchKernelGGL((k_spmv_s<float, double, 2>), dim3(nb), dim3(256), 0, c.stream,
                          ev0, ev1, 0, n, np, (const int32_t *)c.d_col,
                          (const int32_t *)c.d_row, (const float *)c.d_ApT32,
                          (const double *)(c.d_W + vs * k), c.d_W + vs * (k + 1),
                          (const double *)nullptr, (const GmresState *)c.d_gs);
  } else {
    hipExtLaunchKernelGGL((k_spmv_s<double, double, 2>), dim3(nb), dim3(256), 0, c.stream,
                          ev0, ev1, 0, n, np, (const int32_t *)c.d_col,
                          (const int32_t *)c.d_row, (const double *)c.d_ApT,
                          (const double *)(c.d_W + vs * k), c.d_W + vs * (k + 1),
                          (const double *)nullptr, (const GmresState *)c.d_gs);
  }
  return hipGetLastError();
}

// ev0/ev1 (may be null) receive the kernel's own start and end timestamps
hipError_t launch_spmv_arnoldi(Ctx &c, int precision, int k, hipEvent_t ev0,
                               hipEvent_t ev1) {
  if (c.gm_kind == 1) return launch_spmv_arnoldi_s(c, precision, k, ev0, ev1);
  if (c.gm_kind == 2) return launch_spmv_arnoldi_v(c, precision, k, ev0, ev1);
  const unsigned nb = spmv_blocks(c);
  const size_t vs = (size_t)c.nnode * 4;
  if (c.right) {
    // right form: W_{k+1} = A T with T = P^-1 W_k from the previous sweep
    if (c.basis32)
      hipExtLaunchKernelGGL((k_spmv<float, float, 2>), dim3(nb), dim3(256), 0, c.stream, ev0,
                            ev1, 0, c.nnode, (const int32_t *)c.d_col,
                            (const int32_t *)c.d_row, (const float *)c.d_A32,
                            (const float *)c.d_T32, c.d_W32 + vs * (k + 1),
                            (const double *)nullptr, (const GmresState *)c.d_gs);
    else
      hipExtLaunchKernelGGL((k_spmv<double, double, 2>), dim3(nb), dim3(256), 0, c.stream,
                            ev0, ev1, 0, c.nnode, (const int32_t *)c.d_col,
                            (const int32_t *)c.d_row, (const double *)c.d_A64,
                            (const double *)c.d_T, c.d_W + vs * (k + 1),
                            (const double *)nullptr, (const GmresState *)c.d_gs);
    return hipGetLastError();
  }
  if (c.basis32) {
    const float *x = c.d_W32 + vs * k;
    float *y = c.d_W32 + vs * (k + 1);
    hipExtLaunchKernelGGL((k_spmv<float, float, 2>), dim3(nb), dim3(256), 0, c.stream, ev0,
                          ev1, 0, c.nnode, (const int32_t *)c.d_col, (const int32_t *)c.d_row,
                          (const float *)c.d_Ap32, x, y, (const double *)nullptr,
                          (const GmresState *)c.d_gs);
    return hipGetLastError();
  }
  const double *x = c.d_W + vs * k;
  double *y = c.d_W + vs * (k + 1);
  if (precision == OFD_FP32)
    hipExtLaunchKernelGGL((k_spmv<float, double, 2>), dim3(nb), dim3(256), 0, c.stream, ev0,
                          ev1, 0, c.nnode, (const int32_t *)c.d_col, (const int32_t *)c.d_row,
                          (const float *)c.d_Ap32, x, y, (const double *)nullptr,
                          (const GmresState *)c.d_gs);
  else
    hipExtLaunchKernelGGL((k_spmv<double, double, 2>), dim3(nb), dim3(256), 0, c.stream, ev0,
                          ev1, 0, c.nnode, (const int32_t *)c.d_col, (const int32_t *)c.d_row,
                          (const double *)c.d_Ap64, x, y, (const double *)nullptr,
                          (const GmresState *)c.d_gs);
  return hipGetLastError();
}

}  // namespace ofd
