// Block-Jacobi GMRES building blocks on gfx950.
//
// Restates the Arnoldi/Givens loop of SparseGMRES_up
// (SparseGMRES_NS.f90:123-352) with every scalar of the cycle on the device.
// Each sweep writes per-block partial sums; the next kernel adds them in a
// fixed order (every block the same way) before its own sweep, and its
// block 0 writes the Hessenberg column, the Givens rotation, the residual
// estimate and the convergence flag.  The kernel boundary orders producer
// and consumer: no device-scope fences (on a multi-XCD part each one writes
// back a whole L2).  Results are bitwise reproducible run to run.
//
// Krylov loop on the stored left-preconditioned matrix P^-1 A:
//   product   z = (P^-1 A) W_k, stored unscaled as W_{k+1}
//   CGS       dots <z, W_j> (closing step k-1 first), update z -= sum h_j v_j
//             and |z|^2 (CGS1), or two projections (CGS2)
//   MGS       the k+1 sweeps of :246-287, for parity with the reference
// Basis vectors are stored unscaled: v_j = s_j W_j with s_j = 1/|W_j| kept in
// GmresState, which saves the normalisation sweep of :293-294.  The basis is
// FP64, or FP32 in the mixed solver (dot products, Hessenberg, Givens and the
// solution update always accumulate in FP64).
//
// Vectors are node-interleaved [N][4]; one thread per node.
#include "gmres_dev.h"

namespace ofd {

namespace {

__device__ __forceinline__ void ld4(const double *p, double v[4]) {
  const double2 a = reinterpret_cast<const double2 *>(p)[0];
  const double2 b = reinterpret_cast<const double2 *>(p)[1];
  v[0] = a.x;
  v[1] = a.y;
  v[2] = b.x;
  v[3] = b.y;
}

__device__ __forceinline__ void ld4(const float *p, double v[4]) {
  const float4 a = reinterpret_cast<const float4 *>(p)[0];
  v[0] = a.x;
  v[1] = a.y;
  v[2] = a.z;
  v[3] = a.w;
}

// store, leaving v as the stored (possibly rounded) values
__device__ __forceinline__ void st4(double *p, double v[4]) {
  reinterpret_cast<double2 *>(p)[0] = make_double2(v[0], v[1]);
  reinterpret_cast<double2 *>(p)[1] = make_double2(v[2], v[3]);
}

__device__ __forceinline__ void st4(float *p, double v[4]) {
  const float4 f = make_float4((float)v[0], (float)v[1], (float)v[2], (float)v[3]);
  reinterpret_cast<float4 *>(p)[0] = f;
  v[0] = f.x;
  v[1] = f.y;
  v[2] = f.z;
  v[3] = f.w;
}

__device__ __forceinline__ double dot4(const double a[4], const double b[4]) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2] + a[3] * b[3];
}

// four entries of a node as loaded (float4 or two double2), converted to
// FP64 only when used, so loads can be issued early without doubling the
// registers an FP32 basis needs
template <typename V>
struct Raw4;
template <>
struct Raw4<float> {
  float4 a;
  __device__ __forceinline__ void load(const float *p) {
    a = reinterpret_cast<const float4 *>(p)[0];
  }
  __device__ __forceinline__ void get(double v[4]) const {
    v[0] = a.x;
    v[1] = a.y;
    v[2] = a.z;
    v[3] = a.w;
  }
};
template <>
struct Raw4<double> {
  double2 a, b;
  __device__ __forceinline__ void load(const double *p) {
    a = reinterpret_cast<const double2 *>(p)[0];
    b = reinterpret_cast<const double2 *>(p)[1];
  }
  __device__ __forceinline__ void get(double v[4]) const {
    v[0] = a.x;
    v[1] = a.y;
    v[2] = b.x;
    v[3] = b.y;
  }
};

// Partials of <z, W_j>, j0 <= j < m, for the node of this thread, reduced
// over the block in chunks of kCh.  Every load of a chunk is issued
// unconditionally (indices past m clamped to m-1, their products dropped)
// so all of them are in flight together.
template <typename V>
__device__ __forceinline__ void dots_chunks(const V *__restrict__ W, size_t vs, int i,
                                            bool live, const double zz[4], int m,
                                            double *__restrict__ wr, int jstart = 0) {
  for (int j0 = jstart; j0 < m; j0 += kCh) {
    const int n = (m - j0 < kCh) ? (m - j0) : kCh;
    double v[kCh][4];
#pragma unroll
    for (int t = 0; t < kCh; ++t) {
      const int j = (j0 + t < m) ? j0 + t : m - 1;
      ld4(W + vs * j + (size_t)i * 4, v[t]);
    }
    double acc[kCh];
#pragma unroll
    for (int t = 0; t < kCh; ++t) acc[t] = (live && t < n) ? dot4(zz, v[t]) : 0.0;
    block_partials<kCh>(acc, n, wr, j0);
  }
}

// z -= sum_{j<m} coef[j] W_j for the node of this thread, 8 loads in flight
// per batch (indices past m clamped, their weights zero), subtracted in j order
template <typename V>
__device__ __forceinline__ void axpy_batches(const V *__restrict__ W, size_t vs, int i,
                                             const double *coef, int m, double zz[4],
                                             int jstart = 0) {
  for (int j0 = jstart; j0 < m; j0 += 8) {
    double v[8][4];
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      const int j = (j0 + t < m) ? j0 + t : m - 1;
      ld4(W + vs * j + (size_t)i * 4, v[t]);
    }
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      const double cj = (j0 + t < m) ? coef[j0 + t] : 0.0;
#pragma unroll
      for (int c = 0; c < 4; ++c) zz[c] -= cj * v[t][c];
    }
  }
}

}  // namespace

// z = P^-1 r for 4-tuple i of a system vector.
// KIND 0: NS node vectors, P = [N][16] block-Jacobi inverse;
// KIND 1: the T system, P = inverse diagonal of the nsc scalar unknowns;
// KIND 2: the NSVOF system, [N][4] (u, p) node values then [N] phi (nsc = N);
//         P = [N][16] block inverses of the 4x4 (u, p) diagonal blocks, then
//         [N] 1/LS.
// The scalar values of kind 1 and the phi values of kind 2 are taken 4 at a
// time; entries past the last node are 0.
template <int KIND>
__device__ __forceinline__ void apply_prec(int i, int nsc, const double *__restrict__ P,
                                           const double r[4], double z[4]) {
  if (KIND == 1) {
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int t = 4 * i + c;
      z[c] = (t < nsc) ? P[t] * r[c] : 0.0;
    }
  } else if (KIND == 2 && i >= nsc) {
    // phi entries: tuples nsc.. hold phi of nodes 4 (i - nsc) + c, scaled
    // by 1/LS (P + 16 nsc), zero past the last node
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int t = 4 * (i - nsc) + c;
      z[c] = (t < nsc) ? P[(size_t)16 * nsc + t] * r[c] : 0.0;
    }
  } else {
    const double *p = P + (size_t)i * 16;
#pragma unroll
    for (int c = 0; c < 4; ++c)
      z[c] = p[c * 4 + 0] * r[0] + p[c * 4 + 1] * r[1] + p[c * 4 + 2] * r[2] +
             p[c * 4 + 3] * r[3];
  }
}

// right form: T = P^-1 z for tuple i of the NSVOF system, z as stored in
// the basis (the input of the next Arnoldi product)
template <typename V>
__device__ __forceinline__ void store_right(int i, int nsc, const double *__restrict__ P,
                                            const double z[4], V *__restrict__ T) {
  double t[4];
  apply_prec<2>(i, nsc, P, z, t);
  st4(T + (size_t)i * 4, t);
}

// Left form: W_0 = P^-1 b and partials of |W_0|^2 (SparseGMRES_NS.f90:
// 124-155).  Right form (T given, NSVOF only): W_0 = b, T = P^-1 b, the
// input of the first Arnoldi product.  Block 0 also resets the cycle
// scalars.  The first orthogonalisation kernel of iteration 0 closes the
// norm.
template <typename V, int KIND>
__global__ void __launch_bounds__(256)
k_gmres_init(int N, int nsc, const double *__restrict__ P, const double *__restrict__ b,
             V *__restrict__ W0, V *__restrict__ T, double *__restrict__ part,
             GmresState *__restrict__ g, double tol, int kspace, int kmin) {
  if (g->outer_done) {
    // a cycle queued behind the one that met the outer tolerance
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      g->done = 1;
      g->iters = 0;
    }
    return;
  }
  if (blockIdx.x == 0) {
    for (int q = threadIdx.x; q <= kspace; q += blockDim.x) g->e[q] = 0.0;
    if (threadIdx.x == 0) {
      g->tol = tol;
      g->kspace = kspace;
      g->kmin = kmin;
      g->done = 0;
      g->iters = 0;
      g->beta = 0.0;
    }
  }
  double acc[1] = {0.0};
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N;
       i += gridDim.x * blockDim.x) {
    double r[4], z[4];
    ld4(b + (size_t)i * 4, r);
    if (T) {
      st4(W0 + (size_t)i * 4, r);
      apply_prec<KIND>(i, nsc, P, r, z);
      st4(T + (size_t)i * 4, z);
      acc[0] += dot4(r, r);
    } else {
      apply_prec<KIND>(i, nsc, P, r, z);
      st4(W0 + (size_t)i * 4, z);
      acc[0] += dot4(z, z);
    }
  }
  block_partials<1>(acc, 1, part, 0);
}

// One MGS sweep j of Arnoldi step k (SparseGMRES_NS.f90:246-287) on the
// unscaled product z = (P^-1 A) W_k (FP64 basis):
//   j = 0:  close step k-1 first (partials in `rd`)
//   j > 0:  beta = <z, v_{j-1}> from the previous sweep's partials (kept
//           in h1(j-1), h2(j-1) = 0); z -= beta v_{j-1}
//   partials of <z, W_j> (j <= k) or <z, z> (j = k+1) into `wr`
__global__ void __launch_bounds__(256)
k_mgs_step(int N, double *__restrict__ W, const double *__restrict__ rd,
           double *__restrict__ wr, GmresState *__restrict__ g, int k, int j, int nb,
           int32_t *hflag) {
  if (cycle_stopped(g, k, j == 0 ? hflag : nullptr)) return;
  __shared__ double s_b[1];
  const size_t vs = (size_t)N * 4;
  double *z = W + vs * (k + 1);
  double cp = 0.0;
  if (j == 0) {
    if (arnoldi_close(g, rd, nb, k - 1, hflag)) return;
  } else {
    final_sums(rd, nb, 1, s_b);
    const double beta = s_b[0] * g->s[j - 1];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      g->h1[j - 1] = beta;
      g->h2[j - 1] = 0.0;
    }
    cp = beta * g->s[j - 1];
  }
  const double *vp = W + vs * (j > 0 ? j - 1 : 0);
  const double *vj = W + vs * (j <= k ? j : k + 1);
  double acc[1] = {0.0};
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N;
       i += gridDim.x * blockDim.x) {
    double zz[4];
    ld4(z + (size_t)i * 4, zz);
    if (j > 0) {
      double v[4];
      ld4(vp + (size_t)i * 4, v);
#pragma unroll
      for (int c = 0; c < 4; ++c) zz[c] -= cp * v[c];
      st4(z + (size_t)i * 4, zz);
    }
    double v[4];
    if (j <= k) {
      ld4(vj + (size_t)i * 4, v);
    } else {
#pragma unroll
      for (int c = 0; c < 4; ++c) v[c] = zz[c];
    }
    acc[0] += dot4(zz, v);
  }
  block_partials<1>(acc, 1, wr, 0);
}

// First projection of the CGS variants: close step k-1 (norm partials in
// `rd`), then partials of <z, W_j>, j = 0..k, of the unscaled product
// z = (P^-1 A) W_k into `wr`.
template <typename V>
__global__ void __launch_bounds__(256)
k_cgs_dots(int N, const V *__restrict__ W, const double *__restrict__ rd,
           double *__restrict__ wr, GmresState *__restrict__ g, int k, int nb,
           int32_t *hflag) {
  if (cycle_stopped(g, k, hflag)) return;
  const size_t vs = (size_t)N * 4;
  const V *z = W + vs * (k + 1);
  const int m = k + 1;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = i < N;
  const int ii = live ? i : 0;
  // the first chunk's loads do not depend on the close: in flight during it
  Raw4<V> zr, vr[kCh];
  zr.load(z + (size_t)ii * 4);
#pragma unroll
  for (int t = 0; t < kCh; ++t) vr[t].load(W + vs * (t < m ? t : m - 1) + (size_t)ii * 4);
  if (arnoldi_close(g, rd, nb, k - 1, hflag)) return;
  double zz[4];
  zr.get(zz);
  {
    const int n = m < kCh ? m : kCh;
    double acc[kCh];
#pragma unroll
    for (int t = 0; t < kCh; ++t) {
      double v[4];
      vr[t].get(v);
      acc[t] = (live && t < n) ? dot4(zz, v) : 0.0;
    }
    block_partials<kCh>(acc, n, wr, 0);
  }
  if (m > kCh) dots_chunks(W, vs, ii, live, zz, m, wr, kCh);
}

// Update sweeps of the CGS variants after the projection dots of the
// unscaled product z = (P^-1 A) W_k against W_0..W_k (partials in `rd_dots`):
//   mode 1 (CGS1):        h1 = <z, v_j>, h2 = 0; z -= sum h1_j v_j;
//                         partials of |z|^2 -> wr
//   mode 2 (CGS2 first):  h1 = <z, v_j>; z -= sum h1_j v_j; partials of
//                         <z, W_j> -> wr
//   mode 3 (CGS2 second): h2 = <z, v_j>; z -= sum h2_j v_j; partials of |z|^2 -> wr
template <typename V>
__global__ void __launch_bounds__(256)
k_cgs_update(int N, V *__restrict__ W, const double *__restrict__ rd_dots,
             double *__restrict__ wr, GmresState *__restrict__ g, int k, int mode,
             int nb, const double *__restrict__ P, int nsc, V *__restrict__ T) {
  if (g->done) return;
  const size_t vs = (size_t)N * 4;
  V *z = W + vs * (k + 1);
  const int m = k + 1;
  __shared__ double s_coef[kMaxKspace + 1];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = i < N;
  const int ii = live ? i : 0;
  // z and the first 8 basis vectors do not depend on the coefficients:
  // their loads, the scales and the partial sums are all in flight together
  Raw4<V> zr, vr[8];
  zr.load(z + (size_t)ii * 4);
#pragma unroll
  for (int t = 0; t < 8; ++t) vr[t].load(W + vs * (t < m ? t : m - 1) + (size_t)ii * 4);
  const int t0 = threadIdx.x;
  const double s_own = (t0 < m) ? g->s[t0] : 0.0;
  final_sums(rd_dots, nb, m, s_coef);
  for (int j = threadIdx.x; j < m; j += blockDim.x) {
    const double sj = (j == t0) ? s_own : g->s[j];
    const double h = s_coef[j] * sj;
    if (blockIdx.x == 0) {
      if (mode == 3) {
        g->h2[j] = h;
      } else {
        g->h1[j] = h;
        if (mode == 1) g->h2[j] = 0.0;
      }
    }
    s_coef[j] = h * sj;
  }
  __syncthreads();
  double zz[4];
  zr.get(zz);
#pragma unroll
  for (int t = 0; t < 8; ++t) {
    double v[4];
    vr[t].get(v);
    const double cj = (t < m) ? s_coef[t] : 0.0;
#pragma unroll
    for (int c = 0; c < 4; ++c) zz[c] -= cj * v[c];
  }
  axpy_batches(W, vs, ii, s_coef, m, zz, 8);
  if (live) st4(z + (size_t)i * 4, zz);
  if (mode != 2) {
    if (T && live) store_right(i, nsc, P, zz, T);
    double acc[1] = {live ? dot4(zz, zz) : 0.0};
    block_partials<1>(acc, 1, wr, 0);
    return;
  }
  dots_chunks(W, vs, ii, live, zz, m, wr);
}

// k_cgs_dots and k_cgs_update (mode 1) for a step whose basis count
// m = k + 1 = M is a compile-time constant: the basis loads of the sweep are
// issued unconditionally and with no clamped repeats, all of them at kernel
// entry up to kFlight<V> (32 FP32 nodes of 4 VGPRs; 16 FP64 nodes of 8,
// beyond which the registers run out), the rest in exact batches after.
// The block partials and the subtraction order are those of the runtime-m
// kernels; with the FP32 basis the results are the same bits.
template <typename V>
constexpr int kFlight = sizeof(V) == 4 ? 32 : 16;

// partials of <z, W_j>, J0 <= j < M, chunks of kCh, loads of W_j issued per
// chunk (exact counts)
template <typename V, int M, int J0>
__device__ __forceinline__ void dots_tail(const V *__restrict__ W, size_t vs, int ii,
                                          bool live, const double zz[4],
                                          double *__restrict__ wr) {
#pragma unroll
  for (int j0 = J0; j0 < M; j0 += kCh) {
    const int n = (M - j0 < kCh) ? (M - j0) : kCh;
    Raw4<V> vr[kCh];
#pragma unroll
    for (int t = 0; t < kCh; ++t)
      if (t < n) vr[t].load(W + vs * (j0 + t) + (size_t)ii * 4);
    double acc[kCh];
#pragma unroll
    for (int t = 0; t < kCh; ++t) {
      double v[4] = {0.0, 0.0, 0.0, 0.0};
      if (t < n) vr[t].get(v);
      acc[t] = (live && t < n) ? dot4(zz, v) : 0.0;
    }
    block_partials<kCh>(acc, n, wr, j0);
  }
}

template <typename V, int M>
__global__ void __launch_bounds__(256)
k_cgs_dots_m(int N, const V *__restrict__ W, const double *__restrict__ rd,
             double *__restrict__ wr, GmresState *__restrict__ g, int nb,
             int32_t *hflag) {
  constexpr int k = M - 1;
  constexpr int M1 = M < kFlight<V> ? M : kFlight<V>;
  if (cycle_stopped(g, k, hflag)) return;
  const size_t vs = (size_t)N * 4;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = i < N;
  const int ii = live ? i : 0;
  Raw4<V> zr, vr[M1];
  zr.load(W + vs * M + (size_t)ii * 4);
#pragma unroll
  for (int t = 0; t < M1; ++t) vr[t].load(W + vs * t + (size_t)ii * 4);
  if (arnoldi_close(g, rd, nb, k - 1, hflag)) return;
  double zz[4];
  zr.get(zz);
#pragma unroll
  for (int j0 = 0; j0 < M1; j0 += kCh) {
    const int n = (M1 - j0 < kCh) ? (M1 - j0) : kCh;
    double acc[kCh];
#pragma unroll
    for (int t = 0; t < kCh; ++t) {
      const int j = (j0 + t < M1) ? j0 + t : M1 - 1;
      double v[4];
      vr[j].get(v);
      acc[t] = (live && t < n) ? dot4(zz, v) : 0.0;
    }
    block_partials<kCh>(acc, n, wr, j0);
  }
  if constexpr (M > M1) dots_tail<V, M, M1>(W, vs, ii, live, zz, wr);
}

template <typename V, int M>
__global__ void __launch_bounds__(256)
k_cgs_update1_m(int N, V *__restrict__ W, const double *__restrict__ rd_dots,
                double *__restrict__ wr, GmresState *__restrict__ g, int nb,
                const double *__restrict__ P, int nsc, V *__restrict__ T) {
  constexpr int M1 = M < kFlight<V> ? M : kFlight<V>;
  if (g->done) return;
  const size_t vs = (size_t)N * 4;
  V *z = W + vs * M;
  __shared__ double s_coef[M];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = i < N;
  const int ii = live ? i : 0;
  Raw4<V> zr, vr[M1];
  zr.load(z + (size_t)ii * 4);
#pragma unroll
  for (int t = 0; t < M1; ++t) vr[t].load(W + vs * t + (size_t)ii * 4);
  const int t0 = threadIdx.x;
  const double s_own = (t0 < M) ? g->s[t0] : 0.0;
  final_sums(rd_dots, nb, M, s_coef);
  if (t0 < M) {
    const double h = s_coef[t0] * s_own;
    if (blockIdx.x == 0) {
      g->h1[t0] = h;
      g->h2[t0] = 0.0;
    }
    s_coef[t0] = h * s_own;
  }
  __syncthreads();
  double zz[4];
  zr.get(zz);
#pragma unroll
  for (int t = 0; t < M1; ++t) {
    double v[4];
    vr[t].get(v);
    const double cj = s_coef[t];
#pragma unroll
    for (int c = 0; c < 4; ++c) zz[c] -= cj * v[c];
  }
  if constexpr (M > M1) {
#pragma unroll
    for (int j0 = M1; j0 < M; j0 += kCh) {
      const int n = (M - j0 < kCh) ? (M - j0) : kCh;
      Raw4<V> vb[kCh];
#pragma unroll
      for (int t = 0; t < kCh; ++t)
        if (t < n) vb[t].load(W + vs * (j0 + t) + (size_t)ii * 4);
#pragma unroll
      for (int t = 0; t < kCh; ++t) {
        if (t < n) {
          double v[4];
          vb[t].get(v);
          const double cj = s_coef[j0 + t];
#pragma unroll
          for (int c = 0; c < 4; ++c) zz[c] -= cj * v[c];
        }
      }
    }
  }
  if (live) st4(z + (size_t)i * 4, zz);
  if (T && live) store_right(i, nsc, P, zz, T);
  double acc[1] = {live ? dot4(zz, zz) : 0.0};
  block_partials<1>(acc, 1, wr, 0);
}

// closes the last Arnoldi step of a cycle (one block)
__global__ void __launch_bounds__(256)
k_gmres_close(const double *__restrict__ part, int nb, GmresState *__restrict__ g, int k) {
  if (g->done) return;
  arnoldi_close(g, part, nb, k, nullptr);
}

// |v|^2 partials of a [N][4] vector
__global__ void __launch_bounds__(256)
k_norm2(int N, const double *__restrict__ v, double *__restrict__ part) {
  double acc[1] = {0.0};
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N;
       i += gridDim.x * blockDim.x) {
    double z[4];
    ld4(v + (size_t)i * 4, z);
    acc[0] += dot4(z, z);
  }
  block_partials<1>(acc, 1, part, 0);
}

// rec (may be null): the cycle's record for the host, `seq` stored last
// with a system-scope release so the fields are visible before it
__global__ void __launch_bounds__(256)
k_norm_fin(const double *__restrict__ part, int nb, GmresState *__restrict__ g, int slot,
           int outer, CycleRecord *rec, int32_t seq) {
  __shared__ double s_sum[1];
  final_sums(part, nb, 1, s_sum);
  if (threadIdx.x == 0) {
    const double rn = sqrt(s_sum[0]);
    g->scal[slot] = rn;
    // outer residual of a restarted solve: at or below the threshold (or
    // not finite) stops the cycles queued behind this one
    if (outer) g->outer_done = (rn > g->outer_thr) ? 0 : 1;
    if (rec) {
      rec->resid = g->resid;
      rec->rnorm0 = g->rnorm0;
      rec->rnorm = rn;
      rec->iters = g->iters;
      __hip_atomic_store(&rec->seq, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    }
  }
}

// |u|^2 and |p|^2 partials of a [N][4] vector (velocity and pressure parts)
__global__ void __launch_bounds__(256)
k_norm2_up(int N, const double *__restrict__ v, double *__restrict__ part) {
  double acc[2] = {0.0, 0.0};
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N;
       i += gridDim.x * blockDim.x) {
    double z[4];
    ld4(v + (size_t)i * 4, z);
    acc[0] += z[0] * z[0] + z[1] * z[1] + z[2] * z[2];
    acc[1] += z[3] * z[3];
  }
  block_partials<2>(acc, 2, part, 0);
}

// |u|^2, |p|^2, |phi|^2 partials of an NSVOF vector ([N][4] then [N] phi)
__global__ void __launch_bounds__(256)
k_norm2_v(int N, const double *__restrict__ v, double *__restrict__ part) {
  double acc[3] = {0.0, 0.0, 0.0};
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N;
       i += gridDim.x * blockDim.x) {
    const double *z = v + (size_t)i * 4;
    const double f = v[(size_t)4 * N + i];
    acc[0] += z[0] * z[0] + z[1] * z[1] + z[2] * z[2];
    acc[1] += z[3] * z[3];
    acc[2] += f * f;
  }
  block_partials<3>(acc, 3, part, 0);
}

// scal[slot + j] = sqrt of summed partial column j < m
__global__ void __launch_bounds__(256)
k_norm_fin2(const double *__restrict__ part, int nb, GmresState *__restrict__ g, int slot,
            int m) {
  __shared__ double s_sum[3];
  final_sums(part, nb, m, s_sum);
  if ((int)threadIdx.x < m) g->scal[slot + threadIdx.x] = sqrt(s_sum[threadIdx.x]);
}

__global__ void k_outer_reset(GmresState *__restrict__ g, double thr) {
  if (threadIdx.x == 0) {
    g->outer_thr = thr;
    g->outer_done = 0;
  }
}

// back substitution (SparseGMRES_NS.f90:341-346) and x += sum y_j v_j (:349-352)
// Right form (P given, NSVOF only): x += P^-1 (sum y_j v_j).
template <typename V>
__global__ void __launch_bounds__(256)
k_gmres_finish(int N, const V *__restrict__ W, const GmresState *__restrict__ g,
               double *__restrict__ x, const double *__restrict__ P, int nsc) {
  if (g->outer_done) return;  // queued behind the cycle that met the tolerance
  __shared__ double s_y[kMaxKspace];
  __shared__ double s_H[kMaxKspace * kMaxKspace];  // upper triangle, column-major it x it
  __shared__ double s_e[kMaxKspace + 1];
  const int it = g->iters;
  const int K1 = g->kspace + 1;
  for (int t = threadIdx.x; t < it * it; t += blockDim.x) {
    const int r = t % it, cidx = t / it;
    s_H[t] = g->H[r + K1 * cidx];
  }
  for (int t = threadIdx.x; t <= it; t += blockDim.x) s_e[t] = g->e[t];
  // column-oriented back substitution: y_j once e_j is final, then every
  // thread l < j updates its own e_l (the same subtraction order per entry
  // as the serial loop)
  const int t = threadIdx.x;
  const double s_t = (t < it) ? g->s[t] : 0.0;
  for (int j = it - 1; j >= 0; --j) {
    __syncthreads();
    const double yj = s_e[j] / s_H[j + it * j];
    if (t < j) s_e[t] -= yj * s_H[t + it * j];
    if (t == j) s_y[j] = yj * s_t;
  }
  __syncthreads();
  const size_t vs = (size_t)N * 4;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N;
       i += gridDim.x * blockDim.x) {
    double xx[4], x0[4];
    ld4(x + (size_t)i * 4, x0);
    // the left form adds the terms to x itself; the right form sums them
    // from zero and adds P^-1 of the sum
#pragma unroll
    for (int c = 0; c < 4; ++c) xx[c] = P ? 0.0 : x0[c];
    int j = 0;
    for (; j + 4 <= it; j += 4) {
      double v0[4], v1[4], v2[4], v3[4];
      ld4(W + vs * (j + 0) + (size_t)i * 4, v0);
      ld4(W + vs * (j + 1) + (size_t)i * 4, v1);
      ld4(W + vs * (j + 2) + (size_t)i * 4, v2);
      ld4(W + vs * (j + 3) + (size_t)i * 4, v3);
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        xx[c] += s_y[j] * v0[c];
        xx[c] += s_y[j + 1] * v1[c];
        xx[c] += s_y[j + 2] * v2[c];
        xx[c] += s_y[j + 3] * v3[c];
      }
    }
    for (; j < it; ++j) {
      double v[4];
      ld4(W + vs * j + (size_t)i * 4, v);
#pragma unroll
      for (int c = 0; c < 4; ++c) xx[c] += s_y[j] * v[c];
    }
    if (P) {
      double d[4];
      apply_prec<2>(i, nsc, P, xx, d);
#pragma unroll
      for (int c = 0; c < 4; ++c) xx[c] = x0[c] + d[c];
    }
    st4(x + (size_t)i * 4, xx);
  }
}

int nblocks_nodes(const Ctx &c) {
  int nb = (c.nnode + kRedBlock - 1) / kRedBlock;
  if (nb > kMaxRedBlocks) nb = kMaxRedBlocks;
  if (nb < 1) nb = 1;
  return nb;
}

// blocks of the GMRES sweeps over the current system's gm_n tuples
int nblocks_gm(const Ctx &c) {
  int nb = (c.gm_n + kRedBlock - 1) / kRedBlock;
  if (nb > kMaxRedBlocks) nb = kMaxRedBlocks;
  if (nb < 1) nb = 1;
  return nb;
}

// partial-sum buffers: 0/1 alternate along the MGS chain and hold the
// closing norms of the CGS variants (parity of k); 2 holds the projection
// dots, 3 the dots of the second CGS2 projection
static double *part_buf(Ctx &c, int which) {
  return c.d_part + (size_t)kMaxRedBlocks * kPartStride * which;
}

hipError_t launch_gmres_init(Ctx &c, const double *b, double tol, int kspace, int kmin) {
  const int nb = nblocks_gm(c);
  const int kind = c.gm_kind;
  const double *P = kind == 1 ? c.d_PT : (kind == 2 ? c.d_PV : c.d_P64);
  const int nsc = c.nnode;
#define OFD_INIT(V, K, W, T)                                                                \
  hipLaunchKernelGGL((k_gmres_init<V, K>), dim3(nb), dim3(kRedBlock), 0, c.stream, c.gm_n, \
                     nsc, P, b, W, T, part_buf(c, 0), c.d_gs, tol, kspace, kmin)
  if (c.basis32) {
    if (kind == 1)
      OFD_INIT(float, 1, c.d_W32, (float *)nullptr);
    else if (kind == 2)
      OFD_INIT(float, 2, c.d_W32, c.right ? c.d_T32 : nullptr);
    else
      OFD_INIT(float, 0, c.d_W32, c.right ? c.d_T32 : nullptr);
  } else {
    if (kind == 1)
      OFD_INIT(double, 1, c.d_W, (double *)nullptr);
    else if (kind == 2)
      OFD_INIT(double, 2, c.d_W, c.right ? c.d_T : nullptr);
    else
      OFD_INIT(double, 0, c.d_W, c.right ? c.d_T : nullptr);
  }
#undef OFD_INIT
  c.part_cur = 0;
  return hipGetLastError();
}

hipError_t launch_mgs_step(Ctx &c, int k, int j) {
  const int nb = nblocks_gm(c);
  hipLaunchKernelGGL(k_mgs_step, dim3(nb), dim3(kRedBlock), 0, c.stream, c.gm_n, c.d_W,
                     part_buf(c, c.part_cur), part_buf(c, c.part_cur ^ 1), c.d_gs, k, j, nb,
                     c.cur_flags);
  c.part_cur ^= 1;
  return hipGetLastError();
}

// basis counts m = k + 1 up to kExactM take the exact-count kernels
constexpr int kExactM = 32;

#define OFD_M8(X, a) X(a + 1) X(a + 2) X(a + 3) X(a + 4) X(a + 5) X(a + 6) X(a + 7) X(a + 8)
#define OFD_M32(X) OFD_M8(X, 0) OFD_M8(X, 8) OFD_M8(X, 16) OFD_M8(X, 24)

template <typename V>
static void launch_dots_m(Ctx &c, int m, int nb, const V *W, const double *rd, double *wr) {
  switch (m) {
#define OFD_DM(M)                                                                            \
  case M:                                                                                  \
    hipLaunchKernelGGL((k_cgs_dots_m<V, M>), dim3(nb), dim3(kRedBlock), 0, c.stream, c.gm_n, \
                       W, rd, wr, c.d_gs, nb, c.cur_flags);                                  \
    break;
    OFD_M32(OFD_DM)
#undef OFD_DM
  }
}

template <typename V>
static void launch_update1_m(Ctx &c, int m, int nb, V *W, const double *rd, double *wr,
                             V *T) {
  const double *P = T ? (c.gm_kind == 2 ? c.d_PV : c.d_P64) : nullptr;
  switch (m) {
#define OFD_UM(M)                                                                           \
  case M:                                                                                 \
    hipLaunchKernelGGL((k_cgs_update1_m<V, M>), dim3(nb), dim3(kRedBlock), 0, c.stream,     \
                       c.gm_n, W, rd, wr, c.d_gs, nb, P, c.nnode, T);                       \
    break;
    OFD_M32(OFD_UM)
#undef OFD_UM
  }
}

hipError_t launch_cgs_dots(Ctx &c, int k) {
  const int nb = nblocks_gm(c);
  if (k + 1 <= kExactM) {
    if (c.basis32)
      launch_dots_m<float>(c, k + 1, nb, c.d_W32, part_buf(c, k & 1), part_buf(c, 2));
    else
      launch_dots_m<double>(c, k + 1, nb, c.d_W, part_buf(c, k & 1), part_buf(c, 2));
    return hipGetLastError();
  }
  if (c.basis32)
    hipLaunchKernelGGL(k_cgs_dots<float>, dim3(nb), dim3(kRedBlock), 0, c.stream, c.gm_n,
                       c.d_W32, part_buf(c, k & 1), part_buf(c, 2), c.d_gs, k, nb,
                       c.cur_flags);
  else
    hipLaunchKernelGGL(k_cgs_dots<double>, dim3(nb), dim3(kRedBlock), 0, c.stream,
                       c.gm_n, c.d_W, part_buf(c, k & 1), part_buf(c, 2), c.d_gs, k, nb,
                       c.cur_flags);
  return hipGetLastError();
}

hipError_t launch_cgs_update(Ctx &c, int k, int mode) {
  const int nb = nblocks_gm(c);
  const double *rd_dots = part_buf(c, mode == 3 ? 3 : 2);
  double *wr = (mode == 2) ? part_buf(c, 3) : part_buf(c, (k + 1) & 1);
  // right form: the sweep that leaves W_{k+1} also writes T = P^-1 W_{k+1}
  const double *P = c.right ? (c.gm_kind == 2 ? c.d_PV : c.d_P64) : nullptr;
  float *T32 = c.right ? c.d_T32 : nullptr;
  double *T64 = c.right ? c.d_T : nullptr;
  if (mode == 1 && k + 1 <= kExactM) {
    if (c.basis32)
      launch_update1_m<float>(c, k + 1, nb, c.d_W32, rd_dots, wr, T32);
    else
      launch_update1_m<double>(c, k + 1, nb, c.d_W, rd_dots, wr, T64);
  } else if (c.basis32)
    hipLaunchKernelGGL(k_cgs_update<float>, dim3(nb), dim3(kRedBlock), 0, c.stream,
                       c.gm_n, c.d_W32, rd_dots, wr, c.d_gs, k, mode, nb, P, c.nnode, T32);
  else
    hipLaunchKernelGGL(k_cgs_update<double>, dim3(nb), dim3(kRedBlock), 0, c.stream,
                       c.gm_n, c.d_W, rd_dots, wr, c.d_gs, k, mode, nb, P, c.nnode, T64);
  c.part_cur = (k + 1) & 1;
  return hipGetLastError();
}

hipError_t launch_gmres_close(Ctx &c, int k, int ortho) {
  (void)ortho;
  hipLaunchKernelGGL(k_gmres_close, dim3(1), dim3(kRedBlock), 0, c.stream,
                     part_buf(c, c.part_cur), nblocks_gm(c), c.d_gs, k);
  return hipGetLastError();
}

hipError_t launch_gmres_finish(Ctx &c, double *x) {
  const double *P = c.right ? (c.gm_kind == 2 ? c.d_PV : c.d_P64) : nullptr;
  if (c.basis32)
    hipLaunchKernelGGL(k_gmres_finish<float>, dim3(nblocks_gm(c)), dim3(kRedBlock), 0,
                       c.stream, c.gm_n, c.d_W32, c.d_gs, x, P, c.nnode);
  else
    hipLaunchKernelGGL(k_gmres_finish<double>, dim3(nblocks_gm(c)), dim3(kRedBlock), 0,
                       c.stream, c.gm_n, c.d_W, c.d_gs, x, P, c.nnode);
  return hipGetLastError();
}

hipError_t launch_norm2_up(Ctx &c, const double *v, int slot) {
  const int nb = nblocks_nodes(c);
  double *pb = part_buf(c, 0);
  hipLaunchKernelGGL(k_norm2_up, dim3(nb), dim3(kRedBlock), 0, c.stream, c.nnode, v, pb);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_norm_fin2, dim3(1), dim3(kRedBlock), 0, c.stream, pb, nb, c.d_gs,
                     slot, 2);
  return hipGetLastError();
}

// |u|, |p|, |phi| of an NSVOF vector into GmresState::scal[1..3]
hipError_t launch_norm2_v(Ctx &c, const double *v) {
  const int nb = nblocks_nodes(c);
  double *pb = part_buf(c, 0);
  hipLaunchKernelGGL(k_norm2_v, dim3(nb), dim3(kRedBlock), 0, c.stream, c.nnode, v, pb);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_norm_fin2, dim3(1), dim3(kRedBlock), 0, c.stream, pb, nb, c.d_gs, 1,
                     3);
  return hipGetLastError();
}

hipError_t launch_outer_reset(Ctx &c, double thr) {
  hipLaunchKernelGGL(k_outer_reset, dim3(1), dim3(64), 0, c.stream, c.d_gs, thr);
  return hipGetLastError();
}

hipError_t launch_norm2(Ctx &c, const double *v, int slot, int outer, CycleRecord *rec,
                        int32_t seq) {
  const int nb = nblocks_gm(c);
  double *pb = part_buf(c, 0);
  hipLaunchKernelGGL(k_norm2, dim3(nb), dim3(kRedBlock), 0, c.stream, c.gm_n, v, pb);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_norm_fin, dim3(1), dim3(kRedBlock), 0, c.stream, pb, nb, c.d_gs,
                     slot, outer, rec, seq);
  return hipGetLastError();
}

}  // namespace ofd
