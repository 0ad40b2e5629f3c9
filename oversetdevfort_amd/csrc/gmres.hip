// Block-Jacobi GMRES building blocks on gfx950.
//
// Restates the Arnoldi/Givens loop of SparseGMRES_up
// (SparseGMRES_NS.f90:123-352) with every scalar of the cycle on the device:
// each sweep writes per-block partial sums, and a one-block finalize kernel
// adds them in a fixed order and writes the Hessenberg column, the Givens
// rotations, the residual estimate and the convergence flag, so a cycle runs
// without host round trips.  The kernel boundary orders the two (no
// device-scope fences, which on a multi-XCD part write back a whole L2).
// Results are bitwise reproducible run to run.
//
// Basis vectors are stored unscaled: v_j = s_j W_j with s_j = 1/|W_j| kept
// in GmresState, which saves the normalisation sweep of :293-294.
//
// Vectors are node-interleaved [N][4]; one thread per node.
#include "ofd_internal.h"

namespace ofd {

namespace {

constexpr int kCh = 16;  // dot products per reduction chunk

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int m = 32; m > 0; m >>= 1) v += __shfl_xor(v, m);
  return v;
}

__device__ __forceinline__ void ld4(const double *p, double v[4]) {
  const double2 a = reinterpret_cast<const double2 *>(p)[0];
  const double2 b = reinterpret_cast<const double2 *>(p)[1];
  v[0] = a.x;
  v[1] = a.y;
  v[2] = b.x;
  v[3] = b.y;
}

__device__ __forceinline__ void st4(double *p, const double v[4]) {
  reinterpret_cast<double2 *>(p)[0] = make_double2(v[0], v[1]);
  reinterpret_cast<double2 *>(p)[1] = make_double2(v[2], v[3]);
}

// Reduce acc[0..n) over the block into part[blockIdx][off + j].
template <int NV>
__device__ void block_partials(const double (&acc)[NV], int n, double *part, int stride,
                               int off) {
  __shared__ double s_red[4][kCh];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int j = 0; j < NV; ++j) {
    if (j < n) {
      const double v = wave_sum(acc[j]);
      if (lane == 0) s_red[w][j] = v;
    }
  }
  __syncthreads();
  if ((int)threadIdx.x < n) {
    const int j = threadIdx.x;
    part[(size_t)blockIdx.x * stride + off + j] =
        s_red[0][j] + s_red[1][j] + s_red[2][j] + s_red[3][j];
  }
  __syncthreads();
}


// Sum of part[b][j] over the nb producer blocks for j < m <= 65, by one
// 256-thread block: thread (g, j) adds blocks g, g+G, ... (independent
// loads that pipeline), then the G group sums are added in a fixed order,
// so the result is the same on every run.
__device__ void final_sums(const double *part, int stride, int nb, int m, double *s_out) {
  __shared__ double s_acc[kRedBlock];
  const int t = threadIdx.x;
  const int G = (m <= 32) ? 8 : ((m <= 64) ? 4 : 2);
  const int jw = kRedBlock / G;
  const int j = t % jw, g = t / jw;
  double s = 0.0;
  if (j < m) {
#pragma unroll 8
    for (int b = g; b < nb; b += G) s += part[(size_t)b * stride + j];
  }
  s_acc[t] = s;
  __syncthreads();
  if (t < m) {
    double r = 0.0;
    for (int q = 0; q < G; ++q) r += s_acc[q * jw + t];
    s_out[t] = r;
  }
  __syncthreads();
}

// Givens update of column k after H(0..k+1, k) is complete
// (SparseGMRES_NS.f90:305-327); single thread.
__device__ void givens_and_check(GmresState *g, int k) {
  const int K1 = g->kspace + 1;
  double *H = g->H;
  for (int j = 0; j < k; ++j) {
    const double t = g->rc[j] * H[j + K1 * k] + g->rs[j] * H[j + 1 + K1 * k];
    H[j + 1 + K1 * k] = -g->rs[j] * H[j + K1 * k] + g->rc[j] * H[j + 1 + K1 * k];
    H[j + K1 * k] = t;
  }
  const double hk = H[k + K1 * k], hk1 = H[k + 1 + K1 * k];
  const double t = sqrt(hk * hk + hk1 * hk1);
  g->rc[k] = hk / t;
  g->rs[k] = hk1 / t;
  H[k + K1 * k] = t;
  H[k + 1 + K1 * k] = 0.0;
  const double e0 = g->e[k], e1 = g->e[k + 1];
  g->e[k] = g->rc[k] * e0 + g->rs[k] * e1;
  g->e[k + 1] = -g->rs[k] * e0 + g->rc[k] * e1;
  const double er = fabs(g->e[k + 1]);
  g->resid = er;
  g->iters = k + 1;
  if ((er <= g->eps && k + 1 >= g->kmin) || !(er == er) || hk1 == 0.0) g->done = 1;
}

}  // namespace

// W_0 = P^-1 b and partial |W_0|^2 (SparseGMRES_NS.f90:124-155)
template <typename T>
__global__ void __launch_bounds__(256)
k_gmres_init(int N, const T *__restrict__ P, const double *__restrict__ b,
             double *__restrict__ W0, double *__restrict__ part) {
  double acc[1] = {0.0};
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N;
       i += gridDim.x * blockDim.x) {
    double r[4], z[4];
    ld4(b + (size_t)i * 4, r);
    const T *p = P + (size_t)i * 16;
#pragma unroll
    for (int c = 0; c < 4; ++c)
      z[c] = (double)p[c * 4 + 0] * r[0] + (double)p[c * 4 + 1] * r[1] +
             (double)p[c * 4 + 2] * r[2] + (double)p[c * 4 + 3] * r[3];
    st4(W0 + (size_t)i * 4, z);
    acc[0] += z[0] * z[0] + z[1] * z[1] + z[2] * z[2] + z[3] * z[3];
  }
  block_partials<1>(acc, 1, part, kMaxKspace + 1, 0);
}

// One MGS sweep j of Arnoldi step k (SparseGMRES_NS.f90:246-287):
//   j > 0:  z -= beta * v_{j-1};   partial of <z, v_j> (j <= k) or <z, z> (j = k+1)
__global__ void __launch_bounds__(256)
k_mgs_step(int N, double *__restrict__ W, double *__restrict__ part,
           const GmresState *__restrict__ g, int k, int j) {
  if (g->done) return;
  const size_t vs = (size_t)N * 4;
  double *z = W + vs * (k + 1);
  const double *vp = W + vs * (j > 0 ? j - 1 : 0);
  const double *vj = W + vs * (j <= k ? j : k + 1);
  const double cp = (j > 0) ? g->beta * g->s[j - 1] : 0.0;
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
    acc[0] += zz[0] * v[0] + zz[1] * v[1] + zz[2] * v[2] + zz[3] * v[3];
  }
  block_partials<1>(acc, 1, part, kMaxKspace + 1, 0);
}

// CGS2 sweeps of Arnoldi step k:
//   pass 1: partials of <z, W_j>                  j = 0..k
//   pass 2: z -= sum h1_j v_j ; partials of <z, W_j>
//   pass 3: z -= sum h2_j v_j ; partial of |z|^2
__global__ void __launch_bounds__(256)
k_cgs_pass(int N, double *__restrict__ W, double *__restrict__ part,
           const GmresState *__restrict__ g, int k, int pass) {
  if (g->done) return;
  const size_t vs = (size_t)N * 4;
  double *z = W + vs * (k + 1);
  const int m = k + 1;
  __shared__ double s_coef[kMaxKspace + 1];
  for (int j = threadIdx.x; j < m; j += blockDim.x)
    s_coef[j] = (pass == 2) ? g->h1[j] * g->s[j] : (pass == 3 ? g->h2[j] * g->s[j] : 0.0);
  __syncthreads();
  const int stride = kMaxKspace + 1;
  if (pass >= 2) {
    double acc[1] = {0.0};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N;
         i += gridDim.x * blockDim.x) {
      double zz[4];
      ld4(z + (size_t)i * 4, zz);
      for (int j = 0; j < m; ++j) {
        double v[4];
        ld4(W + vs * j + (size_t)i * 4, v);
        const double cj = s_coef[j];
#pragma unroll
        for (int c = 0; c < 4; ++c) zz[c] -= cj * v[c];
      }
      st4(z + (size_t)i * 4, zz);
      acc[0] += zz[0] * zz[0] + zz[1] * zz[1] + zz[2] * zz[2] + zz[3] * zz[3];
    }
    if (pass == 3) {
      block_partials<1>(acc, 1, part, stride, 0);
      return;
    }
  }
  // dots <z, W_j> in chunks of kCh; pass 2 reads the z the same thread wrote
  for (int j0 = 0; j0 < m; j0 += kCh) {
    const int n = (m - j0 < kCh) ? (m - j0) : kCh;
    double acc[kCh];
#pragma unroll
    for (int t = 0; t < kCh; ++t) acc[t] = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N;
         i += gridDim.x * blockDim.x) {
      double zz[4];
      ld4(z + (size_t)i * 4, zz);
#pragma unroll
      for (int t = 0; t < kCh; ++t) {
        if (t < n) {
          double v[4];
          ld4(W + vs * (j0 + t) + (size_t)i * 4, v);
          acc[t] += zz[0] * v[0] + zz[1] * v[1] + zz[2] * v[2] + zz[3] * v[3];
        }
      }
    }
    block_partials<kCh>(acc, n, part, stride, j0);
  }
}

// |v|^2 partials of a [N][4] vector
__global__ void __launch_bounds__(256)
k_norm2(int N, const double *__restrict__ v, double *__restrict__ part) {
  double acc[1] = {0.0};
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N;
       i += gridDim.x * blockDim.x) {
    double z[4];
    ld4(v + (size_t)i * 4, z);
    acc[0] += z[0] * z[0] + z[1] * z[1] + z[2] * z[2] + z[3] * z[3];
  }
  block_partials<1>(acc, 1, part, kMaxKspace + 1, 0);
}

// One block: finish the reduction of the preceding producer kernel and
// update the cycle scalars.  op: 0 init (e, s_0, eps), 1 MGS sweep j,
// 2 CGS pass 1, 3 CGS pass 2, 4 CGS pass 3, 5 norm into scal[slot].
__global__ void __launch_bounds__(256)
k_finalize(const double *__restrict__ part, int nb, GmresState *__restrict__ g, int op,
           int k, int j, double tol, int kspace, int kmin) {
  __shared__ double s_sum[kMaxKspace + 1];
  if (op >= 1 && op <= 4 && g->done) return;
  const int stride = kMaxKspace + 1;
  const int m = (op == 2 || op == 3) ? k + 1 : 1;
  final_sums(part, stride, nb, m, s_sum);
  const int t = threadIdx.x;
  if (op == 0) {
    if (t == 0) {
      const double nrm = sqrt(s_sum[0]);
      g->kspace = kspace;
      g->kmin = kmin;
      g->rnorm0 = nrm;
      g->eps = tol * nrm;
      g->e[0] = nrm;
      g->s[0] = (nrm > 0.0) ? 1.0 / nrm : 0.0;
      g->iters = 0;
      g->resid = nrm;
      g->beta = 0.0;
      g->done = (nrm > 0.0 && nrm == nrm) ? 0 : 1;
    }
    for (int q = t + 1; q <= kspace; q += blockDim.x) g->e[q] = 0.0;
  } else if (op == 1) {
    if (t == 0) {
      const int K1 = g->kspace + 1;
      if (j <= k) {
        const double beta = s_sum[0] * g->s[j];
        g->beta = beta;
        g->H[j + K1 * k] = beta;
      } else {
        const double beta = s_sum[0];
        g->beta = beta;
        const double nrm = sqrt(beta);
        g->H[k + 1 + K1 * k] = nrm;
        g->s[k + 1] = (nrm > 0.0) ? 1.0 / nrm : 0.0;
        givens_and_check(g, k);
      }
    }
  } else if (op == 2 || op == 3) {
    for (int q = t; q < m; q += blockDim.x) {
      const double h = s_sum[q] * g->s[q];
      if (op == 2)
        g->h1[q] = h;
      else
        g->h2[q] = h;
    }
  } else if (op == 4) {
    if (t == 0) {
      const int K1 = g->kspace + 1;
      for (int q = 0; q <= k; ++q) g->H[q + K1 * k] = g->h1[q] + g->h2[q];
      const double nrm = sqrt(s_sum[0]);
      g->H[k + 1 + K1 * k] = nrm;
      g->s[k + 1] = (nrm > 0.0) ? 1.0 / nrm : 0.0;
      givens_and_check(g, k);
    }
  } else {
    if (t == 0) g->scal[j] = sqrt(s_sum[0]);
  }
}

// back substitution (SparseGMRES_NS.f90:341-346) and x += sum y_j v_j (:349-352)
__global__ void __launch_bounds__(256)
k_gmres_finish(int N, const double *__restrict__ W, const GmresState *__restrict__ g,
               double *__restrict__ x) {
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
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int j = it - 1; j >= 0; --j) {
      const double yj = s_e[j] / s_H[j + it * j];
      s_y[j] = yj * g->s[j];
      for (int l = 0; l < j; ++l) s_e[l] -= yj * s_H[l + it * j];
    }
  }
  __syncthreads();
  const size_t vs = (size_t)N * 4;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N;
       i += gridDim.x * blockDim.x) {
    double xx[4];
    ld4(x + (size_t)i * 4, xx);
    for (int j = 0; j < it; ++j) {
      double v[4];
      ld4(W + vs * j + (size_t)i * 4, v);
#pragma unroll
      for (int c = 0; c < 4; ++c) xx[c] += s_y[j] * v[c];
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

static hipError_t fin(Ctx &c, int op, int k, int j, double tol = 0.0, int kspace = 0,
                      int kmin = 0) {
  hipLaunchKernelGGL(k_finalize, dim3(1), dim3(kRedBlock), 0, c.stream, c.d_part,
                     nblocks_nodes(c), c.d_gs, op, k, j, tol, kspace, kmin);
  return hipGetLastError();
}

hipError_t launch_gmres_init(Ctx &c, int precision, const double *b, double tol,
                             int kspace, int kmin) {
  const int nb = nblocks_nodes(c);
  if (precision == OFD_FP32)
    hipLaunchKernelGGL(k_gmres_init<float>, dim3(nb), dim3(kRedBlock), 0, c.stream,
                       c.nnode, c.d_P32, b, c.d_W, c.d_part);
  else
    hipLaunchKernelGGL(k_gmres_init<double>, dim3(nb), dim3(kRedBlock), 0, c.stream,
                       c.nnode, c.d_P64, b, c.d_W, c.d_part);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return fin(c, 0, 0, 0, tol, kspace, kmin);
}

hipError_t launch_mgs_step(Ctx &c, int k, int j) {
  hipLaunchKernelGGL(k_mgs_step, dim3(nblocks_nodes(c)), dim3(kRedBlock), 0, c.stream,
                     c.nnode, c.d_W, c.d_part, c.d_gs, k, j);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return fin(c, 1, k, j);
}

hipError_t launch_cgs_pass(Ctx &c, int k, int pass) {
  hipLaunchKernelGGL(k_cgs_pass, dim3(nblocks_nodes(c)), dim3(kRedBlock), 0, c.stream,
                     c.nnode, c.d_W, c.d_part, c.d_gs, k, pass);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return fin(c, 1 + pass, k, 0);
}

hipError_t launch_gmres_finish(Ctx &c, double *x) {
  hipLaunchKernelGGL(k_gmres_finish, dim3(nblocks_nodes(c)), dim3(kRedBlock), 0,
                     c.stream, c.nnode, c.d_W, c.d_gs, x);
  return hipGetLastError();
}

hipError_t launch_norm2(Ctx &c, const double *v, int slot) {
  hipLaunchKernelGGL(k_norm2, dim3(nblocks_nodes(c)), dim3(kRedBlock), 0, c.stream,
                     c.nnode, v, c.d_part);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return fin(c, 5, 0, slot);
}

}  // namespace ofd
