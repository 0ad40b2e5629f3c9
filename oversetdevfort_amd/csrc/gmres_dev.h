// Device helpers shared by the SpMV and GMRES kernels: fixed-order sums of
// per-block partials and the closing of an Arnoldi step (Hessenberg column,
// Givens rotation, residual estimate, convergence test).
//
// Every block of a consumer kernel repeats these few hundred flops on the
// partials its producer kernel left behind, in the same order, so all
// blocks agree bit for bit without any device-scope fence; block 0 alone
// writes the results back to GmresState.
#pragma once

#include "ofd_internal.h"

namespace ofd {

constexpr int kPartStride = kMaxKspace + 1;
constexpr int kCh = 16;  // dot products per reduction chunk

// Reduce acc[0..n) over a 256-thread block into part[blockIdx][off + j]
// through LDS: thread t stores its values at s[j][t]; thread (j, q) adds
// s[j][q + 16 r], r = 0..15 (consecutive q -> consecutive banks); thread j
// adds its 16 group sums.  A fixed order, no cross-lane shuffles.
template <int NV>
__device__ __forceinline__ void block_partials(const double (&acc)[NV], int n, double *part,
                                               int off) {
  __shared__ double s_v[NV][kRedBlock];
  __shared__ double s_g[NV][16];
  const int t = threadIdx.x;
#pragma unroll
  for (int j = 0; j < NV; ++j) s_v[j][t] = acc[j];
  __syncthreads();
  {
    const int j = t >> 4, q = t & 15;
    if (j < NV && j < n) {
      double r = 0.0;
#pragma unroll
      for (int u = 0; u < 16; ++u) r += s_v[j][q + 16 * u];
      s_g[j][q] = r;
    }
  }
  __syncthreads();
  if (t < n) {
    double r = 0.0;
#pragma unroll
    for (int q = 0; q < 16; ++q) r += s_g[t][q];
    part[(size_t)blockIdx.x * kPartStride + off + t] = r;
  }
  __syncthreads();
}

// s_out[j] = sum over b < nb of part[b][j], j < m <= 65 (all block threads).
// Thread (g, j) adds blocks g, g+G, ... with G = 256 / jw groups (jw = m
// rounded up to a power of two), then the G group sums of each j are added
// by a fixed-shape tree: the result is the same on every run and in every
// block.
__device__ __forceinline__ void final_sums(const double *part, int nb, int m,
                                           double *s_out) {
  __shared__ double s_acc[kRedBlock];
  const int t = threadIdx.x;
  int jw = 1;
  while (jw < m) jw <<= 1;
  const int G = kRedBlock / jw;  // 0 if m > 256 (not used: m <= 65)
  const int j = t % jw, g = t / jw;
  double s = 0.0;
  if (j < m) {
    // independent loads in batches of 16, added in block order
    int b = g;
    for (; b + 15 * G < nb; b += 16 * G) {
      double v[16];
#pragma unroll
      for (int u = 0; u < 16; ++u) v[u] = part[(size_t)(b + u * G) * kPartStride + j];
#pragma unroll
      for (int u = 0; u < 16; ++u) s += v[u];
    }
    for (; b < nb; b += G) s += part[(size_t)b * kPartStride + j];
  }
  s_acc[t] = s;
  __syncthreads();
  for (int h = G >> 1; h > 0; h >>= 1) {
    if (g < h) s_acc[t] += s_acc[t + h * jw];
    __syncthreads();
  }
  if (t < m) s_out[t] = s_acc[t];
  __syncthreads();
}

// Sum over b < nb of part[b][0], in every block thread: thread t adds
// blocks t, t+256, ..., then the 256 values are added through LDS in a
// fixed order -- every block gets the same bits.
__device__ __forceinline__ double sum1(const double *part, int nb) {
  __shared__ double s_v[kRedBlock];
  __shared__ double s_g[16];
  const int t = threadIdx.x;
  double s = 0.0;
  for (int b = t; b < nb; b += kRedBlock) s += part[(size_t)b * kPartStride];
  s_v[t] = s;
  __syncthreads();
  if (t < 16) {
    double r = 0.0;
#pragma unroll
    for (int u = 0; u < 16; ++u) r += s_v[t + 16 * u];
    s_g[t] = r;
  }
  __syncthreads();
  double r = 0.0;
#pragma unroll
  for (int q = 0; q < 16; ++q) r += s_g[q];
  return r;
}

// Entry test of a step kernel: true if the cycle has already stopped (an
// earlier close converged, or the cycle was queued behind one that met the
// outer tolerance).  The host waits on the flag of step k - 1 from the
// kernel that closes it, so block 0 of a kernel that returns here reports
// the stop in that flag: every flag the host waits for gets written.
__device__ __forceinline__ bool cycle_stopped(const GmresState *__restrict__ g, int k,
                                              int32_t *hflag) {
  if (!g->done) return false;
  if (hflag && k >= 1 && blockIdx.x == 0 && threadIdx.x == 0) hflag[k - 1] = 1;
  return true;
}

// Close Arnoldi step k (SparseGMRES_NS.f90:289-327).  The product of step k
// is stored unscaled, z = (P^-1 A) W_k = (P^-1 A v_k) / s_k, so every
// coefficient found for z is multiplied by s_k here: H(j,k) = s_k (h1_j + h2_j)
// for j <= k (the CGS projections, or the MGS betas in h1 with h2 = 0) and
// H(k+1,k) = s_k |z_perp| from the partials of |z_perp|^2.  Then the previous
// rotations are applied to column k, the new one is formed, e is rotated and
// |e_{k+1}| <= eps and k+1 >= kmin is tested.  k = -1 closes the initial
// residual instead (e_0 = |P^-1 b|, eps).  Returns the convergence flag;
// *s_next (if given) receives s_{k+1} in every block.
//
// Every block reads only what earlier kernels wrote: h1, h2, s_k, the
// rotations q < k and ec[k] (e_k before this rotation; e_{k+1} is still the
// zero the cycle started with).  Block 0 writes H(:,k), rc/rs[k], e[k],
// e[k+1], ec[k+1] and s[k+1], none of which another block of the kernel
// reads, so a block that starts after block 0 has finished (a grid larger
// than one round of resident blocks) takes the same decision.
__device__ __forceinline__ bool arnoldi_close(GmresState *__restrict__ g,
                                              const double *__restrict__ part, int nb,
                                              int k, int32_t *hflag,
                                              double *s_next = nullptr) {
  __shared__ double s_res[2];
  __shared__ double s_h[kMaxKspace + 1];   // s_k h_q, q <= k
  __shared__ double s_col[kMaxKspace + 2]; // rotated column
  __shared__ double s_rc[kMaxKspace], s_rs[kMaxKspace];
  const int t = threadIdx.x;
  // issue every global load of the close before the first barrier: the
  // column entries, rotations and s_k (threads q <= k), the cycle scalars
  // (thread 0) and the norm partials (all threads, inside sum1)
  const int K1 = g->kspace + 1;
  double colq = 0.0, rcq = 0.0, rsq = 0.0, sk = 0.0;
  if (k >= 0 && t <= k) {
    colq = g->h1[t] + g->h2[t];
    sk = g->s[k];
    if (t < k) {
      rcq = g->rc[t];
      rsq = g->rs[t];
    }
  }
  double e0 = 0.0, e1 = 0.0, eps = 0.0, tol = 0.0;
  int kmin = 0;
  if (t == 0) {
    if (k >= 0) {
      e0 = g->ec[k];
      eps = g->eps;
      kmin = g->kmin;
    } else {
      tol = g->tol;
    }
  }
  const double nrm = sqrt(sum1(part, nb));
  if (k >= 0 && t <= k) {
    s_h[t] = colq * sk;
    if (t < k) {
      s_rc[t] = rcq;
      s_rs[t] = rsq;
    }
  }
  __syncthreads();
  if (t == 0) {
    const double sc = (nrm > 0.0) ? 1.0 / nrm : 0.0;
    bool done;
    const bool w = blockIdx.x == 0;
    if (k < 0) {
      done = !(nrm > 0.0 && nrm == nrm);
      if (w) {
        g->rnorm0 = nrm;
        g->eps = tol * nrm;
        g->e[0] = nrm;
        g->ec[0] = nrm;
        g->s[0] = sc;
        g->resid = nrm;
        g->iters = 0;
        g->done = done ? 1 : 0;
      }
    } else {
      // previous rotations applied in order; entry q of the running column
      // is carried in a register, the loads of h, c, s do not depend on it
      double carry = s_h[0];
#pragma unroll 8
      for (int q = 0; q < k; ++q) {
        const double c = s_rc[q], s = s_rs[q], h = s_h[q + 1];
        s_col[q] = c * carry + s * h;
        carry = -s * carry + c * h;
      }
      const double hk = carry, hk1 = sk * nrm;
      const double tn = sqrt(hk * hk + hk1 * hk1);
      const double rc = hk / tn, rs = hk1 / tn;
      const double ek = rc * e0 + rs * e1;
      const double ek1 = -rs * e0 + rc * e1;
      const double er = fabs(ek1);
      done = (er <= eps && k + 1 >= kmin) || !(er == er) || hk1 == 0.0;
      s_col[k] = tn;
      s_col[k + 1] = 0.0;
      if (w) {
        g->rc[k] = rc;
        g->rs[k] = rs;
        g->e[k] = ek;
        g->e[k + 1] = ek1;
        g->ec[k + 1] = ek1;
        g->s[k + 1] = sc;
        g->resid = er;
        g->iters = k + 1;
        g->done = done ? 1 : 0;
        if (hflag) hflag[k] = done ? 1 : 0;
      }
    }
    s_res[0] = done ? 1.0 : 0.0;
    s_res[1] = sc;
  }
  __syncthreads();
  if (k >= 0 && blockIdx.x == 0) {
    for (int q = t; q <= k + 1; q += blockDim.x) g->H[q + K1 * k] = s_col[q];
  }
  if (s_next) *s_next = s_res[1];
  return s_res[0] != 0.0;
}

}  // namespace ofd
