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

// s_out[j] = sum over b < nb of part[b][j], j < m <= 65 (all block threads)
__device__ __forceinline__ void final_sums(const double *part, int nb, int m,
                                           double *s_out) {
  __shared__ double s_acc[kRedBlock];
  const int t = threadIdx.x;
  const int G = (m <= 32) ? 8 : ((m <= 64) ? 4 : 2);
  const int jw = kRedBlock / G;
  const int j = t % jw, g = t / jw;
  double s = 0.0;
  if (j < m) {
#pragma unroll 8
    for (int b = g; b < nb; b += G) s += part[(size_t)b * kPartStride + j];
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

// Close Arnoldi step k (SparseGMRES_NS.f90:289-327) from the partials of
// |z_{k+1}|^2: H(k+1,k) = |z|, apply the previous rotations to column k,
// form the new one, rotate e, test |e_{k+1}| <= eps and k+1 >= kmin.
// k = -1 closes the initial residual instead (e_0 = |P^-1 b|, eps).
// Column entries 0..k come from h1 + h2 (CGS2) or H (MGS).  Returns the
// convergence flag; *s_next receives 1/|z| (the scale of basis vector k+1).
__device__ __forceinline__ bool arnoldi_close(GmresState *__restrict__ g,
                                              const double *__restrict__ part, int nb,
                                              int k, int cgs, int32_t *hflag,
                                              double *s_next) {
  __shared__ double s_sum[1];
  __shared__ double s_res[2];
  final_sums(part, nb, 1, s_sum);
  if (threadIdx.x == 0) {
    const double nrm = sqrt(s_sum[0]);
    const double sc = (nrm > 0.0) ? 1.0 / nrm : 0.0;
    bool done;
    const bool w = blockIdx.x == 0;
    if (k < 0) {
      done = !(nrm > 0.0 && nrm == nrm);
      if (w) {
        g->rnorm0 = nrm;
        g->eps = g->tol * nrm;
        g->e[0] = nrm;
        g->s[0] = sc;
        g->resid = nrm;
        g->iters = 0;
        g->done = done ? 1 : 0;
      }
    } else {
      const int K1 = g->kspace + 1;
      double col[kMaxKspace + 1];
      for (int q = 0; q <= k; ++q)
        col[q] = cgs ? g->h1[q] + g->h2[q] : g->H[q + K1 * k];
      col[k + 1] = nrm;
      for (int q = 0; q < k; ++q) {
        const double c = g->rc[q], s = g->rs[q];
        const double t = c * col[q] + s * col[q + 1];
        col[q + 1] = -s * col[q] + c * col[q + 1];
        col[q] = t;
      }
      const double hk = col[k], hk1 = col[k + 1];
      const double t = sqrt(hk * hk + hk1 * hk1);
      const double rc = hk / t, rs = hk1 / t;
      const double e0 = g->e[k], e1 = g->e[k + 1];
      const double ek = rc * e0 + rs * e1;
      const double ek1 = -rs * e0 + rc * e1;
      const double er = fabs(ek1);
      done = (er <= g->eps && k + 1 >= g->kmin) || !(er == er) || hk1 == 0.0;
      if (w) {
        for (int q = 0; q < k; ++q) g->H[q + K1 * k] = col[q];
        g->H[k + K1 * k] = t;
        g->H[k + 1 + K1 * k] = 0.0;
        g->rc[k] = rc;
        g->rs[k] = rs;
        g->e[k] = ek;
        g->e[k + 1] = ek1;
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
  *s_next = s_res[1];
  return s_res[0] != 0.0;
}

}  // namespace ofd
