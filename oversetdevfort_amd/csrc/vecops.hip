// Node-vector kernels: layout conversion between the reference's SoA node
// fields and the interleaved [N][4] vectors of the solver, and the
// generalized-alpha field updates of the time/Newton drivers.
#include "ofd_internal.h"

namespace ofd {

// (NNODE,3) + (NNODE) column-major -> [N][4]
__global__ void k_soa_to_node4(int N, const double *__restrict__ u,
                               const double *__restrict__ p, double *__restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  double2 *o = reinterpret_cast<double2 *>(out + (size_t)i * 4);
  o[0] = make_double2(u[i], u[(size_t)N + i]);
  o[1] = make_double2(u[2 * (size_t)N + i], p ? p[i] : 0.0);
}

__global__ void k_node4_to_soa(int N, const double *__restrict__ in,
                               double *__restrict__ u, double *__restrict__ p) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const double2 *s = reinterpret_cast<const double2 *>(in + (size_t)i * 4);
  const double2 a = s[0], b = s[1];
  u[i] = a.x;
  u[(size_t)N + i] = a.y;
  u[2 * (size_t)N + i] = b.x;
  if (p) p[i] = b.y;
}

// which: 0 Newton update (solveOverset.f90:76-78), 1 predictor
// (driver.f90:83-85), 2 old <- new (driver.f90:98-103)
__global__ void k_fields(int N, int which, double gami, double delt, double alfi,
                         const double *__restrict__ x, double *ug, double *ugold,
                         double *acg, double *acgold, double *pg, double *pgold,
                         double *dg, double *dgold, double *ugm, double *ugmold,
                         double *acgm, double *acgmold) {
#pragma clang fp contract(off)  // products and sums rounded separately
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  if (which == 0) {
    const double fu = gami * delt;
    const double fp = alfi * gami * delt;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const size_t o = (size_t)c * N + i;
      const double d = x[(size_t)i * 4 + c];
      acg[o] = acg[o] + d;
      ug[o] = ug[o] + fu * d;
    }
    pg[i] = pg[i] + fp * x[(size_t)i * 4 + 3];
  } else if (which == 1) {
    const double f = (gami - 1.0) / gami;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const size_t o = (size_t)c * N + i;
      ug[o] = ugold[o];
      acg[o] = f * acgold[o];
    }
    pg[i] = pgold[i];
  } else {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const size_t o = (size_t)c * N + i;
      ugold[o] = ug[o];
      acgold[o] = acg[o];
      dgold[o] = dg[o];
      ugmold[o] = ugm[o];
      acgmold[o] = acgm[o];
    }
    pgold[i] = pg[i];
  }
}

// setBCs_CFD (setBCs.f90:84-91), value part
__global__ void k_apply_dirichlet(int N, const uint8_t *__restrict__ ibc,
                                  const double *__restrict__ val, double dtgl,
                                  double gami, double *ug, const double *ugold,
                                  double *acg, const double *acgold) {
#pragma clang fp contract(off)  // products and sums rounded separately
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const unsigned m = ibc[i];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    if ((m >> c) & 1u) {
      const size_t o = (size_t)c * N + i;
      ug[o] = val[o];
      acg[o] = ((ug[o] - ugold[o]) * dtgl + (gami - 1.0) * acgold[o]) / gami;
    }
  }
}

// plain device copy in 16-byte units, grid-stride: the bandwidth and
// hardware-counter reference for the bandwidth-bound kernels
// (4 independent 16-byte loads in flight per thread and pass)
__global__ void __launch_bounds__(256) k_copy(size_t n16, const double2 *__restrict__ src,
                                              double2 *__restrict__ dst) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (; i + 3 * stride < n16; i += 4 * stride) {
    double2 v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) v[u] = src[i + u * stride];
#pragma unroll
    for (int u = 0; u < 4; ++u) dst[i + u * stride] = v[u];
  }
  for (; i < n16; i += stride) dst[i] = src[i];
}

// read-only stream: every thread sums its grid-stride share of 16-byte
// loads (8 in flight per pass) and each block stores one value, so no load
// can be dropped; the achievable-read reference for the SpMV, whose matrix
// stream is read-only
__global__ void __launch_bounds__(256) k_read(size_t n16, const double2 *__restrict__ src,
                                              double *__restrict__ out) {
  __shared__ double s[256];
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  double acc[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (; i + 7 * stride < n16; i += 8 * stride) {
    double2 v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = src[i + u * stride];
#pragma unroll
    for (int u = 0; u < 8; ++u) acc[u] += v[u].x + v[u].y;
  }
  for (; i < n16; i += stride) acc[0] += src[i].x + src[i].y;
  double t = 0.0;
#pragma unroll
  for (int u = 0; u < 8; ++u) t += acc[u];
  s[threadIdx.x] = t;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) s[threadIdx.x] += s[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = s[0];
}

// out holds kReadBlocks doubles
hipError_t launch_read(Ctx &c, const void *src, size_t nbytes, double *out) {
  const size_t n16 = nbytes / 16;
  size_t want = (n16 + 8 * 256 - 1) / (8 * 256);
  if (want < 1) want = 1;
  const unsigned nb = (unsigned)(want < (size_t)kReadBlocks ? want : (size_t)kReadBlocks);
  hipLaunchKernelGGL(k_read, dim3(nb), dim3(256), 0, c.stream, n16,
                     static_cast<const double2 *>(src), out);
  return hipGetLastError();
}

hipError_t launch_copy(Ctx &c, const void *src, void *dst, size_t nbytes) {
  const size_t n16 = nbytes / 16;
  const size_t want = (n16 + 4 * 256 - 1) / (4 * 256);
  const unsigned nb = (unsigned)(want < 16384 ? (want > 0 ? want : 1) : 16384);
  hipLaunchKernelGGL(k_copy, dim3(nb), dim3(256), 0, c.stream, n16,
                     static_cast<const double2 *>(src), static_cast<double2 *>(dst));
  return hipGetLastError();
}

hipError_t launch_apply_dirichlet(Ctx &c) {
  double **f = c.d_field;
  hipLaunchKernelGGL(k_apply_dirichlet, dim3((c.nnode + 255) / 256), dim3(256), 0,
                     c.stream, c.nnode, c.d_ibc, c.d_bcval, 1.0 / c.prm[OFD_P_DELT],
                     c.prm[OFD_P_GAMI], f[OFD_UG], f[OFD_UGOLD], f[OFD_ACG],
                     f[OFD_ACGOLD]);
  return hipGetLastError();
}

hipError_t launch_soa_to_node4(Ctx &c, const double *u, const double *p, double *out) {
  hipLaunchKernelGGL(k_soa_to_node4, dim3((c.nnode + 255) / 256), dim3(256), 0,
                     c.stream, c.nnode, u, p, out);
  return hipGetLastError();
}

hipError_t launch_node4_to_soa(Ctx &c, const double *in, double *u, double *p) {
  hipLaunchKernelGGL(k_node4_to_soa, dim3((c.nnode + 255) / 256), dim3(256), 0,
                     c.stream, c.nnode, in, u, p);
  return hipGetLastError();
}

// scalar field pair (a, ra) = (Tg, rTg) or (phig, rphig): which 0 Newton
// update ra += x, a += gami*Delt*x (solveFlow.f90:139-140); 1 predictor
// a = aold, ra = (gami-1)/gami raold; 2 old <- new
__global__ void k_fields_s(int N, int which, double gami, double delt,
                           const double *__restrict__ x, double *a, double *aold, double *ra,
                           double *raold) {
#pragma clang fp contract(off)  // products and sums rounded separately
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  if (which == 0) {
    const double d = x[i];
    ra[i] = ra[i] + d;
    a[i] = a[i] + gami * delt * d;
  } else if (which == 1) {
    a[i] = aold[i];
    ra[i] = (gami - 1.0) / gami * raold[i];
  } else {
    aold[i] = a[i];
    raold[i] = ra[i];
  }
}

hipError_t launch_fields_s(Ctx &c, int which, const double *x, int fa, int faold, int fra,
                           int fraold) {
  double **f = c.d_field;
  hipLaunchKernelGGL(k_fields_s, dim3((c.nnode + 255) / 256), dim3(256), 0, c.stream, c.nnode,
                     which, c.prm[OFD_P_GAMI], c.prm[OFD_P_DELT], x, f[fa], f[faold], f[fra],
                     f[fraold]);
  return hipGetLastError();
}

// NSVOF Newton update (solveFlow.f90:99-104) from x = [N][4] (u, p), [N] phi
__global__ void k_fields_v(int N, double gami, double delt, double alfi,
                           const double *__restrict__ x, double *ug, double *acg, double *pg,
                           double *rphig, double *phig) {
#pragma clang fp contract(off)  // products and sums rounded separately
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const double *d = x + (size_t)i * 4;
  const double dphi = x[(size_t)4 * N + i];
  const double fu = gami * delt;
  for (int c = 0; c < 3; ++c) {
    const size_t o = (size_t)c * N + i;
    acg[o] = acg[o] + d[c];
    ug[o] = ug[o] + fu * d[c];
  }
  pg[i] = pg[i] + alfi * gami * delt * d[3];
  rphig[i] = rphig[i] + dphi;
  phig[i] = phig[i] + fu * dphi;
}

// NSVOF vector ([N][4] u, p, then [N] phi) -> u (N,3), p (N), phi (N)
__global__ void k_nsvof_to_soa(int N, const double *__restrict__ in, double *u, double *p,
                               double *phi) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const double2 *s = reinterpret_cast<const double2 *>(in + (size_t)i * 4);
  const double2 a = s[0], b = s[1];
  u[i] = a.x;
  u[(size_t)N + i] = a.y;
  u[2 * (size_t)N + i] = b.x;
  p[i] = b.y;
  phi[i] = in[(size_t)4 * N + i];
}

hipError_t launch_fields_v(Ctx &c) {
  double **f = c.d_field;
  hipLaunchKernelGGL(k_fields_v, dim3((c.nnode + 255) / 256), dim3(256), 0, c.stream, c.nnode,
                     c.prm[OFD_P_GAMI], c.prm[OFD_P_DELT], c.prm[OFD_P_ALFI], c.d_xV, f[OFD_UG],
                     f[OFD_ACG], f[OFD_PG], f[OFD_RPHIG], f[OFD_PHIG]);
  return hipGetLastError();
}

hipError_t launch_nsvof_to_soa(Ctx &c, const double *in) {
  double **f = c.d_field;
  hipLaunchKernelGGL(k_nsvof_to_soa, dim3((c.nnode + 255) / 256), dim3(256), 0, c.stream,
                     c.nnode, in, f[OFD_SOLU], f[OFD_SOLP], f[OFD_SOLPHI]);
  return hipGetLastError();
}

hipError_t launch_axpy_fields(Ctx &c, int which) {
  double **f = c.d_field;
  hipLaunchKernelGGL(k_fields, dim3((c.nnode + 255) / 256), dim3(256), 0, c.stream,
                     c.nnode, which, c.prm[OFD_P_GAMI], c.prm[OFD_P_DELT],
                     c.prm[OFD_P_ALFI], c.d_x, f[OFD_UG], f[OFD_UGOLD], f[OFD_ACG],
                     f[OFD_ACGOLD], f[OFD_PG], f[OFD_PGOLD], f[OFD_DG], f[OFD_DGOLD],
                     f[OFD_UGM], f[OFD_UGMOLD], f[OFD_ACGM], f[OFD_ACGMOLD]);
  return hipGetLastError();
}

}  // namespace ofd
