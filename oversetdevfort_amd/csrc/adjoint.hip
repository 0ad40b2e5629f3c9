// Adjoint of the NS Newton step on gfx950 (no reference counterpart).
//
// The residual is the assembled right-hand side F of IntElmAss_NS_overset
// (assembleNavSto.f90:76-336) as the reference computes it: per Gauss point,
// with w = DetJ * weight,
//   F_u(a,i) -= w (N_a t1_i + sum_j dN_a/dx_j t2_ij),   F_p(a) -= w N_a div u
//   t1_i  = rho (ac_i + sum_j uadv_j du_i/dx_j) - f_i
//   t2_ij = -(p - tauC div u) delta_ij + mu' (du_i/dx_j + du_j/dx_i)
//   tauC  = rho sqrt(uadv.G.uadv + 12 (mu/rho)^2 G:G) / tr G,   mu' = mu + kdc
// (e3Rhs_3D_fluid with uprime = 0, e3STAB_3D); uadv = u - um.  No tauM term
// reaches F, so its derivative needs tauC only.
//
// For lambda, the Gauss-point contribution to lambda . F is
//   phi = -w (Lam . t1 + Gam : t2 + Lam_p div u),
// with Lam_i = sum_b N_b lambda_{b,i}, Gam_ij = sum_b dN_b/dx_j lambda_{b,i},
// Lam_p = sum_b N_b lambda_{b,p}, and (tr = trace of Gam, s = uadv.G.uadv +
// 12 (mu/rho)^2 G:G):
//   dphi/dac_k     = -w rho Lam_k
//   dphi/dp        =  w tr
//   dphi/d(du_k/dx_l) = -w [rho Lam_k uadv_l + (tauC tr + Lam_p) delta_kl
//                           + mu' (Gam_kl + Gam_lk)]
//   dphi/duadv_l   = -w [rho sum_i Lam_i du_i/dx_l
//                        + tr div u rho (G uadv)_l / (sqrt(s) tr G)]
//   dphi/dmu       = -w [sum_ij Gam_ij (du_i/dx_j + du_j/dx_i)
//                        + tr div u 12 mu G:G / (rho sqrt(s) tr G)]
// and node a collects N_a dphi/duadv + grad N_a . dphi/d(grad u), N_a
// dphi/dac, N_a dphi/dp and, for the per-node viscosity scale s_a
// (mu = mu0 sum_a N_a s_a), mu0 N_a dphi/dmu.
//
// The Newton increments move the alpha stage by du = alfi gami Delt d,
// dac = almi d, dp = alfi gami Delt dp, so J = -dF/d(d, dp) is applied
// transposed as -(alfi gami Delt g_u + almi g_ac, alfi gami Delt g_p).  The
// Dirichlet projection of BCLhs_3D (rows and columns of IBC dofs cleared, 1
// on the diagonal per element) transposes to: lambda masked before, the
// output masked after, plus lambda at each constrained dof once per element.
//
// Layout: one 64-lane wavefront per element, lane = (Gauss point q, node
// a), 8-lane butterflies for the Gauss-point sums, a butterfly over q for
// the node sums, and a coloured read-modify-write scatter.  Trilinear hexes
// fill the 8 x 8 lanes; on linear tets (prisms) the lanes a >= 4 (6) carry
// zero shape functions and the lanes q >= 4 (6) zero weight, so they add
// nothing.
#include "element_shape.h"
#include "ofd_internal.h"
#include "wave_dpp.h"

namespace ofd {

namespace {

constexpr int kWaves = 4;

__device__ __forceinline__ double red8(double v) { return wsum8(v); }

struct AdjParams {
  double rho, mu, kdc, delt, cu, ca, cp, gx, gy, gz;
  int vms;  // 1: the residual with the fine-scale velocity uprime = -tauM rLi
};

}  // namespace

// AT[k] = A[tmap[k]]^T for the 4x4 blocks, one thread per entry
__global__ void k_block_transpose(size_t n16, const int32_t *__restrict__ tmap,
                                  const double *__restrict__ A, double *__restrict__ AT) {
  size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; t < n16; t += stride) {
    const size_t k = t >> 4;
    const int r = (int)(t >> 2) & 3, c = (int)t & 3;
    AT[t] = A[(size_t)tmap[k] * 16 + c * 4 + r];
  }
}

// MODE 0: out = [gu (3,N) | gac (3,N) | gp (N) | gmus (N)] += raw gradients
// MODE 1: out [N][4] += J_bc^T lam
// MODE 2: out [N][4] += J_bc lam, lam a Newton increment (forward tangent)
template <int S, int MODE>
__global__ void __launch_bounds__(256)
k_vjp_ns(int N, int E, const int32_t *__restrict__ ien, const int32_t *__restrict__ elm_id,
             const int32_t *__restrict__ elems, int first, int count,
             const double *__restrict__ rec, const double *__restrict__ mus,
             const uint8_t *__restrict__ ibc, AdjParams P, int id1, int id2, int subdomain_flag,
             const double *__restrict__ lam, double *__restrict__ out) {
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const int idx = first + blockIdx.x * kWaves + w;
  const bool valid = idx < first + count;
  const int e = valid ? elems[idx] : 0;
  bool tgt = false;
  if (valid) {
    const int id = elm_id[e];
    if (subdomain_flag & 1) tgt = tgt || id == id1;
    if (subdomain_flag & 2) tgt = tgt || id == id2;
  }
  const int q = lane >> 3;
  const int a = lane & 7;
  // node lanes a >= S (tets) repeat node 0 with zero shape functions
  const int na = valid ? ien[(size_t)(a < S ? a : 0) * E + e] : 0;
  const unsigned m = valid ? ibc[na] : 0u;

  // alpha-stage record: x + d (0-2), u (3-5), ac (6-8), um (9-11), p (12)
  const double2 *r2 = reinterpret_cast<const double2 *>(rec + (size_t)na * 16);
  double xr[13];
#pragma unroll
  for (int t = 0; t < 6; ++t) {
    const double2 v = r2[t];
    xr[2 * t] = v.x;
    xr[2 * t + 1] = v.y;
  }
  xr[12] = r2[6].x;
  const double *xd = xr, *ua = xr + 3, *aca = xr + 6, *uma = xr + 9;
  const double pa = xr[12];
  // lambda at node a; the projected operator sees it with the Dirichlet
  // dofs cleared
  double lr[4];
  {
    const double2 *l2 = reinterpret_cast<const double2 *>(lam + (size_t)na * 4);
    const double2 u0 = l2[0], u1 = l2[1];
    lr[0] = u0.x;
    lr[1] = u0.y;
    lr[2] = u1.x;
    lr[3] = u1.y;
  }
  double la[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) la[c] = (MODE != 0 && ((m >> c) & 1u)) ? 0.0 : lr[c];

  // shape functions at Gauss point q (linhexshl with the truncated abscissa,
  // lintetshl with the 4-point rule)
  double Na, dl0, dl1, dl2, gwq;
  elem_shape<S>(q, a, Na, dl0, dl1, dl2, gwq);
  double J[9];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    J[i + 0] = red8(xd[i] * dl0);
    J[i + 3] = red8(xd[i] * dl1);
    J[i + 6] = red8(xd[i] * dl2);
  }
  double Ji[9];
  Ji[0] = J[4] * J[8] - J[5] * J[7];
  Ji[3] = J[5] * J[6] - J[3] * J[8];
  Ji[6] = J[3] * J[7] - J[6] * J[4];
  const double tmp = 1.0 / (Ji[0] * J[0] + Ji[3] * J[1] + Ji[6] * J[2]);
  Ji[0] *= tmp;
  Ji[3] *= tmp;
  Ji[6] *= tmp;
  Ji[1] = (J[7] * J[2] - J[1] * J[8]) * tmp;
  Ji[4] = (J[0] * J[8] - J[2] * J[6]) * tmp;
  Ji[7] = (J[1] * J[6] - J[0] * J[7]) * tmp;
  Ji[2] = (J[1] * J[5] - J[4] * J[2]) * tmp;
  Ji[5] = (J[2] * J[3] - J[0] * J[5]) * tmp;
  Ji[8] = (J[0] * J[4] - J[3] * J[1]) * tmp;
  const double dw = gwq * (1.0 / tmp);
  double ga[3];
#pragma unroll
  for (int j = 0; j < 3; ++j)
    ga[j] = dl0 * Ji[0 + 3 * j] + dl1 * Ji[1 + 3 * j] + dl2 * Ji[2 + 3 * j];
  double G[9];
#pragma unroll
  for (int j = 0; j < 3; ++j)
#pragma unroll
    for (int i = 0; i < 3; ++i)
      G[i + 3 * j] = Ji[0 + 3 * i] * Ji[0 + 3 * j] + Ji[1 + 3 * i] * Ji[1 + 3 * j] +
                     Ji[2 + 3 * i] * Ji[2 + 3 * j];

  // Gauss-point state and lambda
  double uadv[3], du[9], Lam[3], Gam[9];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    uadv[i] = red8(ua[i] * Na) - red8(uma[i] * Na);
    Lam[i] = red8(la[i] * Na);
  }
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      du[i + 3 * j] = red8(ua[i] * ga[j]);
      Gam[i + 3 * j] = red8(la[i] * ga[j]);
    }
  const double Lp = red8(la[3] * Na);
  // the fine-scale velocity and what it needs: ac, grad p at q, the gradient
  // of the pressure part of the input, tauM
  double aci[3] = {0.0, 0.0, 0.0}, rL[3] = {0.0, 0.0, 0.0}, Psi[3] = {0.0, 0.0, 0.0};
  if (P.vms) {
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      aci[i] = red8(aca[i] * Na);
      Psi[i] = red8(la[3] * ga[i]);
    }
    // grad p enters rLi below (rL holds it until then)
#pragma unroll
    for (int j = 0; j < 3; ++j) rL[j] = red8(pa * ga[j]);
  }
  const double mu = mus ? P.mu * red8(mus[na] * Na) : P.mu;
  const double rho = P.rho;
  const double mup = mu + P.kdc;
  const double divu = du[0] + du[4] + du[8];
  const double tr = Gam[0] + Gam[4] + Gam[8];
  double Gu[3], taua = 0.0, gg = 0.0;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    Gu[i] = G[i + 0] * uadv[0] + G[i + 3] * uadv[1] + G[i + 6] * uadv[2];
    taua += uadv[i] * Gu[i];
  }
#pragma unroll
  for (int t = 0; t < 9; ++t) gg += G[t] * G[t];
  const double nu = mu / rho;
  const double sq = sqrt(taua + nu * nu * 12.0 * gg);
  const double trG = G[0] + G[4] + G[8];
  const double tauC = rho * sq / trG;
  const double isq = (sq > 0.0) ? 1.0 / (sq * trG) : 0.0;
  const double Stot = 4.0 / (P.delt * P.delt) + taua + nu * nu * 12.0 * gg;
  const double tauM = 1.0 / (rho * sqrt(Stot));
  double up[3] = {0.0, 0.0, 0.0}, adv[3];
  if (P.vms) {
    const double fi[3] = {P.gx * rho, P.gy * rho, P.gz * rho};
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      rL[i] = rho * (aci[i] + uadv[0] * du[i + 0] + uadv[1] * du[i + 3] + uadv[2] * du[i + 6]) -
              fi[i] + rL[i];
      up[i] = -tauM * rL[i];
    }
  }
#pragma unroll
  for (int i = 0; i < 3; ++i) adv[i] = uadv[i] + up[i];

  double gu[3] = {0.0, 0.0, 0.0}, gac[3] = {0.0, 0.0, 0.0}, gp = 0.0, gmu = 0.0;
  if (MODE == 2) {
    // tangent of F along the increment: the alpha stage moves by cu d (u),
    // ca d (ac), cp dp (p); Lam, Gam, Lp, Psi interpolate the increment here
    double tu[3], tdu[9], tac[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      tu[i] = P.cu * Lam[i];
      tac[i] = P.ca * Lam[i];
    }
#pragma unroll
    for (int t = 0; t < 9; ++t) tdu[t] = P.cu * Gam[t];
    const double tdiv = P.cu * tr;
    const double gtu = Gu[0] * tu[0] + Gu[1] * tu[1] + Gu[2] * tu[2];
    const double ttau = rho * gtu * isq;
    const double tptot = P.cp * Lp - ttau * divu - tauC * tdiv;
    double tup[3] = {0.0, 0.0, 0.0};
    if (P.vms) {
      const double ttauM = -tauM * gtu / Stot;
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        const double trL = rho * (tac[i] + tu[0] * du[i + 0] + tu[1] * du[i + 3] +
                                  tu[2] * du[i + 6] + uadv[0] * tdu[i + 0] +
                                  uadv[1] * tdu[i + 3] + uadv[2] * tdu[i + 6]) +
                           P.cp * Psi[i];
        tup[i] = -ttauM * rL[i] - tauM * trL;
      }
    }
    double tadv[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) tadv[i] = tu[i] + tup[i];
    double t1[3], t2[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
      t1[i] = rho * (tac[i] + tadv[0] * du[i + 0] + tadv[1] * du[i + 3] + tadv[2] * du[i + 6] +
                     adv[0] * tdu[i + 0] + adv[1] * tdu[i + 3] + adv[2] * tdu[i + 6]);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j)
        t2[i + 3 * j] = ((i == j) ? -tptot : 0.0) + mup * (tdu[i + 3 * j] + tdu[j + 3 * i]) -
                        rho * (tadv[j] * up[i] + adv[j] * tup[i]);
    if (tgt) {
#pragma unroll
      for (int i = 0; i < 3; ++i)
        gu[i] = -dw * (Na * t1[i] + ga[0] * t2[i] + ga[1] * t2[i + 3] + ga[2] * t2[i + 6]);
      gp = -dw * (Na * tdiv - (ga[0] * tup[0] + ga[1] * tup[1] + ga[2] * tup[2]));
    }
  } else {
    // reverse sweep: dphi/d(.) of the Gauss-point quantities, w = -wn
    const double wn = -dw;
    double bUadv[3], bDu[9], bAc[3], bUp[3], bDp[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      // t1 (through adv) and the fine-scale stress
      double s1 = 0.0, s2 = 0.0, s3 = 0.0;
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        s1 += Lam[i] * du[i + 3 * j];
        s2 += Gam[i + 3 * j] * up[i];
        s3 += Gam[j + 3 * i] * adv[i];
      }
      bUadv[j] = wn * rho * s1 - wn * rho * s2;  // b_adv
      bUp[j] = -wn * rho * s3 - wn * Psi[j];
      bAc[j] = wn * rho * Lam[j];
    }
    const double bPtot = -wn * tr;
    const double bDiv = wn * Lp - tauC * bPtot;
    double bTauC = -divu * bPtot;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j)
        bDu[i + 3 * j] = wn * rho * Lam[i] * adv[j] + wn * mup * (Gam[i + 3 * j] + Gam[j + 3 * i]) +
                         ((i == j) ? bDiv : 0.0);
    double bMu = 0.0;
    if (P.vms) {
      // adv = uadv + up, up = -tauM rLi, rLi = rho (ac + uadv . grad u) - f + grad p
      double bTauM = 0.0, bRL[3];
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        const double bu = bUp[i] + bUadv[i];
        bTauM -= bu * rL[i];
        bRL[i] = -tauM * bu;
      }
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        bAc[i] += rho * bRL[i];
        bDp[i] = bRL[i];
      }
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        double s1 = 0.0;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          s1 += bRL[i] * du[i + 3 * j];
          bDu[i + 3 * j] += rho * bRL[i] * uadv[j];
        }
        bUadv[j] += rho * s1 - bTauM * tauM * Gu[j] / Stot;
      }
      if (MODE == 0 && mus) bMu -= bTauM * 12.0 * tauM * mu * gg / (Stot * rho * rho);
    }
#pragma unroll
    for (int l = 0; l < 3; ++l) bUadv[l] += bTauC * rho * Gu[l] * isq;
    if (tgt) {
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        gu[k] = Na * bUadv[k] + ga[0] * bDu[k + 0] + ga[1] * bDu[k + 3] + ga[2] * bDu[k + 6];
        gac[k] = Na * bAc[k];
      }
      gp = Na * bPtot + ga[0] * bDp[0] + ga[1] * bDp[1] + ga[2] * bDp[2];
    }
    if (MODE == 0 && mus && tgt) {
      double sg = 0.0;
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) sg += Gam[i + 3 * j] * (du[i + 3 * j] + du[j + 3 * i]);
      bMu += wn * sg + bTauC * 12.0 * mu * gg * isq / rho;
      gmu = Na * P.mu * bMu;
    }
  }
  // node sums over the 8 Gauss points (lanes a, a+8, ..., a+56)
#pragma unroll
  for (int s = 8; s < 64; s <<= 1) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      gu[k] += __shfl_xor(gu[k], s);
      gac[k] += __shfl_xor(gac[k], s);
    }
    gp += __shfl_xor(gp, s);
    gmu += __shfl_xor(gmu, s);
  }
  if (!valid || lane >= S) return;
  if (MODE == 0) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      out[(size_t)k * N + na] += gu[k];
      out[(size_t)(3 + k) * N + na] += gac[k];
    }
    out[(size_t)6 * N + na] += gp;
    if (mus) out[(size_t)7 * N + na] += gmu;
  } else {
    // J = -dF/d(increment): transposed (MODE 1) or applied (MODE 2), with
    // the Dirichlet projection
    double o[4];
#pragma unroll
    for (int k = 0; k < 3; ++k)
      o[k] = ((m >> k) & 1u) ? lr[k]
                             : ((MODE == 1) ? -(P.cu * gu[k] + P.ca * gac[k]) : -gu[k]);
    o[3] = ((m >> 3) & 1u) ? lr[3] : ((MODE == 1) ? -(P.cp * gp) : -gp);
    double2 *y2 = reinterpret_cast<double2 *>(out + (size_t)na * 4);
    double2 y0 = y2[0], y1 = y2[1];
    y0.x += o[0];
    y0.y += o[1];
    y1.x += o[2];
    y1.y += o[3];
    y2[0] = y0;
    y2[1] = y1;
  }
}

// out = x + a y
__global__ void k_axpby(size_t n, const double *__restrict__ x, const double *__restrict__ y,
                        double a, double *__restrict__ out) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) out[i] = x[i] + a * y[i];
}

// <w, V_j> for j < m (V_j at stride vs): block (b, j) sums chunk b of the
// entries in a fixed order; k_dot_final adds the chunk sums in order
constexpr int kDotChunks = 64;
__global__ void __launch_bounds__(256)
k_multidot(size_t n, const double *__restrict__ w, const double *__restrict__ V, size_t vs,
           double *__restrict__ part) {
  __shared__ double s[256];
  const int j = blockIdx.y, b = blockIdx.x, t = threadIdx.x;
  const double *v = V + vs * j;
  double acc = 0.0;
  for (size_t i = (size_t)b * 256 + t; i < n; i += (size_t)256 * kDotChunks) acc += w[i] * v[i];
  s[t] = acc;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (t < h) s[t] += s[t + h];
    __syncthreads();
  }
  if (t == 0) part[(size_t)j * kDotChunks + b] = s[0];
}

__global__ void k_dot_final(int m, const double *__restrict__ part, double *__restrict__ out) {
  const int j = threadIdx.x;
  if (j >= m) return;
  double a = 0.0;
  for (int b = 0; b < kDotChunks; ++b) a += part[(size_t)j * kDotChunks + b];
  out[j] = a;
}

// w += sign sum_j h_j V_j
__global__ void k_multi_axpy(size_t n, double *__restrict__ w, const double *__restrict__ V,
                             size_t vs, const double *__restrict__ h, int m, double sign) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) {
    double acc = 0.0;
    for (int j = 0; j < m; ++j) acc += h[j] * V[vs * j + i];
    w[i] += sign * acc;
  }
}

__global__ void k_scale(size_t n, const double *__restrict__ x, double a,
                        double *__restrict__ y) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) y[i] = a * x[i];
}

static AdjParams adj_params(const Ctx &c) {
  AdjParams P;
  P.rho = c.prm[OFD_P_RHO];
  P.mu = c.prm[OFD_P_MU];
  P.kdc = c.prm[OFD_P_KDC];
  P.delt = c.prm[OFD_P_DELT];
  P.cu = c.prm[OFD_P_ALFI] * c.prm[OFD_P_GAMI] * c.prm[OFD_P_DELT];
  P.ca = c.prm[OFD_P_ALMI];
  P.cp = P.cu;
  P.gx = c.prm[OFD_P_GX];
  P.gy = c.prm[OFD_P_GY];
  P.gz = c.prm[OFD_P_GZ];
  P.vms = c.res_form;
  return P;
}

static unsigned grid_for(size_t n) {
  const size_t b = (n + 255) / 256;
  return (unsigned)(b < 8192 ? (b > 0 ? b : 1) : 8192);
}

hipError_t launch_block_transpose(Ctx &c) {
  const size_t n16 = (size_t)c.icnt * 16;
  hipLaunchKernelGGL(k_block_transpose, dim3(grid_for(n16)), dim3(256), 0, c.stream, n16,
                     c.d_tmap, c.d_A64, c.t_A64);
  return hipGetLastError();
}

template <int MODE>
static hipError_t vjp_colours(Ctx &c, int subdomain_flag, const double *lam, double *out) {
  const AdjParams P = adj_params(c);
  const double *mus = c.mus_on ? c.d_field[OFD_MUSCALE] : nullptr;
  for (int col = 0; col + 1 < (int)c.color_off.size(); ++col) {
    const int first = c.color_off[col];
    const int count = c.color_off[col + 1] - first;
    if (count <= 0) continue;
    hipLaunchKernelGGL((c.nshl == 4   ? k_vjp_ns<4, MODE>
                        : c.nshl == 6 ? k_vjp_ns<6, MODE>
                                      : k_vjp_ns<8, MODE>),
                       dim3((count + kWaves - 1) / kWaves),
                       dim3(64 * kWaves), 0, c.stream, c.nnode, c.nelem, c.d_ien, c.d_elm_id,
                       c.d_color_elems, first, count, c.d_noderec, mus, c.d_ibc, P,
                       c.elm_id_set[0], c.elm_id_set[1], subdomain_flag, lam, out);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t launch_vjp_ns(Ctx &c, int subdomain_flag, const double *lam) {
  return vjp_colours<0>(c, subdomain_flag, lam, c.d_vjp);
}

hipError_t launch_jt_ns(Ctx &c, int subdomain_flag, const double *lam, double *y) {
  const hipError_t e = launch_zero(c, y, (size_t)c.nnode * 4);
  if (e != hipSuccess) return e;
  return vjp_colours<1>(c, subdomain_flag, lam, y);
}

hipError_t launch_jv_ns(Ctx &c, int subdomain_flag, const double *v, double *y) {
  const hipError_t e = launch_zero(c, y, (size_t)c.nnode * 4);
  if (e != hipSuccess) return e;
  return vjp_colours<2>(c, subdomain_flag, v, y);
}

hipError_t launch_multidot(Ctx &c, size_t n, const double *w, const double *V, size_t vs,
                           int m, double *out) {
  hipLaunchKernelGGL(k_multidot, dim3(kDotChunks, m), dim3(256), 0, c.stream, n, w, V, vs,
                     c.d_part);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_dot_final, dim3(1), dim3(128), 0, c.stream, m, c.d_part, out);
  return hipGetLastError();
}

hipError_t launch_multi_axpy(Ctx &c, size_t n, double *w, const double *V, size_t vs,
                             const double *h, int m, double sign) {
  hipLaunchKernelGGL(k_multi_axpy, dim3(grid_for(n)), dim3(256), 0, c.stream, n, w, V, vs, h,
                     m, sign);
  return hipGetLastError();
}

hipError_t launch_scale(Ctx &c, size_t n, const double *x, double a, double *y) {
  hipLaunchKernelGGL(k_scale, dim3(grid_for(n)), dim3(256), 0, c.stream, n, x, a, y);
  return hipGetLastError();
}

hipError_t launch_axpby(Ctx &c, const double *x, const double *y, double a, double *out,
                        size_t n) {
  hipLaunchKernelGGL(k_axpby, dim3(grid_for(n)), dim3(256), 0, c.stream, n, x, y, a, out);
  return hipGetLastError();
}

}  // namespace ofd
