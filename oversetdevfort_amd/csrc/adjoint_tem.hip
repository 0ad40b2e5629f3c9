// Adjoint of the quenching temperature equation on gfx950 (no reference
// counterpart).
//
// The residual is the assembled RHSGTem of IntElmAss_Tem_Quenching_STAB
// (solveQuenching.f90:579-882) as the forward kernel computes it.  Per Gauss
// point, with w = DetJ * weight, node a receives
//   fluid  r_a = -w [N_a res + rc tau sc_a res + hk grad N_a . grad T]
//   solid  r_a = -w [N_a rc rT + hk grad N_a . grad T]
//   res = rc (rT + uadv . grad T) - Se(T),  sc_a = grad N_a . uadv,
//   Se  = ((cpw - cpa)(T - Ts) - lh) mdot(T)   (mdot linear in T - Ts on each
//         side of Ts, :742-752),
//   tau = 1 / (rho cp sqrt(tt)),  tt = uadv.G.uadv + 4/Delt^2
//         + 12 (k/(rho cp))^2 G:G    (e3STAB_3D_TEM, e3STAB.f90:338-367),
//   hk  = k + kdc (fluid), k (solid);  kdc = TEM_KDC |res| / sqrt(G:G) / tref,
//   k   = kappa(phi) sum_a N_a s_a    (s the per-node conductivity scale),
// and rc = (rho cp)(phi).  For lambda, with L0 = sum_a lambda_a N_a,
// Lg = sum_a lambda_a grad N_a, Lsc = Lg . uadv, LdT = Lg . grad T, the
// contribution to lambda . R is
//   fluid  phi = -w [res (L0 + rc tau Lsc) + hk LdT]
//   solid  phi = -w [L0 rc rT + k LdT]
// and its reverse sweep (b_x = dphi/dx):
//   b_hk  = -w LdT
//   b_res = -w (L0 + rc tau Lsc) + b_hk dkdc/dres
//   b_tau = -w res rc Lsc,  b_tt = -b_tau tau / (2 tt)
//   b_k   = b_hk + b_tt 24 G:G k / (rho cp)^2,   b_s = kappa b_k
//   b_rT  = rc b_res,  b_T = -b_res dSe/dT (the exact derivative),
//   b_gradT = rc b_res uadv + b_LdT Lg,           b_LdT = -w hk
//   b_uadv  = rc b_res grad T - w res rc tau Lg + 2 b_tt G uadv.
// Node a collects N_a b_T + grad N_a . b_gradT, N_a b_rT, N_a b_uadv and
// N_a b_s.  The Newton increments move the alpha stage by
// dT = alfi gami Delt d and drT = almi d, so J_T = -dR/dd is applied
// transposed as -(alfi gami Delt g_T + almi g_rT); the BCLhs_tem projection
// (rows and columns of IBC column 6 cleared, 1 on the diagonal per element)
// transposes to lambda masked before, the output masked after, plus lambda
// at each constrained node once per element.
//
// Layout as the forward kernel: one 64-lane wavefront per element, lane =
// (Gauss point q, node a), 8-lane butterflies for the Gauss-point sums, a
// butterfly over q for the node sums, a coloured read-modify-write scatter.
#include "ofd_internal.h"
#include "wave_dpp.h"

namespace ofd {

namespace {

constexpr int kWaves = 4;
__constant__ int8_t c_sgn_tadj[8][3] = {{-1, 1, -1}, {1, 1, -1}, {1, 1, 1},  {-1, 1, 1},
                                        {-1, -1, -1}, {1, -1, -1}, {1, -1, 1}, {-1, -1, 1}};

__device__ __forceinline__ double red8t(double v) { return wsum8(v); }

// Gauss-point value summed over the element's 8 nodes in node order, as the
// forward kernel forms it (the order of the reference's e3int_qr).  Used
// for the values compared against a threshold (the VOF field clamped to
// [0, 1], T against T_sat): inside one phase every nodal value is exactly 0
// or 1, and a tree-ordered sum can land one rounding on the other side of
// the clamp and flip the branch.
__device__ __forceinline__ double seq8t(double v) {
  const int b = (int)(threadIdx.x & 63) & ~7;
  double s = 0.0;
#pragma unroll
  for (int j = 0; j < 8; ++j) s = s + __shfl(v, b + j);
  return s;
}

__device__ __forceinline__ double interp(double a, double b, double phi) {
  const double pc = fmax(0.0, fmin(1.0, phi));
  return a * (1.0 - pc) + b * pc;
}

struct TemAdjParams {
  double almi, cT;  // almi, alfi gami Delt
  double delt;
  double rhoa, rhow, cpa, cpw, kappaa, kappaw, rhos, cps, kappas;
  double Ts, lh, c_cond, c_evap, tem_kdc, tref;
  int fluid_id;
};

}  // namespace

// tT_A[k] = A[tmap[k]]: the scalar LHSTem transposed
__global__ void k_scalar_transpose(int64_t n, const int32_t *__restrict__ tmap,
                                   const double *__restrict__ A, double *__restrict__ AT) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; k < n; k += stride) AT[k] = A[tmap[k]];
}

// MODE 0: out = [gT (N) | grT (N) | gu (3,N) | gks (N)] += raw gradients
// MODE 1: out [N] += J_T,bc^T lam
template <int MODE>
__global__ void __launch_bounds__(256)
k_vjp_tem_hex(int N, int E, const int32_t *__restrict__ ien, const int32_t *__restrict__ elm_id,
              const int32_t *__restrict__ elems, int first, int count,
              const double *__restrict__ rec, const double *__restrict__ ks,
              const uint8_t *__restrict__ ibc, TemAdjParams P, const double *__restrict__ lam,
              double *__restrict__ out) {
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const int idx = first + blockIdx.x * kWaves + w;
  const bool valid = idx < first + count;
  const int e = valid ? elems[idx] : 0;
  const bool fluid = valid && elm_id[e] == P.fluid_id;
  const int q = lane >> 3;
  const int a = lane & 7;
  const int na = valid ? ien[(size_t)a * E + e] : 0;
  const bool dir = valid && ((ibc[na] >> 5) & 1u);

  // alpha-stage record: x + d (0-2), u (3-5), ac (6-8), um (9-11), p, phi, T, rT
  const double2 *r2 = reinterpret_cast<const double2 *>(rec + (size_t)na * 16);
  const double2 v0 = r2[0], v1 = r2[1], v2 = r2[2], v4 = r2[4], v5 = r2[5], v6 = r2[6],
                v7 = r2[7];
  const double xd[3] = {v0.x, v0.y, v1.x};
  const double ua[3] = {v1.y, v2.x, v2.y};
  const double uma[3] = {v4.y, v5.x, v5.y};
  const double phia = v6.y, Ta = v7.x, rTa = v7.y;
  const double lr = valid ? lam[na] : 0.0;
  const double la = (MODE != 0 && dir) ? 0.0 : lr;

  // linhexshl shape functions at Gauss point q (the truncated abscissa)
  const double gpt = 0.57735026918963;
  const double xi = (q & 1) ? gpt : -gpt;
  const double eta = (q & 2) ? gpt : -gpt;
  const double zeta = (q & 4) ? gpt : -gpt;
  const double sx = c_sgn_tadj[a][0], sy = c_sgn_tadj[a][1], sz = c_sgn_tadj[a][2];
  const double fx = (sx > 0) ? (xi + 1.0) : (1.0 - xi);
  const double fy = (sy > 0) ? (eta + 1.0) : (1.0 - eta);
  const double fz = (sz > 0) ? (zeta + 1.0) : (1.0 - zeta);
  const double Na = fx * fy * fz * 0.125;
  const double dl0 = sx * (fy * fz * 0.125);
  const double dl1 = sy * (fx * fz * 0.125);
  const double dl2 = sz * (fx * fy * 0.125);
  double J[9];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    J[i + 0] = red8t(xd[i] * dl0);
    J[i + 3] = red8t(xd[i] * dl1);
    J[i + 6] = red8t(xd[i] * dl2);
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
  const double dw = 1.0 / tmp;  // 8-point rule weights are 1
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

  // Gauss-point values
  const double phii = seq8t(phia * Na);
  const double Ti = seq8t(Ta * Na);
  const double rTi = red8t(rTa * Na);
  double uadv[3], dT[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) uadv[i] = red8t(ua[i] * Na) - red8t(uma[i] * Na);
#pragma unroll
  for (int j = 0; j < 3; ++j) dT[j] = red8t(Ta * ga[j]);
  const double sq = ks ? red8t(ks[na] * Na) : 1.0;
  // lambda at the Gauss point and its gradient
  const double L0 = red8t(la * Na);
  double Lg[3];
#pragma unroll
  for (int j = 0; j < 3; ++j) Lg[j] = red8t(la * ga[j]);

  // phase change with the exact T derivative of mdot
  const double phic = fmax(fmin(phii, 1.0), 0.0);
  double mdot, dmdT;
  if (Ti > P.Ts) {
    mdot = P.c_evap * (1 - phic) * P.rhow * (Ti - P.Ts) / P.Ts;
    dmdT = P.c_evap * (1 - phic) * P.rhow / P.Ts;
  } else {
    mdot = P.c_cond * (phic)*P.rhoa * (Ti - P.Ts) / P.Ts;
    dmdT = P.c_cond * (phic)*P.rhoa / P.Ts;
  }
  double rhoi, cpi, k0, rc;
  if (fluid) {
    rhoi = interp(P.rhow, P.rhoa, phic);
    cpi = interp(P.cpw, P.cpa, phic);
    k0 = interp(P.kappaw, P.kappaa, phic);
    rc = interp(P.rhow * P.cpw, P.rhoa * P.cpa, phic);
  } else {
    rhoi = P.rhos;
    cpi = P.cps;
    k0 = P.kappas;
    rc = P.rhos * P.cps;
  }
  const double k = k0 * sq;
  const double LdT = Lg[0] * dT[0] + Lg[1] * dT[1] + Lg[2] * dT[2];

  double bT = 0.0, brT = 0.0, bdT[3] = {0.0, 0.0, 0.0}, bu[3] = {0.0, 0.0, 0.0}, bk = 0.0;
  if (fluid) {
    const double Se = ((P.cpw - P.cpa) * (Ti - P.Ts) - P.lh) * mdot;
    const double dSe = (P.cpw - P.cpa) * mdot + ((P.cpw - P.cpa) * (Ti - P.Ts) - P.lh) * dmdT;
    const double udT = uadv[0] * dT[0] + uadv[1] * dT[1] + uadv[2] * dT[2];
    const double res = rc * (rTi + udT) - Se;
    double Gu[3];
#pragma unroll
    for (int l = 0; l < 3; ++l)
      Gu[l] = G[l + 0] * uadv[0] + G[l + 3] * uadv[1] + G[l + 6] * uadv[2];
    double gg = 0.0;
#pragma unroll
    for (int t = 0; t < 9; ++t) gg = gg + G[t] * G[t];
    const double dcp = rhoi * cpi;
    const double kk = k / dcp;
    const double tt = uadv[0] * Gu[0] + uadv[1] * Gu[1] + uadv[2] * Gu[2] +
                      4.0 / (P.delt * P.delt) + kk * kk * 12.0 * gg;
    const double tau = 1.0 / (dcp * sqrt(tt));
    double kdc = 0.0, dkdc = 0.0;
    if (fabs(P.tem_kdc) > 0.0) {
      kdc = P.tem_kdc * fabs(res) / sqrt(gg) / P.tref;
      dkdc = (res > 0.0 ? 1.0 : (res < 0.0 ? -1.0 : 0.0)) * P.tem_kdc / sqrt(gg) / P.tref;
    }
    const double hk = k + kdc;
    const double Lsc = Lg[0] * uadv[0] + Lg[1] * uadv[1] + Lg[2] * uadv[2];
    const double bhk = -dw * LdT;
    const double bres = -dw * (L0 + rc * tau * Lsc) + bhk * dkdc;
    const double btau = -dw * res * rc * Lsc;
    const double btt = -0.5 * btau * tau / tt;
    const double bLsc = -dw * res * rc * tau;
    bk = bhk + btt * 24.0 * gg * k / (dcp * dcp);
    brT = rc * bres;
    bT = -bres * dSe;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      bdT[j] = rc * bres * uadv[j] - dw * hk * Lg[j];
      bu[j] = rc * bres * dT[j] + bLsc * Lg[j] + 2.0 * btt * Gu[j];
    }
  } else {
    brT = -dw * L0 * rc;
#pragma unroll
    for (int j = 0; j < 3; ++j) bdT[j] = -dw * k * Lg[j];
    bk = -dw * LdT;
  }
  double gT = 0.0, grT = 0.0, gu[3] = {0.0, 0.0, 0.0}, gk = 0.0;
  if (valid) {
    gT = Na * bT + ga[0] * bdT[0] + ga[1] * bdT[1] + ga[2] * bdT[2];
    grT = Na * brT;
#pragma unroll
    for (int j = 0; j < 3; ++j) gu[j] = Na * bu[j];
    gk = Na * k0 * bk;
  }
  // node sums over the 8 Gauss points (lanes a, a+8, ..., a+56)
#pragma unroll
  for (int s = 8; s < 64; s <<= 1) {
    gT += __shfl_xor(gT, s);
    grT += __shfl_xor(grT, s);
#pragma unroll
    for (int j = 0; j < 3; ++j) gu[j] += __shfl_xor(gu[j], s);
    gk += __shfl_xor(gk, s);
  }
  if (!valid || lane >= 8) return;
  if (MODE == 0) {
    out[na] += gT;
    out[(size_t)N + na] += grT;
#pragma unroll
    for (int j = 0; j < 3; ++j) out[(size_t)(2 + j) * N + na] += gu[j];
    if (ks) out[(size_t)5 * N + na] += gk;
  } else {
    out[na] += dir ? lr : -(P.cT * gT + P.almi * grT);
  }
}

static TemAdjParams tem_adj_params(const Ctx &c) {
  const double *q = c.qprm;
  TemAdjParams P;
  P.almi = c.prm[OFD_P_ALMI];
  P.cT = c.prm[OFD_P_ALFI] * c.prm[OFD_P_GAMI] * c.prm[OFD_P_DELT];
  P.delt = c.prm[OFD_P_DELT];
  P.rhoa = q[OFD_Q_RHOA];
  P.rhow = q[OFD_Q_RHOW];
  P.cpa = q[OFD_Q_CPA];
  P.cpw = q[OFD_Q_CPW];
  P.kappaa = q[OFD_Q_KAPPAA];
  P.kappaw = q[OFD_Q_KAPPAW];
  P.rhos = q[OFD_Q_RHOS];
  P.cps = q[OFD_Q_CPS];
  P.kappas = q[OFD_Q_KAPPAS];
  P.Ts = q[OFD_Q_TS];
  P.lh = q[OFD_Q_LH];
  P.c_cond = q[OFD_Q_CCOND];
  P.c_evap = q[OFD_Q_CEVAP];
  P.tem_kdc = q[OFD_Q_TEMKDC];
  P.tref = q[OFD_Q_TREF];
  P.fluid_id = (int)q[OFD_Q_FLUID_ID];
  return P;
}

template <int MODE>
static hipError_t vjp_tem_colours(Ctx &c, const double *lam, double *out) {
  const TemAdjParams P = tem_adj_params(c);
  const double *ks = c.ks_on ? c.d_field[OFD_KSCALE] : nullptr;
  for (int col = 0; col + 1 < (int)c.color_off.size(); ++col) {
    const int first = c.color_off[col];
    const int count = c.color_off[col + 1] - first;
    if (count <= 0) continue;
    hipLaunchKernelGGL((k_vjp_tem_hex<MODE>), dim3((count + kWaves - 1) / kWaves),
                       dim3(64 * kWaves), 0, c.stream, c.nnode, c.nelem, c.d_ien, c.d_elm_id,
                       c.d_color_elems, first, count, c.d_noderec, ks, c.d_ibc, P, lam, out);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t launch_scalar_transpose(Ctx &c) {
  const int64_t n = c.icnt;
  const int64_t b = (n + 255) / 256;
  const unsigned grid = (unsigned)(b < 8192 ? (b > 0 ? b : 1) : 8192);
  hipLaunchKernelGGL(k_scalar_transpose, dim3(grid), dim3(256), 0, c.stream, n, c.d_tmap,
                     c.d_AT, c.tT_A);
  return hipGetLastError();
}

hipError_t launch_vjp_tem(Ctx &c, const double *lam, double *out) {
  return vjp_tem_colours<0>(c, lam, out);
}

// tets and prisms: the dual-number operator (d_vjpT, [7N], as its scratch)
hipError_t launch_jt_tem(Ctx &c, const double *lam, double *y) {
  if (c.nshl != 8) return launch_jt_tem_dual(c, lam, y, c.d_vjpT);
  const hipError_t e = launch_zero(c, y, (size_t)c.temT_n4 * 4);
  if (e != hipSuccess) return e;
  return vjp_tem_colours<1>(c, lam, y);
}

}  // namespace ofd
