// Coupled Navier-Stokes + VOF assembly of the quenching solver on gfx950,
// hex elements.
//
// Restates IntElmAss_NSVOF_Quenching_STAB (solveQuenching.f90:122-573):
// the VMS momentum/continuity equations with phase-change mass transfer
// (vdot, dmdot/dphi; :343-366), the SUPG-stabilised VOF transport (:386-387,
// :452-494), tau's from e3STAB_3D_NSVOF (e3STAB.f90:292-333), LSIC terms;
// then the 5-field Dirichlet projection BCLhs_NS_conv (BCLhs_3D.f90:278-
// 428; IBC columns 1-3 = u, 7 = p, 8 = phi) and the scatter
// FillSparseMat_NS_conv / LocaltoGlobalNSVOF_3D.  Only fluid elements
// (ELM_ID 102) contribute; solid elements scatter their projected zero
// element matrices.  The NS discontinuity-capturing coefficient is 0.
//
// Storage of the system, 24 doubles per nonzero k in two arrays (one
// allocation, the pairs after the Z blocks):
//   A[k][0..15]     the 4x4 block [K11 G; D1 M] row-major, as the NS system
//   Apair[k][0..7]  the phi couplings as pairs (col phi of row r, row phi of
//                   col r): (ULS(0), LSU(0)) (ULS(1), LSU(1)) (ULS(2), LSU(2))
//                   (PLS, LS)
// so a block is 128 and a pair group 64 aligned bytes (FP32 copies: 64 and
// 32), which the products read whole.
// Node vectors are [N][4] (u0 u1 u2 p) followed by [N] phi.
//
// One 64-lane wavefront per element, as the NS kernel: lane = (Gauss point,
// node) for shape functions and Gauss-point data (LDS), lane = (a, b) for
// the 24 entries of block (a, b), masks for BCLhs_NS_conv, and a coloured
// read-modify-write scatter.
#include "ofd_internal.h"
#include "wave_dpp.h"
#include "element_shape.h"

namespace ofd {

namespace {

constexpr int kWaves = 4;
constexpr int kQs = 14;  // per-Gauss-point scalars staged in LDS

__device__ __forceinline__ double s8(double v) { return wsum8(v); }

// Gauss-point value summed over the element's 8 nodes in node order, the
// order of the reference's e3int_qr.  Used for the values compared against a
// threshold (the VOF field clamped to [0, 1], T against T_sat): inside one
// phase every nodal value is exactly 0 or 1, and a tree-ordered sum can land
// one rounding on the other side of the clamp and flip the branch.
__device__ __forceinline__ double seq8(double v) {
  const int b = (int)(threadIdx.x & 63) & ~7;
  double s = 0.0;
#pragma unroll
  for (int j = 0; j < 8; ++j) s = s + __shfl(v, b + j);
  return s;
}

__device__ __forceinline__ double pinterp(double a, double b, double phi) {
  const double pc = fmax(0.0, fmin(1.0, phi));
  return a * (1.0 - pc) + b * pc;
}

}  // namespace

struct NsvofParams {
  double almi, alfi, gami, delt, gx, gy, gz;
  double rhoa, rhow, mua, muw, rhos, mus, Ts, c_cond, c_evap, c_dmdot, lsc_kdc;
  int fluid_id;
};

template <int NSHL>
__global__ void __launch_bounds__(256)
k_assemble_nsvof_el(int E, const int32_t *__restrict__ ien, const int32_t *__restrict__ elm_id,
                     const int32_t *__restrict__ elems, int first, int count,
                     const int32_t *__restrict__ emap,
                     const uint64_t *__restrict__ firstmask, const double *__restrict__ rec,
                     const uint8_t *__restrict__ ibc, NsvofParams P, int do_mat, int do_vec,
                     int N, double *__restrict__ A, double *__restrict__ Apair,
                     double *__restrict__ rhs,
                     double *__restrict__ cfl_out, int32_t *__restrict__ negdet) {
  __shared__ double s_sh[kWaves][8][8][5];  // [q][a]: N, dN/dx (3), shconv
  __shared__ double s_q[kWaves][8][kQs];
  __shared__ int32_t s_node[kWaves][8];

  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const int idx = first + blockIdx.x * kWaves + w;
  const bool valid = idx < first + count;
  const int e = valid ? elems[idx] : 0;
  const bool fluid = valid && elm_id[e] == P.fluid_id;

  // ---------------- phase A: lane = (q, a) ----------------
  const int q = lane >> 3;
  const int a = lane & 7;
  const int na = valid ? ien[(size_t)(a < NSHL ? a : 0) * E + e] : 0;
  if (lane < 8) s_node[w][lane] = na;
  // record: x+d (0-2), u (3-5), ac (6-8), um (9-11), p, phi, T, rphi (12-15)
  const double2 *r2 = reinterpret_cast<const double2 *>(rec + (size_t)na * 16);
  double rv[16];
#pragma unroll
  for (int t = 0; t < 8; ++t) {
    const double2 v = r2[t];
    rv[2 * t] = v.x;
    rv[2 * t + 1] = v.y;
  }
  const double *xd = rv, *ua = rv + 3, *aca = rv + 6, *uma = rv + 9;
  const double pa = rv[12], phia = rv[13], Ta = rv[14], rphia = rv[15];

  double Na, dl0, dl1, dl2, gwq;
  elem_shape<NSHL>(q, a, Na, dl0, dl1, dl2, gwq);
  double J[9];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    J[i + 0] = s8(xd[i] * dl0);
    J[i + 3] = s8(xd[i] * dl1);
    J[i + 6] = s8(xd[i] * dl2);
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
  const double detj = 1.0 / tmp;
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
  double ui[3], aci[3], umi[3], du[9], dp[3], dphi[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    ui[i] = s8(ua[i] * Na);
    aci[i] = s8(aca[i] * Na);
    umi[i] = s8(uma[i] * Na);
  }
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) du[i + 3 * j] = s8(ua[i] * ga[j]);
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    dp[j] = s8(pa * ga[j]);
    dphi[j] = s8(phia * ga[j]);
  }
  const double pri = s8(pa * Na), phii = seq8(phia * Na), Ti = seq8(Ta * Na);
  const double rphii = s8(rphia * Na);

  // phase change and properties (:343-366)
  const double phic = fmax(fmin(phii, 1.0), 0.0);
  double mdot, dmdphi;
  if (Ti > P.Ts) {
    mdot = P.c_evap * (1 - phic) * P.rhow * (Ti - P.Ts) / P.Ts;
    dmdphi = -P.c_evap * P.rhow * (Ti - P.Ts) / P.Ts * P.c_dmdot;
  } else {
    mdot = P.c_cond * (phic)*P.rhoa * (Ti - P.Ts) / P.Ts;
    dmdphi = P.c_cond * P.rhoa * (Ti - P.Ts) / P.Ts * P.c_dmdot;
  }
  const double vdot = mdot / P.rhoa - mdot / P.rhow;
  double rhoi, mui;
  if (fluid) {
    rhoi = pinterp(P.rhow, P.rhoa, phic);
    mui = pinterp(P.muw, P.mua, phic);
  } else {
    rhoi = P.rhos;
    mui = P.mus;
  }
  const double lam = -2.0 / 3.0 * mui;
  if (phii < 0.0 || phii > 1.0) dmdphi = 0.0;
  double uadv[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) uadv[i] = ui[i] - umi[i];
  const double divu = du[0] + du[4] + du[8];
  const double fi[3] = {P.gx * rhoi, P.gy * rhoi, P.gz * rhoi};
  double rLi[3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
    rLi[i] = rhoi * aci[i] +
             rhoi * (du[i + 0] * uadv[0] + du[i + 3] * uadv[1] + du[i + 6] * uadv[2]) + dp[i] -
             fi[i];
  const double res = rphii + (uadv[0] * dphi[0] + uadv[1] * dphi[1] + uadv[2] * dphi[2]) +
                     phii * vdot - mdot / P.rhoa;
  // e3STAB_3D_NSVOF
  double taua = 0.0, gg = 0.0;
#pragma unroll
  for (int j = 0; j < 3; ++j)
#pragma unroll
    for (int i = 0; i < 3; ++i) taua = taua + uadv[i] * G[i + 3 * j] * uadv[j];
#pragma unroll
  for (int t = 0; t < 9; ++t) gg = gg + G[t] * G[t];
  const double nu = mui / rhoi;
  const double taud = nu * nu * 12.0 * gg;
  const double taut = 4.0 / (P.delt * P.delt);
  const double tauM = 1.0 / (rhoi * sqrt(taut + taua + taud));
  const double tauLS = 1.0 / sqrt(taut + taua);
  const double tauC = rhoi * sqrt(taua + taud) / (G[0] + G[4] + G[8]);
  double kdcphi = 0.0;
  if (fabs(P.lsc_kdc) > 0.0) kdcphi = P.lsc_kdc * fabs(res) / sqrt(gg);
  // e3CFL
  double cf = 0.0;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) cf = cf + uadv[i] * G[i + 3 * j] * uadv[j];
  cf = P.delt * sqrt(cf);
  const double sc = uadv[0] * ga[0] + uadv[1] * ga[1] + uadv[2] * ga[2];
  const double dw = detj * gwq;  // hex weights are 1; padded tet and prism points 0
  if (valid && a == 0 && q < NSHL && detj < 0.0) atomicAdd(negdet, 1);

  s_sh[w][q][a][0] = Na;
  s_sh[w][q][a][1] = ga[0];
  s_sh[w][q][a][2] = ga[1];
  s_sh[w][q][a][3] = ga[2];
  s_sh[w][q][a][4] = sc;
  if (a == 0) {
    double *sq = s_q[w][q];
    sq[0] = rhoi;
    sq[1] = mui;
    sq[2] = tauM;
    sq[3] = tauC;
    sq[4] = lam;
    sq[5] = tauLS;
    sq[6] = vdot;
    sq[7] = phii;
    sq[8] = dmdphi;
    sq[9] = kdcphi;
    sq[10] = dphi[0];
    sq[11] = dphi[1];
    sq[12] = dphi[2];
    sq[13] = dw;
  }

  // RHS of node a at Gauss point q (:496-521), then summed over q
  double ru[3] = {0.0, 0.0, 0.0}, rp = 0.0, rl = 0.0;
  if (do_vec && fluid) {
    const double cr = 1 / P.rhoa - 1 / P.rhow;
    double up[3], t1[3], S[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) up[i] = -tauM * rLi[i];
#pragma unroll
    for (int i = 0; i < 3; ++i)
      t1[i] = rhoi * aci[i] + rhoi * uadv[0] * du[i + 0] + rhoi * uadv[1] * du[i + 3] +
              rhoi * uadv[2] * du[i + 6] - fi[i];
    // the viscous sums are scalars taken off every component, as the
    // reference writes them (:500-502)
#pragma unroll
    for (int k = 0; k < 3; ++k)
      S[k] = ga[0] * (du[0 + 3 * k] + du[k + 0]) + ga[1] * (du[1 + 3 * k] + du[k + 3]) +
             ga[2] * (du[2 + 3 * k] + du[k + 6]);
#pragma unroll
    for (int i = 0; i < 3; ++i)
      ru[i] = ru[i] - Na * t1[i] * dw - ga[i] * (-pri + tauC * divu) * dw +
              ga[i] * tauC * vdot * dw - ga[i] * lam * divu * dw - mui * S[0] * dw -
              mui * S[1] * dw - mui * S[2] * dw + rhoi * sc * up[i] * dw;
    rp = rp - Na * (divu - vdot) * dw + (ga[0] * up[0] + ga[1] * up[1] + ga[2] * up[2]) * dw;
    rl = rl - (Na + tauLS * sc) * res * dw -
         kdcphi * (ga[0] * dphi[0] + ga[1] * dphi[1] + ga[2] * dphi[2]) * dw +
         Na * dmdphi * cr * tauC * (divu - vdot) * dw;
  }
#pragma unroll
  for (int m = 8; m < 64; m <<= 1) {
#pragma unroll
    for (int i = 0; i < 3; ++i) ru[i] += __shfl_xor(ru[i], m);
    rp += __shfl_xor(rp, m);
    rl += __shfl_xor(rl, m);
    cf = fmax(cf, __shfl_xor(cf, m));
  }
  cf = wmax8(cf);
  if (valid && lane == 0) cfl_out[e] = cf;

  __syncthreads();

  // ---------------- phase B: lane = (a, b) ----------------
  const int pa_ = lane & 7;
  const int pb_ = lane >> 3;
  double K[9], Gv[3], D[3], Mv = 0.0, LS = 0.0, LSU[3], ULS[3], PLS = 0.0;
#pragma unroll
  for (int t = 0; t < 9; ++t) K[t] = 0.0;
#pragma unroll
  for (int t = 0; t < 3; ++t) Gv[t] = D[t] = LSU[t] = ULS[t] = 0.0;
  if (do_mat && fluid) {
    const double f1 = P.almi, f2 = P.alfi * P.delt * P.gami;
    const double cr = 1 / P.rhoa - 1 / P.rhow;
    for (int qq = 0; qq < NSHL; ++qq) {
      const double Nq = s_sh[w][qq][pa_][0];
      const double ga0 = s_sh[w][qq][pa_][1], ga1 = s_sh[w][qq][pa_][2],
                   ga2 = s_sh[w][qq][pa_][3], sa = s_sh[w][qq][pa_][4];
      const double Mb = s_sh[w][qq][pb_][0];
      const double gb0 = s_sh[w][qq][pb_][1], gb1 = s_sh[w][qq][pb_][2],
                   gb2 = s_sh[w][qq][pb_][3], sb = s_sh[w][qq][pb_][4];
      const double *sq = s_q[w][qq];
      const double rh = sq[0], mu = sq[1], tM = sq[2], tC = sq[3], lm = sq[4], tL = sq[5],
                   vd = sq[6], ph = sq[7], dm = sq[8], kp = sq[9], dj = sq[13];
      const double dph[3] = {sq[10], sq[11], sq[12]};
      const double gav[3] = {ga0, ga1, ga2}, gbv[3] = {gb0, gb1, gb2};
      const double gab = ga0 * gb0 + ga1 * gb1 + ga2 * gb2;
#pragma unroll
      for (int i = 0; i < 3; ++i) {
#pragma unroll
        for (int j = 0; j < 3; ++j)
          K[i * 3 + j] = K[i * 3 + j] + f2 * gav[j] * gbv[i] * mu * dj +
                         f2 * gav[i] * gbv[j] * (tC + lm) * dj;
        K[i * 3 + i] = K[i * 3 + i] + f1 * Nq * rh * Mb * dj + f2 * Nq * rh * sb * dj +
                       f2 * mu * gab * dj + f1 * rh * sa * tM * rh * Mb * dj +
                       f2 * rh * sa * tM * rh * sb * dj;
      }
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        Gv[i] = Gv[i] - f2 * gav[i] * Mb * dj + f2 * rh * sa * tM * gbv[i] * dj;
        D[i] = D[i] + f2 * Nq * gbv[i] * dj + gav[i] * tM * rh * f1 * Mb * dj +
               gav[i] * tM * rh * f2 * sb * dj;
      }
      Mv = Mv + f2 * tM * gab * dj;
      const double wa = Nq + tL * sa;
      LS = LS + wa * f1 * Mb * dj + wa * f2 * sb * dj;
      LS = LS + wa * f2 * Mb * vd * dj + wa * f2 * ph * cr * dm * Mb * dj -
           wa * f2 * dm * Mb / P.rhoa * dj;
      LS = LS + f2 * kp * gab * dj;
      LS = LS + f2 * dm * cr * Nq * tC * dm * cr * Mb * dj;
#pragma unroll
      for (int i = 0; i < 3; ++i) ULS[i] = ULS[i] + f2 * gav[i] * tC * dm * Mb * cr * dj;
      PLS = PLS - f2 * Nq * dm * Mb * cr * dj;
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        LSU[i] = LSU[i] + f2 * wa * Mb * dph[i] * dj;
        LSU[i] = LSU[i] - dm * cr * f2 * Nq * tC * gbv[i] * dj;
      }
    }
  }

  // ---------------- phase C: BCLhs_NS_conv + scatter ----------------
  if (valid && pa_ < NSHL && pb_ < NSHL) {
    const unsigned ma = ibc[s_node[w][pa_]], mb = ibc[s_node[w][pb_]];
    const bool diag = pa_ == pb_;
    if (do_mat) {
      double v[16], w8[8];
#pragma unroll
      for (int r = 0; r < 3; ++r) {
        const bool rz = (ma >> r) & 1u;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          double x = (rz || ((mb >> c) & 1u)) ? 0.0 : K[r * 3 + c];
          if (diag && r == c && rz) x = 1.0;
          v[4 * r + c] = x;
        }
        v[4 * r + 3] = (rz || ((mb >> 6) & 1u)) ? 0.0 : Gv[r];
        w8[2 * r] = ((mb >> 7) & 1u) ? 0.0 : ULS[r];
      }
      const bool pz = (ma >> 6) & 1u, fz = (ma >> 7) & 1u;
#pragma unroll
      for (int c = 0; c < 3; ++c) v[12 + c] = (pz || ((mb >> c) & 1u)) ? 0.0 : D[c];
      v[15] = pz ? (diag ? 1.0 : 0.0) : Mv;
      w8[6] = (pz || ((mb >> 7) & 1u)) ? 0.0 : PLS;
#pragma unroll
      for (int c = 0; c < 3; ++c)
        w8[2 * c + 1] = (((ma >> c) & 1u) || ((mb >> c) & 1u) || fz) ? 0.0 : LSU[c];
      w8[7] = (fz || ((mb >> 7) & 1u)) ? ((fz && diag) ? 1.0 : 0.0) : LS;
      const int k = emap[(size_t)e * 64 + lane];
      // the first element (in colour order) to reach the nonzero stores
      const bool firstk = (firstmask[e] >> lane) & 1ull;
      double2 *b16 = reinterpret_cast<double2 *>(A + (size_t)k * 16);
      double2 *b8 = reinterpret_cast<double2 *>(Apair + (size_t)k * 8);
#pragma unroll
      for (int t = 0; t < 8; ++t) {
        double2 o = firstk ? make_double2(0.0, 0.0) : b16[t];
        o.x += v[2 * t];
        o.y += v[2 * t + 1];
        b16[t] = o;
      }
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        double2 o = firstk ? make_double2(0.0, 0.0) : b8[t];
        o.x += w8[2 * t];
        o.y += w8[2 * t + 1];
        b8[t] = o;
      }
    }
    if (do_vec && lane < NSHL) {
      const unsigned m0 = ibc[na];
      double *rr = rhs + (size_t)na * 4;
#pragma unroll
      for (int c = 0; c < 3; ++c) rr[c] += ((m0 >> c) & 1u) ? 0.0 : ru[c];
      rr[3] += ((m0 >> 6) & 1u) ? 0.0 : rp;
      rhs[(size_t)4 * N + na] += ((m0 >> 7) & 1u) ? 0.0 : rl;
    }
  }
}

hipError_t launch_assemble_nsvof_hex(Ctx &c, int tensor_flag, int color) {
  const int first = c.color_off[color];
  const int count = c.color_off[color + 1] - first;
  if (count <= 0) return hipSuccess;
  const double *q = c.qprm;
  NsvofParams P;
  P.almi = c.prm[OFD_P_ALMI];
  P.alfi = c.prm[OFD_P_ALFI];
  P.gami = c.prm[OFD_P_GAMI];
  P.delt = c.prm[OFD_P_DELT];
  P.gx = c.prm[OFD_P_GX];
  P.gy = c.prm[OFD_P_GY];
  P.gz = c.prm[OFD_P_GZ];
  P.rhoa = q[OFD_Q_RHOA];
  P.rhow = q[OFD_Q_RHOW];
  P.mua = q[OFD_Q_MUA];
  P.muw = q[OFD_Q_MUW];
  P.rhos = q[OFD_Q_RHOS];
  P.mus = q[OFD_Q_MUS];
  P.Ts = q[OFD_Q_TS];
  P.c_cond = q[OFD_Q_CCOND];
  P.c_evap = q[OFD_Q_CEVAP];
  P.c_dmdot = q[OFD_Q_CDMDOT];
  P.lsc_kdc = q[OFD_Q_LSCKDC];
  P.fluid_id = (int)q[OFD_Q_FLUID_ID];
  const int blocks = (count + kWaves - 1) / kWaves;
  hipLaunchKernelGGL((c.nshl == 4   ? k_assemble_nsvof_el<4>
                      : c.nshl == 6 ? k_assemble_nsvof_el<6>
                                    : k_assemble_nsvof_el<8>),
                     dim3(blocks), dim3(64 * kWaves), 0, c.stream,
                     c.nelem, c.d_ien, c.d_elm_id, c.d_color_elems, first, count, c.d_emap,
                     c.d_first, c.d_noderec, c.d_ibc, P, (tensor_flag & OFD_ASSEMBLE_MAT) ? 1 : 0,
                     (tensor_flag & OFD_ASSEMBLE_VEC) ? 1 : 0, c.nnode, c.d_AV,
                     c.d_AV + (size_t)16 * c.icnt, c.d_rhsV, c.d_cfl, c.d_negdet);
  return hipGetLastError();
}

}  // namespace ofd
