// Temperature equation of the quenching solver on gfx950, hex elements.
//
// Restates IntElmAss_Tem_Quenching_STAB (solveQuenching.f90:579-882): the
// SUPG-stabilised energy equation with the evaporation / condensation
// source and its temperature derivative in the LHS (:742-752, :788-838),
// fluid elements (ELM_ID 102) with VOF-interpolated properties (prop_interp,
// e3int.f90:4-15) and solid elements with constant ones, tau from
// e3STAB_3D_TEM (e3STAB.f90:338-367), then BCLhs_tem (BCLhs_3D.f90:247-274;
// IBC column 6) and FillSparseMat_tem (FillSparseMat.f90:175-201).
//
// The layout is the NS kernel's: one 64-lane wavefront per element.
//   phase A  lane = (Gauss point q, node a): shape functions, Jacobian,
//            Gauss-point values (8-lane butterflies), properties, phase
//            change, tau; per-(q,a) data staged in LDS; the RHS of node a
//            reduced over q.
//   phase B  lane = (a, b): the element matrix entry accumulated over the 8
//            Gauss points in the reference's term order.
//   phase C  Dirichlet rows/columns of T, then the read-modify-write of
//            LHSTem at emap[e][a + 8 b]; one colour per launch, so no two
//            elements in flight share a node.
#include "ofd_internal.h"
#include "wave_dpp.h"
#include "element_shape.h"

namespace ofd {

namespace {

constexpr int kWaves = 4;
__device__ __forceinline__ double sum8(double v) { return wsum8(v); }

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

__device__ __forceinline__ double prop_interp(double a, double b, double phi) {
  const double pc = fmax(0.0, fmin(1.0, phi));
  return a * (1.0 - pc) + b * pc;
}

}  // namespace

struct TemParams {
  double almi, alfi, gami, delt;
  double rhoa, rhow, cpa, cpw, kappaa, kappaw, rhos, cps, kappas;
  double Ts, lh, c_cond, c_evap, c_dmdot, tem_kdc, tref;
  int fluid_id;
};

template <int S>
__global__ void __launch_bounds__(256)
k_assemble_tem_el(int E, const int32_t *__restrict__ ien, const int32_t *__restrict__ elm_id,
                   const int32_t *__restrict__ elems, int first, int count,
                   const int32_t *__restrict__ emap, const uint64_t *__restrict__ firstmask,
                   const double *__restrict__ rec,
                   const double *__restrict__ ks, const uint8_t *__restrict__ ibc,
                   TemParams P, int do_mat, int do_vec,
                   double *__restrict__ A, double *__restrict__ rhs,
                   int32_t *__restrict__ negdet) {
  __shared__ double s_sh[kWaves][8][8][5];  // [q][a]: N, dN/dx (3), shconv
  __shared__ double s_q[kWaves][8][5];      // [q]: rho cp, tau, dSe/dT, k (+k_dc), DetJ*w
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
  const int na = valid ? ien[(size_t)(a < S ? a : 0) * E + e] : 0;
  if (lane < 8) s_node[w][lane] = na;
  // node record: x+d (0-2), u (3-5), ac (6-8), um (9-11), p, phi, T, rT (12-15)
  const double2 *r2 = reinterpret_cast<const double2 *>(rec + (size_t)na * 16);
  const double2 v0 = r2[0], v1 = r2[1], v2 = r2[2], v4 = r2[4], v5 = r2[5], v6 = r2[6],
                v7 = r2[7];
  const double xd[3] = {v0.x, v0.y, v1.x};
  const double ua[3] = {v1.y, v2.x, v2.y};
  const double uma[3] = {v4.y, v5.x, v5.y};
  const double phia = v6.y, Ta = v7.x, rTa = v7.y;

  double Na, dl0, dl1, dl2, gwq;
  elem_shape<S>(q, a, Na, dl0, dl1, dl2, gwq);

  double J[9];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    J[i + 0] = sum8(xd[i] * dl0);
    J[i + 3] = sum8(xd[i] * dl1);
    J[i + 6] = sum8(xd[i] * dl2);
  }
  // get_inverse_3x3 (eval_shape.f90:85-112), column-major
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

  // Gauss-point values (e3int_qr, e3int_qr_grad)
  const double phii = seq8(phia * Na);
  const double Ti = seq8(Ta * Na);
  const double rTi = sum8(rTa * Na);
  double uadv[3], dT[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) uadv[i] = sum8(ua[i] * Na) - sum8(uma[i] * Na);
#pragma unroll
  for (int j = 0; j < 3; ++j) dT[j] = sum8(Ta * ga[j]);

  // phase change (solveQuenching.f90:742-752)
  const double phic = fmax(fmin(phii, 1.0), 0.0);
  double mdot, dmdT;
  if (Ti > P.Ts) {
    mdot = P.c_evap * (1 - phic) * P.rhow * (Ti - P.Ts) / P.Ts;
    dmdT = P.c_evap * (1 - phic) * P.rhow / P.Ts * P.c_dmdot;
  } else {
    mdot = P.c_cond * (phic)*P.rhoa * (Ti - P.Ts) / P.Ts;
    dmdT = P.c_cond * (phic)*P.rhoa / P.Ts * P.c_dmdot;
  }
  // properties (:756-769)
  double rhoi, cpi, hki, rhocpi;
  if (fluid) {
    rhoi = prop_interp(P.rhow, P.rhoa, phic);
    cpi = prop_interp(P.cpw, P.cpa, phic);
    hki = prop_interp(P.kappaw, P.kappaa, phic);
    rhocpi = prop_interp(P.rhow * P.cpw, P.rhoa * P.cpa, phic);
  } else {
    rhoi = P.rhos;
    cpi = P.cps;
    hki = P.kappas;
    rhocpi = P.rhos * P.cps;
  }
  // per-node conductivity scale (an extension; none: 1)
  if (ks) hki = hki * sum8(ks[na] * Na);
  const double Se = ((P.cpw - P.cpa) * (Ti - P.Ts) - P.lh) * mdot;
  const double udT = uadv[0] * dT[0] + uadv[1] * dT[1] + uadv[2] * dT[2];
  // the second-derivative term of the residual is zero for these elements
  const double res = rhocpi * (rTi + udT) - Se;
  // e3STAB_3D_TEM
  double tt = 0.0, gg = 0.0;
#pragma unroll
  for (int j = 0; j < 3; ++j)
#pragma unroll
    for (int i = 0; i < 3; ++i) tt = tt + uadv[i] * G[i + 3 * j] * uadv[j];
  tt = tt + 4.0 / (P.delt * P.delt);
  const double kk = hki / (rhoi * cpi);
#pragma unroll
  for (int t = 0; t < 9; ++t) gg = gg + G[t] * G[t];
  tt = tt + kk * kk * 12.0 * gg;
  const double tau = 1.0 / (rhoi * cpi * sqrt(tt));
  double kdc = 0.0;
  if (fabs(P.tem_kdc) > 0.0 && fluid) kdc = P.tem_kdc * fabs(res) / sqrt(gg) / P.tref;
  const double dSedT = (P.cpw - P.cpa) * mdot + ((P.cpw - P.cpa) * (Ti - P.Ts) - P.lh) * dmdT;
  const double sc = ga[0] * uadv[0] + ga[1] * uadv[1] + ga[2] * uadv[2];
  const double hk = fluid ? hki + kdc : hki;
  const double dw = detj * gwq;  // hex weights are 1; padded tet points 0
  if (valid && a == 0 && (S == 8 || q < 4) && detj < 0.0) atomicAdd(negdet, 1);

  s_sh[w][q][a][0] = Na;
  s_sh[w][q][a][1] = ga[0];
  s_sh[w][q][a][2] = ga[1];
  s_sh[w][q][a][3] = ga[2];
  s_sh[w][q][a][4] = sc;
  if (a == 0) {
    s_q[w][q][0] = rhocpi;
    s_q[w][q][1] = tau;
    s_q[w][q][2] = dSedT;
    s_q[w][q][3] = hk;
    s_q[w][q][4] = dw;
  }

  // RHS of node a at Gauss point q (:841-855), then summed over q
  double rt = 0.0;
  if (do_vec) {
    if (fluid) {
      rt = rt - Na * rhocpi * (rTi + udT) * dw;
      rt = rt + Na * Se * dw;
      rt = rt - rhocpi * tau * sc * res * dw;
    } else {
      rt = rt - Na * rhocpi * rTi * dw;
    }
    rt = rt - hk * ga[0] * dT[0] * dw;
    rt = rt - hk * ga[1] * dT[1] * dw;
    rt = rt - hk * ga[2] * dT[2] * dw;
  }
  rt += __shfl_xor(rt, 8);
  rt += __shfl_xor(rt, 16);
  rt += __shfl_xor(rt, 32);

  __syncthreads();

  // ---------------- phase B: lane = (a, b) ----------------
  const int pa = lane & 7;
  const int pb = lane >> 3;
  double acc = 0.0;
  if (do_mat) {
    const double f1 = P.almi, f2 = P.alfi * P.gami * P.delt;
#pragma unroll 2
    for (int qq = 0; qq < (S == 8 ? 8 : 4); ++qq) {
      const double Nq = s_sh[w][qq][pa][0], g0 = s_sh[w][qq][pa][1],
                   g1 = s_sh[w][qq][pa][2], g2 = s_sh[w][qq][pa][3],
                   sca = s_sh[w][qq][pa][4];
      const double Mb = s_sh[w][qq][pb][0], h0 = s_sh[w][qq][pb][1],
                   h1 = s_sh[w][qq][pb][2], h2 = s_sh[w][qq][pb][3],
                   scb = s_sh[w][qq][pb][4];
      const double rc = s_q[w][qq][0], tq = s_q[w][qq][1], dS = s_q[w][qq][2],
                   kq = s_q[w][qq][3], dj = s_q[w][qq][4];
      const double gab = g0 * h0 + g1 * h1 + g2 * h2;
      if (fluid) {
        const double lb = f1 * Mb + f2 * scb;
        acc = acc + Nq * rc * lb * dj - Nq * f2 * dS * Mb * dj + tq * rc * sca * rc * lb * dj -
              tq * rc * sca * f2 * dS * Mb * dj + f2 * kq * gab * dj;
      } else {
        acc = acc + Nq * rc * f1 * Mb * dj + f2 * kq * gab * dj;
      }
    }
  }

  // ---------------- phase C: BCLhs_tem + scatter ----------------
  if (valid && pa < S && pb < S) {
    const unsigned ca = (ibc[s_node[w][pa]] >> 5) & 1u;
    const unsigned cb = (ibc[s_node[w][pb]] >> 5) & 1u;
    if (do_mat) {
      double v = (ca || cb) ? 0.0 : acc;
      if (pa == pb && ca) v = 1.0;
      const int k = emap[(size_t)e * 64 + lane];
      // the first element (in colour order) to reach the nonzero stores
      A[k] = ((firstmask[e] >> lane) & 1ull) ? v : A[k] + v;
    }
    if (do_vec && lane < S) rhs[na] += ca ? 0.0 : rt;
  }
}

hipError_t launch_assemble_tem_hex(Ctx &c, int tensor_flag, int color) {
  const int first = c.color_off[color];
  const int count = c.color_off[color + 1] - first;
  if (count <= 0) return hipSuccess;
  const double *q = c.qprm;
  TemParams P;
  P.almi = c.prm[OFD_P_ALMI];
  P.alfi = c.prm[OFD_P_ALFI];
  P.gami = c.prm[OFD_P_GAMI];
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
  P.c_dmdot = q[OFD_Q_CDMDOT];
  P.tem_kdc = q[OFD_Q_TEMKDC];
  P.tref = q[OFD_Q_TREF];
  P.fluid_id = (int)q[OFD_Q_FLUID_ID];
  const int blocks = (count + kWaves - 1) / kWaves;
  hipLaunchKernelGGL((c.nshl == 4 ? k_assemble_tem_el<4> : k_assemble_tem_el<8>), dim3(blocks),
                     dim3(64 * kWaves), 0, c.stream,
                     c.nelem, c.d_ien, c.d_elm_id, c.d_color_elems, first, count, c.d_emap,
                     c.d_first, c.d_noderec, c.ks_on ? c.d_field[OFD_KSCALE] : nullptr,
                     c.d_ibc, P,
                     (tensor_flag & OFD_ASSEMBLE_MAT) ? 1 : 0,
                     (tensor_flag & OFD_ASSEMBLE_VEC) ? 1 : 0, c.d_AT, c.d_rhsT, c.d_negdet);
  return hipGetLastError();
}

}  // namespace ofd
