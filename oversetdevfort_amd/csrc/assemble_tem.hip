// Temperature equation of the quenching solver on gfx950, hexes and tets.
//
// Restates IntElmAss_Tem_Quenching_STAB (solveQuenching.f90:579-882): the
// SUPG-stabilised energy equation with the evaporation / condensation
// source and its temperature derivative in the LHS (:742-752, :788-838),
// fluid elements (ELM_ID 102) with VOF-interpolated properties (prop_interp,
// e3int.f90:4-15) and solid elements with constant ones, tau from
// e3STAB_3D_TEM (e3STAB.f90:338-367), then BCLhs_tem (BCLhs_3D.f90:247-274;
// IBC column 6) and FillSparseMat_tem (FillSparseMat.f90:175-201).
//
// Three kernels (S = 8 or 4 nodes and Gauss points), as the NS assembly:
//   k_tem_points<S>  one launch over all elements, 64 / S per wavefront,
//                    lane = (element, q): Jacobian, Gauss-point values,
//                    properties, phase change, tau once per point; the sums
//                    over the S nodes are register sums in the association
//                    of the lane butterfly, the VOF value and T (compared
//                    against thresholds) in node order.  It writes each
//                    element's point data at its position in colour order
//                    (J^-1, the advective velocity and five point factors per
//                    point, the RHS per node summed over the points, the node
//                    numbers).
//   k_tem_pairs<S>   one launch per colour, 64 / S^2 elements per wavefront:
//                    lane = (element, q, a) forms dN_a/dx and the streamline
//                    derivative from the point data, then lane = (element, a,
//                    b) the entry (a, b) over the S points, BCLhs_tem and the
//                    first-touch read-modify-write of LHSTem.
//   k_tem_rhs_gather one launch, thread per node: the node's contributions in
//                    colour order, the order of a coloured scatter.
// The one-wavefront-per-element kernel they replace (lane = (q, a), every
// Gauss-point quantity in S copies, the interpolations as lane reductions)
// took 0.304 / 0.207 ms (VEC+MAT / VEC) on the 48^3 tank.
#include "ofd_internal.h"
#include "wave_dpp.h"
#include "element_shape.h"

namespace ofd {

namespace {

constexpr int kWaves = 4;
constexpr int kTRec = 13;  // staged node record: x + d, u, um, phi, T, rT, ks

__device__ __forceinline__ double prop_interp(double a, double b, double phi) {
  const double pc = fmax(0.0, fmin(1.0, phi));
  return a * (1.0 - pc) + b * pc;
}

// point data of one element, in doubles: per point [18] (J^-1 column-major,
// uadv, rho cp, tau, dSe/dT, k (+ k_dc), DetJ w, 0); the RHS per node; the
// node numbers and the fluid flag as int32
template <int S>
struct TemPts {
  static constexpr int Q = 18, R = 18 * S, N = 18 * S + S;
  static constexpr int stride = S == 8 ? 160 : S == 6 ? 120 : 80;
};
static_assert(TemPts<8>::N + 5 <= TemPts<8>::stride && TemPts<8>::stride <= 192, "hex layout");
static_assert(TemPts<6>::N + 4 <= TemPts<6>::stride && TemPts<6>::stride <= 136,
              "prism layout");
static_assert(TemPts<4>::N + 3 <= TemPts<4>::stride && TemPts<4>::stride <= 96, "tet layout");

}  // namespace

struct TemParams {
  double almi, alfi, gami, delt;
  double rhoa, rhow, cpa, cpw, kappaa, kappaw, rhos, cps, kappas;
  double Ts, lh, c_cond, c_evap, c_dmdot, tem_kdc, tref;
  int fluid_id;
};

template <int S>
__global__ void __launch_bounds__(64)
k_tem_points(int E, const int32_t *__restrict__ ien, const int32_t *__restrict__ elm_id,
             const int32_t *__restrict__ cpos, const double *__restrict__ rec,
             const double *__restrict__ ks, TemParams P, int do_mat, int do_vec,
             double *__restrict__ pts, int32_t *__restrict__ negdet) {
  using L = TemPts<S>;
  constexpr int PL = point_lanes<S>();  // lanes per element (prisms: 6 points of 8)
  __shared__ __align__(16) double s_rec[64 * kTRec];
  const int lane = threadIdx.x;
  const int el = lane / PL;  // 64 / PL elements per wavefront
  const int q = lane % PL;   // Gauss point, and node a = q for the loads
  const int e = blockIdx.x * (64 / PL) + el;
  const bool valid = e < E;
  const bool fluid = valid && elm_id[e] == P.fluid_id;
  double *po = pts + (size_t)(valid ? cpos[e] : 0) * L::stride;
  if (q < S) {
    // node record: x+d (0-2), u (3-5), ac (6-8), um (9-11), p, phi, T, rT (12-15)
    const int n = valid ? ien[(size_t)q * E + e] : 0;
    const double2 *r2 = reinterpret_cast<const double2 *>(rec + (size_t)n * 16);
    const double2 v0 = r2[0], v1 = r2[1], v2 = r2[2], v4 = r2[4], v5 = r2[5], v6 = r2[6],
                  v7 = r2[7];
    double *sr = s_rec + (el * S + q) * kTRec;
    sr[0] = v0.x;
    sr[1] = v0.y;
    sr[2] = v1.x;
    sr[3] = v1.y;
    sr[4] = v2.x;
    sr[5] = v2.y;
    sr[6] = v4.y;
    sr[7] = v5.x;
    sr[8] = v5.y;
    sr[9] = v6.y;
    sr[10] = v7.x;
    sr[11] = v7.y;
    sr[12] = ks ? ks[n] : 1.0;
    if (valid && do_mat) {
      int32_t *pn = reinterpret_cast<int32_t *>(po + L::N);
      pn[q] = n;
      if (q == 0) pn[S] = fluid ? 1 : 0;
    }
  }
  __syncthreads();
  const double *re = s_rec + el * S * kTRec;

  double Nq[S];
#pragma unroll
  for (int a = 0; a < S; ++a) Nq[a] = shape_N<S>(q, a);
  double J[9];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    double p0[S], p1[S], p2[S];
#pragma unroll
    for (int a = 0; a < S; ++a) {
      double Na, d0, d1, d2, gw;
      elem_shape<S>(q, a, Na, d0, d1, d2, gw);
      const double x = re[a * kTRec + i];
      p0[a] = x * d0;
      p1[a] = x * d1;
      p2[a] = x * d2;
    }
    J[i + 0] = tree_sum<S>(p0);
    J[i + 3] = tree_sum<S>(p1);
    J[i + 6] = tree_sum<S>(p2);
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
  // dN_a/dx at this point, formed where it is used (the same expression each
  // time, so the same bits) rather than kept for all S nodes
  auto grad = [&](int a, double g[3]) {
    double Na, d0, d1, d2, gw;
    elem_shape<S>(q, a, Na, d0, d1, d2, gw);
#pragma unroll
    for (int j = 0; j < 3; ++j) g[j] = d0 * Ji[0 + 3 * j] + d1 * Ji[1 + 3 * j] + d2 * Ji[2 + 3 * j];
  };
  double gwq;
  {
    double Na, d0, d1, d2;
    elem_shape<S>(q, 0, Na, d0, d1, d2, gwq);
  }
  double G[9];
#pragma unroll
  for (int j = 0; j < 3; ++j)
#pragma unroll
    for (int i = 0; i < 3; ++i)
      G[i + 3 * j] = Ji[0 + 3 * i] * Ji[0 + 3 * j] + Ji[1 + 3 * i] * Ji[1 + 3 * j] +
                     Ji[2 + 3 * i] * Ji[2 + 3 * j];

  // Gauss-point values (e3int_qr, e3int_qr_grad); the VOF value and T in
  // node order, since they are compared against thresholds
  double phii, Ti, rTi, ksi, uadv[3], dT[3];
  {
    double pf[S], pt[S], pr[S], pk[S];
#pragma unroll
    for (int a = 0; a < S; ++a) {
      pf[a] = re[a * kTRec + 9] * Nq[a];
      pt[a] = re[a * kTRec + 10] * Nq[a];
      pr[a] = re[a * kTRec + 11] * Nq[a];
      pk[a] = re[a * kTRec + 12] * Nq[a];
    }
    phii = seq_sum<S>(pf);
    Ti = seq_sum<S>(pt);
    rTi = tree_sum<S>(pr);
    ksi = tree_sum<S>(pk);
  }
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    double pu[S], pm[S];
#pragma unroll
    for (int a = 0; a < S; ++a) {
      pu[a] = re[a * kTRec + 3 + i] * Nq[a];
      pm[a] = re[a * kTRec + 6 + i] * Nq[a];
    }
    uadv[i] = tree_sum<S>(pu) - tree_sum<S>(pm);
  }
  {
    double pd[3][S];
#pragma unroll
    for (int a = 0; a < S; ++a) {
      double g[3];
      grad(a, g);
#pragma unroll
      for (int j = 0; j < 3; ++j) pd[j][a] = re[a * kTRec + 10] * g[j];
    }
#pragma unroll
    for (int j = 0; j < 3; ++j) dT[j] = tree_sum<S>(pd[j]);
  }

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
  if (ks) hki = hki * ksi;
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
  const double hk = fluid ? hki + kdc : hki;
  const double dw = detj * gwq;  // hexes: unit weights
  if (valid && q < S && detj < 0.0) atomicAdd(negdet, 1);
  if (valid && do_mat && q < S) {
    double2 *pq = reinterpret_cast<double2 *>(po + q * L::Q);
    pq[0] = make_double2(Ji[0], Ji[1]);
    pq[1] = make_double2(Ji[2], Ji[3]);
    pq[2] = make_double2(Ji[4], Ji[5]);
    pq[3] = make_double2(Ji[6], Ji[7]);
    pq[4] = make_double2(Ji[8], uadv[0]);
    pq[5] = make_double2(uadv[1], uadv[2]);
    pq[6] = make_double2(rhocpi, tau);
    pq[7] = make_double2(dSedT, hk);
    pq[8] = make_double2(dw, 0.0);
  }

  // RHS of node a at this point (:841-855), summed over the element's point
  // lanes in the butterfly association; lane q keeps node a = q
  if (do_vec) {
    double own = 0.0;
#pragma unroll
    for (int a = 0; a < S; ++a) {
      const double Na = Nq[a];
      double ga[3];
      grad(a, ga);
      double rt = 0.0;
      if (fluid) {
        const double sc = ga[0] * uadv[0] + ga[1] * uadv[1] + ga[2] * uadv[2];
        rt = rt - Na * rhocpi * (rTi + udT) * dw;
        rt = rt + Na * Se * dw;
        rt = rt - rhocpi * tau * sc * res * dw;
      } else {
        rt = rt - Na * rhocpi * rTi * dw;
      }
      rt = rt - hk * ga[0] * dT[0] * dw;
      rt = rt - hk * ga[1] * dT[1] * dw;
      rt = rt - hk * ga[2] * dT[2] * dw;
      const double sm = point_sum<S>(rt);
      if (q == a) own = sm;
    }
    if (valid && q < S) po[L::R + q] = own;
  }
}

// One colour's elements at positions first .. first + count - 1 of the
// colour order, 64 / S^2 per wavefront (one hex or prism, four tets); the
// point data, the scatter map and the first-touch masks are indexed by
// position.
template <int S>
__global__ void __launch_bounds__(256)
k_tem_pairs(int first, int count, const int32_t *__restrict__ emap_c,
            const uint64_t *__restrict__ first_c, const double *__restrict__ pts,
            const uint8_t *__restrict__ ibc, double f1, double f2, double *__restrict__ A) {
  using L = TemPts<S>;
  constexpr int SS = S * S;
  constexpr int EW = 64 / SS;  // elements per wavefront
  __shared__ __align__(16) double s_g[kWaves][64 * 4];      // [slot][q][a]: dN/dx, streamline
  __shared__ __align__(16) double s_f[kWaves][EW * S * 6];  // [slot][q]: 5 point factors
  __shared__ double s_n[kWaves][SS];                         // [q][a]: N_a at point q
  __shared__ uint32_t s_m[kWaves][16];                       // [slot][a]: T Dirichlet flag
  __shared__ int32_t s_fl[kWaves][4];                        // [slot]: fluid element
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const int t0 = first + (blockIdx.x * kWaves + w) * EW;
  if (t0 >= first + count) return;  // whole wavefronts; no block barrier below
  const int nv = (first + count - t0 < EW) ? first + count - t0 : EW;  // live slots
  // lane = (slot, a, b): its nonzero, first-touch flag and old value
  const int sl = lane / SS, rr = lane % SS;
  const int pa = rr % S, pb = rr / S;
  const bool live = sl < nv;
  const int tl = t0 + (live ? sl : 0);
  const int k = emap_c[(size_t)tl * 64 + pa + 8 * pb];
  const bool fst = (first_c[tl] >> (pa + 8 * pb)) & 1ull;
  const double old = (fst || !live) ? 0.0 : A[k];
  if (lane < EW * S) {
    const int s2 = lane / S, x2 = lane % S;
    if (s2 < nv) {
      const double *pe = pts + (size_t)(t0 + s2) * L::stride;
      const int32_t *pn = reinterpret_cast<const int32_t *>(pe + L::N);
      s_m[w][lane] = (ibc[pn[x2]] >> 5) & 1u;
      if (x2 == 0) s_fl[w][s2] = pn[S];
      const double *pf = pe + x2 * L::Q + 12;
      double *sf = s_f[w] + lane * 6;
#pragma unroll
      for (int u = 0; u < 5; ++u) sf[u] = pf[u];
    }
  }
  if (lane < SS) s_n[w][lane] = shape_N<S>(lane / S, lane % S);
  // lane = (slot, q, a): dN_a/dx and uadv . dN_a/dx at point q
  {
    const int s3 = lane / SS, r3 = lane % SS, q3 = r3 / S, a3 = r3 % S;
    if (s3 < nv) {
      const double2 *pj =
          reinterpret_cast<const double2 *>(pts + (size_t)(t0 + s3) * L::stride + q3 * L::Q);
      const double2 j0 = pj[0], j1 = pj[1], j2 = pj[2], j3 = pj[3], j4 = pj[4], j5 = pj[5];
      const double Ji[9] = {j0.x, j0.y, j1.x, j1.y, j2.x, j2.y, j3.x, j3.y, j4.x};
      const double uadv[3] = {j4.y, j5.x, j5.y};
      double Na, d0, d1, d2, gw;
      elem_shape<S>(q3, a3, Na, d0, d1, d2, gw);
      double g[3];
#pragma unroll
      for (int j = 0; j < 3; ++j)
        g[j] = d0 * Ji[0 + 3 * j] + d1 * Ji[1 + 3 * j] + d2 * Ji[2 + 3 * j];
      const double sc = g[0] * uadv[0] + g[1] * uadv[1] + g[2] * uadv[2];
      double2 *sg = reinterpret_cast<double2 *>(s_g[w]) + lane * 2;
      sg[0] = make_double2(g[0], g[1]);
      sg[1] = make_double2(g[2], sc);
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  __builtin_amdgcn_wave_barrier();

  if (!live) return;
  // entry (a, b) over the points, the reference's term order (:788-838)
  double acc = 0.0;
  const bool fluid = s_fl[w][sl] != 0;
  const double *sg = s_g[w] + sl * SS * 4;  // [q][a][4]
  const double *sf = s_f[w] + sl * S * 6;   // [q][6]
  const double *sn = s_n[w];                // [q][a]
#pragma unroll 2
  for (int qq = 0; qq < S; ++qq) {
    const double *ra = sg + (qq * S + pa) * 4, *rb = sg + (qq * S + pb) * 4;
    const double g0 = ra[0], g1 = ra[1], g2 = ra[2], sca = ra[3];
    const double h0 = rb[0], h1 = rb[1], h2 = rb[2], scb = rb[3];
    const double Nq = sn[qq * S + pa], Mb = sn[qq * S + pb];
    const double rc = sf[qq * 6 + 0], tq = sf[qq * 6 + 1], dS = sf[qq * 6 + 2],
                 kq = sf[qq * 6 + 3], dj = sf[qq * 6 + 4];
    const double gab = g0 * h0 + g1 * h1 + g2 * h2;
    if (fluid) {
      const double lb = f1 * Mb + f2 * scb;
      acc = acc + Nq * rc * lb * dj - Nq * f2 * dS * Mb * dj + tq * rc * sca * rc * lb * dj -
            tq * rc * sca * f2 * dS * Mb * dj + f2 * kq * gab * dj;
    } else {
      acc = acc + Nq * rc * f1 * Mb * dj + f2 * kq * gab * dj;
    }
  }
  // BCLhs_tem: rows and columns of fixed-T nodes cleared, 1 on the diagonal
  const unsigned ca = s_m[w][sl * S + pa], cb = s_m[w][sl * S + pb];
  double v = (ca || cb) ? 0.0 : acc;
  if (pa == pb && ca) v = 1.0;
  // the first element (in colour order) to reach the nonzero stores
  A[k] = fst ? v : old + v;
}

// rhs[n] = sum of the node's element contributions in colour order
// (inc_off/inc: entries t * 8 + a), 0 for a fixed-T node
__global__ void __launch_bounds__(256)
k_tem_rhs_gather(int N, const int32_t *__restrict__ inc_off, const int32_t *__restrict__ inc,
                 const double *__restrict__ pts, int stride, int roff,
                 const uint8_t *__restrict__ ibc, double *__restrict__ rhs) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  const bool fixed = (ibc[n] >> 5) & 1u;
  double s = 0.0;
  for (int t = inc_off[n]; t < inc_off[n + 1]; ++t) {
    const int ea = inc[t];
    const double v = pts[(size_t)(ea >> 3) * stride + roff + (ea & 7)];
    s += fixed ? 0.0 : v;
  }
  rhs[n] = s;
}

// The volume loop of assembleQuenching's temperature field: the point
// kernel, the pair kernel per colour (MAT), the RHS gather (VEC; the caller
// zeroes the padded RHS).  c.d_pts holds pts_stride(nshl) doubles per
// element, at least TemPts' stride.
hipError_t launch_assemble_tem(Ctx &c, int tensor_flag) {
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
  const int do_mat = (tensor_flag & OFD_ASSEMBLE_MAT) ? 1 : 0;
  const int do_vec = (tensor_flag & OFD_ASSEMBLE_VEC) ? 1 : 0;
  const int S = c.nshl;
  const int epw = 64 / (S == 4 ? 4 : 8);                  // point kernel: per wavefront
  const int epb = 64 / (S * S) * kWaves;                  // pair kernel: per block
  const double *ks = c.ks_on ? c.d_field[OFD_KSCALE] : nullptr;
  hipLaunchKernelGGL((S == 8 ? k_tem_points<8> : S == 6 ? k_tem_points<6> : k_tem_points<4>),
                     dim3((c.nelem + epw - 1) / epw), dim3(64), 0, c.stream, c.nelem, c.d_ien,
                     c.d_elm_id, c.d_cpos, c.d_noderec, ks, P, do_mat, do_vec, c.d_pts,
                     c.d_negdet);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  if (do_mat) {
    const double f1 = P.almi, f2 = P.alfi * P.gami * P.delt;
    for (int col = 0; col + 1 < (int)c.color_off.size(); ++col) {
      const int first = c.color_off[col];
      const int count = c.color_off[col + 1] - first;
      if (count <= 0) continue;
      hipLaunchKernelGGL((S == 8 ? k_tem_pairs<8> : S == 6 ? k_tem_pairs<6> : k_tem_pairs<4>),
                         dim3((count + epb - 1) / epb), dim3(64 * kWaves), 0, c.stream, first,
                         count, c.d_emap_c, c.d_first_c, c.d_pts, c.d_ibc, f1, f2, c.d_AT);
      if ((e = hipGetLastError()) != hipSuccess) return e;
    }
  }
  if (do_vec) {
    hipLaunchKernelGGL(k_tem_rhs_gather, dim3((c.nnode + 255) / 256), dim3(256), 0, c.stream,
                       c.nnode, c.d_inc_off, c.d_inc, c.d_pts,
                       S == 8 ? TemPts<8>::stride : S == 6 ? TemPts<6>::stride
                                                         : TemPts<4>::stride,
                       S == 8 ? TemPts<8>::R : S == 6 ? TemPts<6>::R : TemPts<4>::R, c.d_ibc,
                       c.d_rhsT);
    e = hipGetLastError();
  }
  return e;
}

}  // namespace ofd
