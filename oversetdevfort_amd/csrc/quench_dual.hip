// Exact derivatives of the quenching solver's element residuals by forward-
// mode dual numbers on gfx950 (no reference counterpart), trilinear hexes,
// linear prisms and linear tets (S = 8, 6 or 4 nodes and Gauss points, the
// forward kernels' rules).
//
// Residuals differentiated, as the forward kernels assemble them:
//   NS + VOF  the RHSGu / RHSGp / RHSGls of IntElmAss_NSVOF_Quenching_STAB
//             (solveQuenching.f90:122-573; the forward restatement is
//             k_assemble_nsvof_el), at the alpha stage of u, ac, p, phi,
//             rphi and T, fluid elements (ELM_ID 102) only;
//   T         the RHSGTem of IntElmAss_Tem_Quenching_STAB (solveQuenching.f90:
//             579-882; forward k_tem_points), at the alpha stage of T, rT,
//             u and phi, with the per-node conductivity scale.
// Every lane of a wavefront evaluates its element's whole residual in
// (value, tangent) pairs, with the tangent seeded on one input of one node
// (its direction): the tangent of the residual is one column of the element
// Jacobian, exact to rounding, branches (phase change side, VOF clamp) taken
// where the values put them.  Nothing is hand-differentiated.
//
//   k_nsvof_dual<JAC>  lane j < 5 S: the Newton increment (d_u (3), d_p,
//                      d_phi) of node j / 5, i.e. the alpha-stage tangents
//                      ac = almi, u = alfi gami Delt for d_u; p = alfi gami
//                      Delt for d_p; rphi = almi, phi = alfi gami Delt for
//                      d_phi.  The column of J = -dR/d(increment), projected
//                      (rows and columns of the IBC columns 1-3, 7, 8
//                      cleared, 1 on the diagonal of a constrained dof once
//                      per element), is written into the NS + VOF matrix
//                      layout (the 4x4 block [K11 G; D1 M] and the phi pair
//                      group of each nonzero) by the first-touch read-modify-
//                      write of the colour launches.
//   k_nsvof_dual<VJP>  lanes j (and j + 64 < 10 S): the alpha-stage
//                      input j % 10 (u (3), ac (3), p, phi, rphi, T) of node
//                      j / 10; the lane's tangent contracted with lambda
//                      gives d(lambda . R)/d(input), added per node into
//                      out[input][N] (colours: no two elements of a launch
//                      share a node).
//   k_tem_dual         lane j < 7 S: input j % 7 (T, rT, u (3), phi, the
//                      conductivity scale) of node j / 7, contracted with
//                      lambda as the VJP mode; with the held-T mask it gives,
//                      through k_tem_jt_combine, J_T,bc^T lambda (the
//                      temperature adjoint's operator on tets).
// One wavefront per element: lanes (q, a) first stage the shape functions,
// dN/dx, the metric G and DetJ w of the 8 Gauss points in LDS (the forward
// kernel's expressions, so the same bits), and the element's node records.
// The Gauss-point values are then summed in node order (e3int_qr's order).
#include "ofd_internal.h"
#include "wave_dpp.h"
#include "element_shape.h"

namespace ofd {

namespace {

constexpr int kDW = 4;  // wavefronts (elements) per block

struct Dd {
  double v, d;
};
__device__ __forceinline__ Dd dd(double v) { return {v, 0.0}; }
__device__ __forceinline__ Dd operator+(Dd a, Dd b) { return {a.v + b.v, a.d + b.d}; }
__device__ __forceinline__ Dd operator-(Dd a, Dd b) { return {a.v - b.v, a.d - b.d}; }
__device__ __forceinline__ Dd operator-(Dd a) { return {-a.v, -a.d}; }
__device__ __forceinline__ Dd operator*(Dd a, Dd b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
__device__ __forceinline__ Dd operator+(Dd a, double b) { return {a.v + b, a.d}; }
__device__ __forceinline__ Dd operator+(double a, Dd b) { return {a + b.v, b.d}; }
__device__ __forceinline__ Dd operator-(Dd a, double b) { return {a.v - b, a.d}; }
__device__ __forceinline__ Dd operator-(double a, Dd b) { return {a - b.v, -b.d}; }
__device__ __forceinline__ Dd operator*(Dd a, double b) { return {a.v * b, a.d * b}; }
__device__ __forceinline__ Dd operator*(double a, Dd b) { return {a * b.v, a * b.d}; }
__device__ __forceinline__ Dd operator/(Dd a, Dd b) {
  const double r = a.v / b.v;
  return {r, (a.d - r * b.d) / b.v};
}
__device__ __forceinline__ Dd operator/(Dd a, double b) { return {a.v / b, a.d / b}; }
__device__ __forceinline__ Dd operator/(double a, Dd b) {
  const double r = a / b.v;
  return {r, -r * b.d / b.v};
}
__device__ __forceinline__ Dd dsqrt(Dd a) {
  const double s = sqrt(a.v);
  return {s, a.d / (2.0 * s)};
}
__device__ __forceinline__ Dd dabs(Dd a) { return a.v < 0.0 ? -a : a; }
// fmax(fmin(x, 1), 0) with the derivative of the branch the value takes
__device__ __forceinline__ Dd dclamp01(Dd x) {
  if (x.v > 1.0) return {1.0, 0.0};
  if (x.v < 0.0) return {0.0, 0.0};
  return x;
}
// prop_interp (e3int.f90:4-15): a (1 - pc) + b pc, pc = phi clamped to [0, 1]
__device__ __forceinline__ Dd dinterp(double a, double b, Dd phi) {
  const Dd pc = dclamp01(phi);
  return a * (1.0 - pc) + b * pc;
}

// Shape data of the wavefront's element at its 8 Gauss points, staged by
// lanes (q, a) with the forward kernels' expressions.
struct ShapeLds {
  double N[8][8];      // [q][a]
  double g[8][8][3];   // [q][a] dN_a/dx
  double G[8][9];      // [q] metric G = J^-T J^-1, column-major
  double dw[8];        // [q] DetJ * weight
};

template <int S>
__device__ __forceinline__ void stage_shape(ShapeLds &s, const double *xd, int q, int a) {
  double Na, dl0, dl1, dl2, gwq;
  elem_shape<S>(q, a, Na, dl0, dl1, dl2, gwq);
  double J[9];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    J[i + 0] = wsum8(xd[i] * dl0);
    J[i + 3] = wsum8(xd[i] * dl1);
    J[i + 6] = wsum8(xd[i] * dl2);
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
  s.N[q][a] = Na;
#pragma unroll
  for (int j = 0; j < 3; ++j)
    s.g[q][a][j] = dl0 * Ji[0 + 3 * j] + dl1 * Ji[1 + 3 * j] + dl2 * Ji[2 + 3 * j];
  if (a == 0) {
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
      for (int i = 0; i < 3; ++i)
        s.G[q][i + 3 * j] = Ji[0 + 3 * i] * Ji[0 + 3 * j] + Ji[1 + 3 * i] * Ji[1 + 3 * j] +
                            Ji[2 + 3 * i] * Ji[2 + 3 * j];
    s.dw[q] = detj * gwq;
  }
}

}  // namespace

struct NsvofDualParams {
  double almi, alfi, gami, delt, gx, gy, gz;
  double rhoa, rhow, mua, muw, rhos, mus, Ts, c_cond, c_evap, c_dmdot, lsc_kdc;
  int fluid_id;
};

enum { kDualJac = 0, kDualVjp = 1 };

// node record (alpha stage): x+d (0-2), u (3-5), ac (6-8), um (9-11), p (12),
// phi (13), T (14), rphi (15)
template <int S, int MODE>
__global__ void __launch_bounds__(64 * kDW)
k_nsvof_dual(int E, const int32_t *__restrict__ ien, const int32_t *__restrict__ elm_id,
             const int32_t *__restrict__ elems, int first, int count,
             const int32_t *__restrict__ emap, const uint64_t *__restrict__ firstmask,
             const double *__restrict__ rec, const uint8_t *__restrict__ ibc,
             NsvofDualParams P, int N, double *__restrict__ A, double *__restrict__ Apair,
             const double *__restrict__ lam4, const double *__restrict__ lamphi,
             double *__restrict__ out) {
  __shared__ ShapeLds s_shape[kDW];
  __shared__ double s_rec[kDW][8][16];
  __shared__ int32_t s_node[kDW][8];
  __shared__ double s_lam[kDW][8][5];
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const int idx = first + blockIdx.x * kDW + w;
  const bool valid = idx < first + count;
  const int e = valid ? elems[idx] : 0;
  const bool fluid = valid && elm_id[e] == P.fluid_id;
  ShapeLds &sh = s_shape[w];
  {
    const int q = lane >> 3, a = lane & 7;
    // node lanes a >= S (tets) repeat node 0 with zero shape functions
    const int na = valid ? ien[(size_t)(a < S ? a : 0) * E + e] : 0;
    const double *r = rec + (size_t)na * 16;
    double xd[3] = {r[0], r[1], r[2]};
    stage_shape<S>(sh, xd, q, a);
    if (q == 0) {
      s_node[w][a] = na;
#pragma unroll
      for (int t = 0; t < 16; ++t) s_rec[w][a][t] = r[t];
      if (MODE == kDualVjp) {
#pragma unroll
        for (int c = 0; c < 4; ++c) s_lam[w][a][c] = lam4[(size_t)na * 4 + c];
        s_lam[w][a][4] = lamphi[na];
      }
    }
  }
  __syncthreads();

  // the lane's directions and their seeds on its node's alpha-stage inputs:
  // tangents of u (3), ac (3), p, phi, rphi, T
  constexpr int kNdir = MODE == kDualJac ? 5 * S : 10 * S;
  constexpr int kPasses = (kNdir + 63) / 64;
  const double f2 = P.alfi * P.gami * P.delt;
  for (int pass = 0; pass < kPasses; ++pass) {
    const int dir = lane + 64 * pass;
    const int ndir = kNdir;
    const bool act = dir < ndir;
    const int nb = MODE == kDualJac ? dir / 5 : dir / 10;
    const int cb = MODE == kDualJac ? dir % 5 : dir % 10;
    const int b = act ? nb : 0;
    double tu[3] = {0.0, 0.0, 0.0}, tac[3] = {0.0, 0.0, 0.0};
    double tp = 0.0, tphi = 0.0, trphi = 0.0, tT = 0.0;
    if (act) {
      if (MODE == kDualJac) {
        if (cb < 3) {
          tu[cb] = f2;
          tac[cb] = P.almi;
        } else if (cb == 3) {
          tp = f2;
        } else {
          tphi = f2;
          trphi = P.almi;
        }
      } else {
        if (cb < 3) tu[cb] = 1.0;
        else if (cb < 6) tac[cb - 3] = 1.0;
        else if (cb == 6) tp = 1.0;
        else if (cb == 7) tphi = 1.0;
        else if (cb == 8) trphi = 1.0;
        else tT = 1.0;
      }
    }
    // JAC: tangents of the 8 x 5 residual entries; VJP: lambda . tangent
    double acc[S][5];
#pragma unroll
    for (int a = 0; a < S; ++a)
#pragma unroll
      for (int c = 0; c < 5; ++c) acc[a][c] = 0.0;
    double vj = 0.0;

    if (fluid) {
      const double cr = 1 / P.rhoa - 1 / P.rhow;
      for (int q = 0; q < S; ++q) {
        const double Nb = sh.N[q][b];
        const double gb[3] = {sh.g[q][b][0], sh.g[q][b][1], sh.g[q][b][2]};
        // Gauss-point values in node order, their tangents from node b
        double ui[3] = {0, 0, 0}, aci[3] = {0, 0, 0}, umi[3] = {0, 0, 0};
        double du[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, dpv[3] = {0, 0, 0}, dphv[3] = {0, 0, 0};
        double pri = 0.0, phii = 0.0, Tiv = 0.0, rphii = 0.0;
        for (int a = 0; a < S; ++a) {
          const double Na = sh.N[q][a];
          const double *r = s_rec[w][a];
#pragma unroll
          for (int i = 0; i < 3; ++i) {
            ui[i] = ui[i] + r[3 + i] * Na;
            aci[i] = aci[i] + r[6 + i] * Na;
            umi[i] = umi[i] + r[9 + i] * Na;
          }
          pri = pri + r[12] * Na;
          phii = phii + r[13] * Na;
          Tiv = Tiv + r[14] * Na;
          rphii = rphii + r[15] * Na;
#pragma unroll
          for (int j = 0; j < 3; ++j) {
            const double gj = sh.g[q][a][j];
#pragma unroll
            for (int i = 0; i < 3; ++i) du[i + 3 * j] = du[i + 3 * j] + r[3 + i] * gj;
            dpv[j] = dpv[j] + r[12] * gj;
            dphv[j] = dphv[j] + r[13] * gj;
          }
        }
        Dd u[3], ac[3], D[9], dp[3], dph[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          u[i] = {ui[i], Nb * tu[i]};
          ac[i] = {aci[i], Nb * tac[i]};
        }
#pragma unroll
        for (int j = 0; j < 3; ++j) {
#pragma unroll
          for (int i = 0; i < 3; ++i) D[i + 3 * j] = {du[i + 3 * j], gb[j] * tu[i]};
          dp[j] = {dpv[j], gb[j] * tp};
          dph[j] = {dphv[j], gb[j] * tphi};
        }
        const Dd pr = {pri, Nb * tp}, ph = {phii, Nb * tphi}, Ti = {Tiv, Nb * tT};
        const Dd rph = {rphii, Nb * trphi};
        const double *G = sh.G[q];
        const double dw = sh.dw[q];

        // phase change and properties (solveQuenching.f90:343-366)
        const Dd phic = dclamp01(ph);
        Dd mdot, dmdphi;
        if (Ti.v > P.Ts) {
          mdot = P.c_evap * (1.0 - phic) * P.rhow * (Ti - P.Ts) / P.Ts;
          dmdphi = -P.c_evap * P.rhow * (Ti - P.Ts) / P.Ts * P.c_dmdot;
        } else {
          mdot = P.c_cond * phic * P.rhoa * (Ti - P.Ts) / P.Ts;
          dmdphi = P.c_cond * P.rhoa * (Ti - P.Ts) / P.Ts * P.c_dmdot;
        }
        const Dd vdot = mdot / P.rhoa - mdot / P.rhow;
        const Dd rhoi = dinterp(P.rhow, P.rhoa, phic);
        const Dd mui = dinterp(P.muw, P.mua, phic);
        const Dd lam = -2.0 / 3.0 * mui;
        if (ph.v < 0.0 || ph.v > 1.0) dmdphi = dd(0.0);
        Dd uadv[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) uadv[i] = u[i] - dd(umi[i]);
        const Dd divu = D[0] + D[4] + D[8];
        Dd fi[3] = {P.gx * rhoi, P.gy * rhoi, P.gz * rhoi};
        Dd rLi[3];
#pragma unroll
        for (int i = 0; i < 3; ++i)
          rLi[i] = rhoi * ac[i] + rhoi * (D[i + 0] * uadv[0] + D[i + 3] * uadv[1] +
                                          D[i + 6] * uadv[2]) + dp[i] - fi[i];
        const Dd res = rph + (uadv[0] * dph[0] + uadv[1] * dph[1] + uadv[2] * dph[2]) +
                       ph * vdot - mdot / P.rhoa;
        // e3STAB_3D_NSVOF (e3STAB.f90:292-333)
        Dd taua = dd(0.0);
        double gg = 0.0;
#pragma unroll
        for (int j = 0; j < 3; ++j)
#pragma unroll
          for (int i = 0; i < 3; ++i) taua = taua + uadv[i] * G[i + 3 * j] * uadv[j];
#pragma unroll
        for (int t = 0; t < 9; ++t) gg = gg + G[t] * G[t];
        const Dd nu = mui / rhoi;
        const Dd taud = nu * nu * 12.0 * gg;
        const double taut = 4.0 / (P.delt * P.delt);
        const Dd tauM = 1.0 / (rhoi * dsqrt(taud + taua + taut));
        const Dd tauLS = 1.0 / dsqrt(taua + taut);
        const Dd tauC = rhoi * dsqrt(taua + taud) / (G[0] + G[4] + G[8]);
        Dd kdcphi = dd(0.0);
        if (fabs(P.lsc_kdc) > 0.0) kdcphi = P.lsc_kdc * dabs(res) / sqrt(gg);
        Dd up[3], t1[3], Sv[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) up[i] = -(tauM * rLi[i]);
#pragma unroll
        for (int i = 0; i < 3; ++i)
          t1[i] = rhoi * ac[i] + rhoi * uadv[0] * D[i + 0] + rhoi * uadv[1] * D[i + 3] +
                  rhoi * uadv[2] * D[i + 6] - fi[i];
        const Dd dv = divu - vdot;
        // the residual rows of the S nodes at this point (:496-521)
#pragma unroll
        for (int a = 0; a < S; ++a) {
          const double Na = sh.N[q][a];
          const double ga[3] = {sh.g[q][a][0], sh.g[q][a][1], sh.g[q][a][2]};
          const Dd sc = uadv[0] * ga[0] + uadv[1] * ga[1] + uadv[2] * ga[2];
#pragma unroll
          for (int k = 0; k < 3; ++k)
            Sv[k] = ga[0] * (D[0 + 3 * k] + D[k + 0]) + ga[1] * (D[1 + 3 * k] + D[k + 3]) +
                   ga[2] * (D[2 + 3 * k] + D[k + 6]);
          const Dd visc = mui * Sv[0] * dw + mui * Sv[1] * dw + mui * Sv[2] * dw;
          Dd r5[5];
#pragma unroll
          for (int i = 0; i < 3; ++i)
            r5[i] = -(Na * t1[i] * dw) - ga[i] * (tauC * divu - pr) * dw +
                    ga[i] * tauC * vdot * dw - ga[i] * lam * divu * dw - visc +
                    rhoi * sc * up[i] * dw;
          r5[3] = -(Na * dv * dw) + (ga[0] * up[0] + ga[1] * up[1] + ga[2] * up[2]) * dw;
          r5[4] = -((Na + tauLS * sc) * res * dw) -
                  kdcphi * (ga[0] * dph[0] + ga[1] * dph[1] + ga[2] * dph[2]) * dw +
                  Na * dmdphi * cr * tauC * dv * dw;
          if (MODE == kDualJac) {
#pragma unroll
            for (int c = 0; c < 5; ++c) acc[a][c] += r5[c].d;
          } else {
#pragma unroll
            for (int c = 0; c < 5; ++c) vj += s_lam[w][a][c] * r5[c].d;
          }
        }
      }
    }

    if (MODE == kDualJac) {
      // J = -dR/d(increment), projected, into column (b, cb) of the element's
      // blocks (row node a); the first element of the colour order stores
      if (valid && act) {
        const unsigned mb = ibc[s_node[w][b]];
        const int colbit = cb < 3 ? cb : (cb == 3 ? 6 : 7);
        const bool colz = (mb >> colbit) & 1u;
#pragma unroll
        for (int a = 0; a < S; ++a) {
          const unsigned ma = ibc[s_node[w][a]];
          const int pl = a + 8 * b;
          const int k = emap[(size_t)e * 64 + pl];
          const bool firstk = (firstmask[e] >> pl) & 1ull;
#pragma unroll
          for (int c = 0; c < 5; ++c) {
            // (phi row, p column) has no slot: the VOF residual does not
            // depend on p, so the entry is zero
            if (c == 4 && cb == 3) continue;
            const int rowbit = c < 3 ? c : (c == 3 ? 6 : 7);
            const bool rowz = (ma >> rowbit) & 1u;
            double v = (rowz || colz) ? 0.0 : -acc[a][c];
            if (rowz && a == b && c == cb) v = 1.0;
            double *dst;
            if (c < 4 && cb < 4) dst = A + (size_t)k * 16 + 4 * c + cb;
            else if (c < 3) dst = Apair + (size_t)k * 8 + 2 * c;          // ULS(c)
            else if (c == 3) dst = Apair + (size_t)k * 8 + 6;             // PLS
            else if (cb < 3) dst = Apair + (size_t)k * 8 + 2 * cb + 1;    // LSU(cb)
            else dst = Apair + (size_t)k * 8 + 7;                          // LS
            *dst = firstk ? v : *dst + v;
          }
        }
      }
    } else {
      if (valid && act) {
        double *o = out + (size_t)cb * N + s_node[w][b];
        *o = *o + vj;
      }
    }
  }
}

struct TemDualParams {
  double delt;
  double rhoa, rhow, cpa, cpw, kappaa, kappaw, rhos, cps, kappas;
  double Ts, lh, c_cond, c_evap, c_dmdot, tem_kdc, tref;
  int fluid_id;
};

// VJP of the temperature residual: lane j < 7 S = (node j / 7, input j % 7:
// T, rT, u (3), phi, conductivity scale); out[input][N] += d(lambda . R)/d.
// Node record as above with rT in slot 15 (the alpha stage of the
// temperature system); the scale from ks (none: 1).
template <int S>
__global__ void __launch_bounds__(64 * kDW)
k_tem_dual(int E, const int32_t *__restrict__ ien, const int32_t *__restrict__ elm_id,
           const int32_t *__restrict__ elems, int first, int count,
           const double *__restrict__ rec, const double *__restrict__ ks, TemDualParams P, int N,
           const double *__restrict__ lam, const uint8_t *__restrict__ tmask,
           double *__restrict__ out) {
  __shared__ ShapeLds s_shape[kDW];
  __shared__ double s_rec[kDW][8][12];  // x+d, u, um, phi, T, rT, ks
  __shared__ int32_t s_node[kDW][8];
  __shared__ double s_lam[kDW][8];
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const int idx = first + blockIdx.x * kDW + w;
  const bool valid = idx < first + count;
  const int e = valid ? elems[idx] : 0;
  const bool fluid = valid && elm_id[e] == P.fluid_id;
  ShapeLds &sh = s_shape[w];
  {
    const int q = lane >> 3, a = lane & 7;
    const int na = valid ? ien[(size_t)(a < S ? a : 0) * E + e] : 0;
    const double *r = rec + (size_t)na * 16;
    double xd[3] = {r[0], r[1], r[2]};
    stage_shape<S>(sh, xd, q, a);
    if (q == 0) {
      s_node[w][a] = na;
      double *sr = s_rec[w][a];
      sr[0] = r[3];
      sr[1] = r[4];
      sr[2] = r[5];
      sr[3] = r[9];
      sr[4] = r[10];
      sr[5] = r[11];
      sr[6] = r[13];
      sr[7] = r[14];
      sr[8] = r[15];
      sr[9] = ks ? ks[na] : 1.0;
      // tmask: lambda of the nodes whose temperature is held (IBC column 6)
      // taken as 0, the projection J_T,bc^T applies before the element
      s_lam[w][a] = (tmask && ((tmask[na] >> 5) & 1u)) ? 0.0 : lam[na];
    }
  }
  __syncthreads();
  const int dir = lane;
  const bool act = dir < 7 * S;
  const int b = act ? dir / 7 : 0;
  const int cb = dir % 7;
  double tT = 0.0, trT = 0.0, tu[3] = {0.0, 0.0, 0.0}, tphi = 0.0, tks = 0.0;
  if (act) {
    if (cb == 0) tT = 1.0;
    else if (cb == 1) trT = 1.0;
    else if (cb < 5) tu[cb - 2] = 1.0;
    else if (cb == 5) tphi = 1.0;
    else tks = 1.0;
  }
  double vj = 0.0;
  if (valid) {
    for (int q = 0; q < S; ++q) {
      const double Nb = sh.N[q][b];
      const double gb[3] = {sh.g[q][b][0], sh.g[q][b][1], sh.g[q][b][2]};
      double ui[3] = {0, 0, 0}, umi[3] = {0, 0, 0}, dTv[3] = {0, 0, 0};
      double phii = 0.0, Tiv = 0.0, rTi = 0.0, ksi = 0.0;
      for (int a = 0; a < S; ++a) {
        const double Na = sh.N[q][a];
        const double *r = s_rec[w][a];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          ui[i] = ui[i] + r[i] * Na;
          umi[i] = umi[i] + r[3 + i] * Na;
          dTv[i] = dTv[i] + r[7] * sh.g[q][a][i];
        }
        phii = phii + r[6] * Na;
        Tiv = Tiv + r[7] * Na;
        rTi = rTi + r[8] * Na;
        ksi = ksi + r[9] * Na;
      }
      Dd uadv[3], dT[3];
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        uadv[i] = {ui[i] - umi[i], Nb * tu[i]};
        dT[i] = {dTv[i], gb[i] * tT};
      }
      const Dd ph = {phii, Nb * tphi}, Ti = {Tiv, Nb * tT}, rT = {rTi, Nb * trT};
      const Dd ksv = {ksi, Nb * tks};
      const double *G = sh.G[q];
      const double dw = sh.dw[q];
      // phase change (solveQuenching.f90:742-752)
      const Dd phic = dclamp01(ph);
      Dd mdot;
      if (Ti.v > P.Ts) mdot = P.c_evap * (1.0 - phic) * P.rhow * (Ti - P.Ts) / P.Ts;
      else mdot = P.c_cond * phic * P.rhoa * (Ti - P.Ts) / P.Ts;
      // properties (:756-769) and the per-node conductivity scale
      Dd rhoi, cpi, hki, rhocpi;
      if (fluid) {
        rhoi = dinterp(P.rhow, P.rhoa, phic);
        cpi = dinterp(P.cpw, P.cpa, phic);
        hki = dinterp(P.kappaw, P.kappaa, phic);
        rhocpi = dinterp(P.rhow * P.cpw, P.rhoa * P.cpa, phic);
      } else {
        rhoi = dd(P.rhos);
        cpi = dd(P.cps);
        hki = dd(P.kappas);
        rhocpi = dd(P.rhos * P.cps);
      }
      if (ks) hki = hki * ksv;
      const Dd Se = ((P.cpw - P.cpa) * (Ti - P.Ts) - P.lh) * mdot;
      const Dd udT = uadv[0] * dT[0] + uadv[1] * dT[1] + uadv[2] * dT[2];
      const Dd res = rhocpi * (rT + udT) - Se;
      // e3STAB_3D_TEM (e3STAB.f90:338-367)
      Dd tt = dd(0.0);
      double gg = 0.0;
#pragma unroll
      for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int i = 0; i < 3; ++i) tt = tt + uadv[i] * G[i + 3 * j] * uadv[j];
      tt = tt + 4.0 / (P.delt * P.delt);
      const Dd kk = hki / (rhoi * cpi);
#pragma unroll
      for (int t = 0; t < 9; ++t) gg = gg + G[t] * G[t];
      tt = tt + kk * kk * 12.0 * gg;
      const Dd tau = 1.0 / (rhoi * cpi * dsqrt(tt));
      Dd kdc = dd(0.0);
      if (fabs(P.tem_kdc) > 0.0 && fluid) kdc = P.tem_kdc * dabs(res) / sqrt(gg) / P.tref;
      const Dd hk = fluid ? hki + kdc : hki;
      // the residual of node a at this point (:841-855)
#pragma unroll
      for (int a = 0; a < S; ++a) {
        const double Na = sh.N[q][a];
        const double ga[3] = {sh.g[q][a][0], sh.g[q][a][1], sh.g[q][a][2]};
        Dd rt = dd(0.0);
        if (fluid) {
          const Dd sc = ga[0] * uadv[0] + ga[1] * uadv[1] + ga[2] * uadv[2];
          rt = rt - Na * rhocpi * (rT + udT) * dw;
          rt = rt + Na * Se * dw;
          rt = rt - rhocpi * tau * sc * res * dw;
        } else {
          rt = rt - Na * rhocpi * rT * dw;
        }
        rt = rt - hk * (ga[0] * dT[0] + ga[1] * dT[1] + ga[2] * dT[2]) * dw;
        vj += s_lam[w][a] * rt.d;
      }
    }
  }
  if (valid && act) {
    double *o = out + (size_t)cb * N + s_node[w][b];
    *o = *o + vj;
  }
}

namespace {

NsvofDualParams nsvof_dual_params(const Ctx &c) {
  const double *q = c.qprm;
  NsvofDualParams P;
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
  return P;
}

}  // namespace

hipError_t launch_nsvof_jac(Ctx &c, int color) {
  const int first = c.color_off[color];
  const int count = c.color_off[color + 1] - first;
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL((c.nshl == 4   ? k_nsvof_dual<4, kDualJac>
                      : c.nshl == 6 ? k_nsvof_dual<6, kDualJac>
                                    : k_nsvof_dual<8, kDualJac>),
                     dim3((count + kDW - 1) / kDW), dim3(64 * kDW), 0,
                     c.stream, c.nelem, c.d_ien, c.d_elm_id, c.d_color_elems, first, count,
                     c.d_emap, c.d_first, c.d_noderec, c.d_ibc, nsvof_dual_params(c), c.nnode,
                     c.d_AV, c.d_AV + (size_t)16 * c.icnt, nullptr, nullptr, nullptr);
  return hipGetLastError();
}

hipError_t launch_nsvof_vjp(Ctx &c, int color, const double *lam4, const double *lamphi,
                            double *out) {
  const int first = c.color_off[color];
  const int count = c.color_off[color + 1] - first;
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL((c.nshl == 4   ? k_nsvof_dual<4, kDualVjp>
                      : c.nshl == 6 ? k_nsvof_dual<6, kDualVjp>
                                    : k_nsvof_dual<8, kDualVjp>),
                     dim3((count + kDW - 1) / kDW), dim3(64 * kDW), 0,
                     c.stream, c.nelem, c.d_ien, c.d_elm_id, c.d_color_elems, first, count,
                     c.d_emap, c.d_first, c.d_noderec, c.d_ibc, nsvof_dual_params(c), c.nnode,
                     nullptr, nullptr, lam4, lamphi, out);
  return hipGetLastError();
}

hipError_t launch_tem_dual_vjp(Ctx &c, int color, const double *ks, const double *lam,
                               double *out, const uint8_t *tmask) {
  const int first = c.color_off[color];
  const int count = c.color_off[color + 1] - first;
  if (count <= 0) return hipSuccess;
  const double *q = c.qprm;
  TemDualParams P;
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
  hipLaunchKernelGGL((c.nshl == 4 ? k_tem_dual<4> : c.nshl == 6 ? k_tem_dual<6> : k_tem_dual<8>),
                     dim3((count + kDW - 1) / kDW), dim3(64 * kDW), 0, c.stream,
                     c.nelem, c.d_ien, c.d_elm_id, c.d_color_elems, first, count, c.d_noderec,
                     ks, P, c.nnode, lam, tmask, out);
  return hipGetLastError();
}

// y = J_T,bc^T lambda from the raw VJP g [7][N] of the masked lambda: free
// rows -(alfi gami Delt g_T + almi g_rT) (the increments rTg += d, Tg +=
// gami Delt d), held rows lambda times the number of elements at the node
// (BCLhs_tem puts 1 on the diagonal once per element); padding rows 0
__global__ void k_tem_jt_combine(int N, int npad, const double *__restrict__ g,
                                 const double *__restrict__ lam,
                                 const uint8_t *__restrict__ ibc,
                                 const int32_t *__restrict__ inc_off, double cT, double almi,
                                 double *__restrict__ y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npad) return;
  if (i >= N) {
    y[i] = 0.0;
    return;
  }
  const bool held = (ibc[i] >> 5) & 1u;
  y[i] = held ? (double)(inc_off[i + 1] - inc_off[i]) * lam[i]
              : -(cT * g[i] + almi * g[(size_t)N + i]);
}

// the temperature adjoint's operators by dual numbers (any element the
// dual kernels take): raw gradients into out [7][N] (zeroed here), or
// y = J_T,bc^T lambda with out as scratch
hipError_t launch_vjp_tem_dual(Ctx &c, const double *lam, double *out) {
  hipError_t e = launch_zero(c, out, (size_t)c.nnode * 7);
  if (e != hipSuccess) return e;
  const double *ks = c.ks_on ? c.d_field[OFD_KSCALE] : nullptr;
  for (int col = 0; col + 1 < (int)c.color_off.size(); ++col) {
    e = launch_tem_dual_vjp(c, col, ks, lam, out, nullptr);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t launch_jt_tem_dual(Ctx &c, const double *lam, double *y, double *scratch) {
  hipError_t e = launch_zero(c, scratch, (size_t)c.nnode * 7);
  if (e != hipSuccess) return e;
  const double *ks = c.ks_on ? c.d_field[OFD_KSCALE] : nullptr;
  for (int col = 0; col + 1 < (int)c.color_off.size(); ++col) {
    e = launch_tem_dual_vjp(c, col, ks, lam, scratch, c.d_ibc);
    if (e != hipSuccess) return e;
  }
  const int npad = c.temT_n4 * 4;
  hipLaunchKernelGGL(k_tem_jt_combine, dim3((npad + 255) / 256), dim3(256), 0, c.stream,
                     c.nnode, npad, scratch, lam, c.d_ibc, c.d_inc_off,
                     c.prm[OFD_P_ALFI] * c.prm[OFD_P_GAMI] * c.prm[OFD_P_DELT],
                     c.prm[OFD_P_ALMI], y);
  return hipGetLastError();
}

// The stored NS + VOF matrix mirrored through tmap with each 5 x 5 block
// transposed, in the same two-array layout: block k of the result is block
// tmap[k] of A transposed.  The layout has no slot for a (phi row, p
// column) entry; the transpose's, PLS of the mirrored block, is dropped (the
// (p row, phi column) entry of the transpose, the (phi, p) entry of A, is
// zero: the VOF residual does not depend on p).  The result preconditions
// solves with A^T; k_spmv_v_t applies A^T exactly.
__global__ void k_nsvof_transpose(int64_t Z, const int32_t *__restrict__ tmap,
                                  const double *__restrict__ A, const double *__restrict__ Ap,
                                  double *__restrict__ At, double *__restrict__ Apt) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= Z) return;
  const int64_t m = tmap[k];
  const double *s = A + m * 16;
  const double *sp = Ap + m * 8;
  double *d = At + k * 16;
  double *dp = Apt + k * 8;
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int c = 0; c < 4; ++c) d[4 * r + c] = s[4 * c + r];
  // ULS(r) of the transpose (row u_r, col phi) is LSU(r) of A and back
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    dp[2 * r] = sp[2 * r + 1];
    dp[2 * r + 1] = sp[2 * r];
  }
  dp[6] = 0.0;
  dp[7] = sp[7];
}

hipError_t launch_nsvof_transpose(Ctx &c, double *At) {
  const int64_t Z = c.icnt;
  hipLaunchKernelGGL(k_nsvof_transpose, dim3((unsigned)((Z + 255) / 256)), dim3(256), 0,
                     c.stream, Z, c.d_tmap, c.d_AV, c.d_AV + 16 * Z, At, At + 16 * Z);
  return hipGetLastError();
}

// y = A^T x for the stored NS + VOF matrix (FP64), row i gathering the
// mirrored blocks of its row: y_i = sum_k B_k^T x_row[k], B_k = block
// tmap[k] as a 5 x 5 (u0 u1 u2 p phi) matrix.  Vectors: [N][4] then [N] phi,
// zero-padded to npad doubles.
__global__ void __launch_bounds__(256)
k_spmv_v_t(int N, int npad, const int32_t *__restrict__ col, const int32_t *__restrict__ row,
           const int32_t *__restrict__ tmap, const double *__restrict__ A,
           const double *__restrict__ Ap, const double *__restrict__ x,
           double *__restrict__ y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) {
    const int t = 5 * N + (i - N);
    if (t < npad) y[t] = 0.0;
    return;
  }
  double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  for (int k = col[i]; k < col[i + 1]; ++k) {
    const int j = row[k];
    const int64_t m = tmap[k];
    const double *B = A + m * 16;
    const double *P = Ap + m * 8;
    const double x0 = x[(size_t)j * 4], x1 = x[(size_t)j * 4 + 1], x2 = x[(size_t)j * 4 + 2],
                 x3 = x[(size_t)j * 4 + 3], x4 = x[(size_t)4 * N + j];
#pragma unroll
    for (int r = 0; r < 4; ++r)
      acc[r] += B[r] * x0 + B[4 + r] * x1 + B[8 + r] * x2 + B[12 + r] * x3;
#pragma unroll
    for (int r = 0; r < 3; ++r) acc[r] += P[2 * r + 1] * x4;
    acc[4] += P[0] * x0 + P[2] * x1 + P[4] * x2 + P[6] * x3 + P[7] * x4;
  }
#pragma unroll
  for (int r = 0; r < 4; ++r) y[(size_t)i * 4 + r] = acc[r];
  y[(size_t)4 * N + i] = acc[4];
}

hipError_t launch_spmv_v_t(Ctx &c, const double *x, double *y) {
  const int npad = c.nsvV_n4 * 4;
  const int threads = c.nnode + (npad - 5 * c.nnode);
  hipLaunchKernelGGL(k_spmv_v_t, dim3((threads + 255) / 256), dim3(256), 0, c.stream, c.nnode,
                     npad, c.d_col, c.d_row, c.d_tmap, c.d_AV, c.d_AV + (size_t)16 * c.icnt, x,
                     y);
  return hipGetLastError();
}

}  // namespace ofd
