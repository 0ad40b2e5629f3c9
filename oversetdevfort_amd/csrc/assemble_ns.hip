// VMS/SUPG Navier-Stokes element assembly on gfx950, hex elements.
//
// Restates IntElmAss_NS_overset (assembleNavSto.f90:76-336) with
// eval_shape_fem/linhexshl (eval_shape_FEM.f90:4-61,148-196), the 8-point
// rule (genGP_GW.f90:355-389), e3STAB_3D (e3STAB.f90:4-44), e3LHS_3D_fluid_Old
// (e3LHS.f90:4-161), e3Rhs_3D_fluid (e3RHS.f90:4-97) with the live call's
// zero fine-scale velocity, BCLhs_3D (BCLhs_3D.f90:4-122) and FillSparseMat_3D
// (FillSparseMat.f90:33-67).
//
// One 64-lane wavefront per element:
//   phase A  lane = (gauss point q, node a): shape functions, Jacobian,
//            Gauss-point interpolation (8-lane butterfly reductions), tau's;
//            shape data staged in LDS; RHS contributions reduced over q.
//   phase B  lane = (a, b) block pair: the 16 entries of the 4x4 block (a,b)
//            accumulated over the 8 Gauss points in registers.
//   phase C  Dirichlet projection as row/column masks (BCLhs_3D zeroes rows and
//            columns of constrained dofs and sets their diagonal to 1, in an
//            order that makes the masks exact), then the scatter of the 64
//            blocks to their nonzeros emap[e][a + 8 b], 8 lanes per 128-byte
//            block (a transpose through LDS): the first element (in colour
//            order) to reach a nonzero stores, the later ones add.  Elements
//            are launched one colour at a time and no two elements of a colour
//            share a node, so no atomics and a fixed summation order.
#include "ofd_internal.h"
#include "wave_dpp.h"
#include "element_shape.h"

namespace ofd {

namespace {

constexpr int kWavesPerBlock = 4;

}  // namespace

// alpha stage (assembleNavSto.f90:25-30; solveQuenching.f90:25-37 for the
// quenching scalars) packed into one 128-byte record per node
__global__ void k_alpha_stage(int N, const double *__restrict__ xg,
                              const double *__restrict__ dg,
                              const double *__restrict__ dgold,
                              const double *__restrict__ ug,
                              const double *__restrict__ ugold,
                              const double *__restrict__ ugm,
                              const double *__restrict__ ugmold,
                              const double *__restrict__ acg,
                              const double *__restrict__ acgold,
                              const double *__restrict__ pg,
                              const double *__restrict__ phig,
                              const double *__restrict__ phigold,
                              const double *__restrict__ tg,
                              const double *__restrict__ tgold,
                              const double *__restrict__ rtg,
                              const double *__restrict__ rtgold, double almi,
                              double alfi, double *__restrict__ rec) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  double r[16];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const size_t o = (size_t)c * N + i;
    const double d = dgold[o] + alfi * (dg[o] - dgold[o]);
    r[c] = xg[o] + d;
    r[3 + c] = ugold[o] + alfi * (ug[o] - ugold[o]);
    r[6 + c] = acgold[o] + almi * (acg[o] - acgold[o]);
    r[9 + c] = ugmold[o] + alfi * (ugm[o] - ugmold[o]);
  }
  r[12] = pg[i];
  // quenching scalars (solveQuenching.f90:33-37): phi, T at alfi, and rT or
  // rphi at almi
  if (phig) {
    r[13] = phigold[i] + alfi * (phig[i] - phigold[i]);
    r[14] = tgold[i] + alfi * (tg[i] - tgold[i]);
    r[15] = rtgold[i] + almi * (rtg[i] - rtgold[i]);
  } else {
    r[13] = r[14] = r[15] = 0.0;
  }
  double2 *out = reinterpret_cast<double2 *>(rec + (size_t)i * 16);
#pragma unroll
  for (int t = 0; t < 8; ++t) out[t] = make_double2(r[2 * t], r[2 * t + 1]);
}

struct AsmParams {
  double almi, alfi, gami, delt, rho, mu, kdc, gx, gy, gz;
  int vms;  // 1: the residual with the fine-scale velocity uprime = -tauM rLi
};

// ---------------------------------------------------------------------------
// Hexes, prisms and tets in three kernels (S = 8, 6 or 4 nodes and Gauss
// points; a prism's 6 points take a group of 8 lanes in the point kernel).
//
// The kernel above gives each element a whole wavefront with lane = (q, a):
// everything that belongs to a Gauss point alone (J^-1, G, the tau's, the
// RHS terms before the node split) runs in S copies, every Gauss-point
// interpolation is a lane reduction, and a tet leaves three quarters of the
// lanes idle.  For hexes and tets the volume loop is split:
//
//   k_ns_points<S>  one launch over all elements, 64 / S per wavefront,
//                   lane = (element, q): the Gauss-point work runs once per
//                   lane and the sums over the S nodes are register loops,
//                   added in the association of the lane butterfly
//                   (((p0+p1)+(p2+p3))+((p4+p5)+(p6+p7)) for hexes) with
//                   every product rounded first.  It writes each element's
//                   point data at its position in colour order (PtsLayout):
//                   J^-1 and the advective velocity, the point factors, the
//                   RHS contributions per node (summed over q in the same
//                   association, e3Rhs_3D_fluid), the node numbers; and its
//                   CFL number.
//   k_ns_pairs<S>   one launch per colour, 64 / S^2 elements per wavefront:
//                   lane = (element, q, a) forms dN_a/dx and tmp1_a from the
//                   point data (the same expressions, so the same bits), then
//                   lane = (element, a, b) the 16 entries of block (a, b) over
//                   the S points (e3LHS_3D_fluid_Old), the Dirichlet masks and
//                   the coalesced first-touch scatter of the kernel above.
//   k_ns_pairs_cl   hexes: one launch per cluster colour, one workgroup per
//                   cluster of up to 8 elements around a node, one
//                   wavefront per element forming its blocks as
//                   k_ns_pairs<8> does; the blocks are summed in LDS, one
//                   slot per distinct nonzero, and each slot written once.
//   k_ns_rhs_gather one launch, thread per node: the node's contributions in
//                   colour order (the order of the coloured scatter, so the
//                   same bits), masked by its Dirichlet flags.
// ---------------------------------------------------------------------------
namespace {

constexpr int kRecStride = 14;  // doubles per staged node record

}  // namespace

template <int S>
__global__ void __launch_bounds__(64)
k_ns_points(int E, const int32_t *__restrict__ ien, const int32_t *__restrict__ elm_id,
                const int32_t *__restrict__ cpos, const double *__restrict__ rec,
                const double *__restrict__ mus, AsmParams P, int id1, int id2,
                int subdomain_flag, int do_mat, int do_vec, double *__restrict__ pts,
                double *__restrict__ cfl_out, int32_t *__restrict__ negdet) {
  using L = PtsLayout<S>;
  // the S node records of each element: [el][a][14] (x + d, u, ac, um, p,
  // viscosity scale)
  __shared__ __align__(16) double s_rec[64 * kRecStride];
  constexpr int PL = point_lanes<S>();  // lanes per element (prisms: 6 points of 8)
  const int lane = threadIdx.x;
  const int el = lane / PL;  // 64 / PL elements per wavefront
  const int q = lane % PL;   // Gauss point, and node a = q for the loads
  const int e = blockIdx.x * (64 / PL) + el;
  const bool valid = e < E;
  bool tgt = false;
  if (valid) {
    const int id = elm_id[e];
    if (subdomain_flag & 1) tgt = tgt || id == id1;
    if (subdomain_flag & 2) tgt = tgt || id == id2;
  }
  const bool vec = do_vec && tgt;
  // this element's point data go to its position in colour order
  double *po = pts + (size_t)(valid ? cpos[e] : 0) * L::stride;
  if (q < S) {
    const int n = valid ? ien[(size_t)q * E + e] : 0;
    const double2 *r2 = reinterpret_cast<const double2 *>(rec + (size_t)n * 16);
    double2 *sr = reinterpret_cast<double2 *>(s_rec + (el * S + q) * kRecStride);
#pragma unroll
    for (int t = 0; t < 6; ++t) sr[t] = r2[t];
    sr[6] = make_double2(r2[6].x, mus ? mus[n] : 1.0);
    if (valid && do_mat) {
      int32_t *pn = reinterpret_cast<int32_t *>(po + L::N);
      pn[q] = n;
      if (q == 0) pn[S] = tgt ? 1 : 0;
    }
  }
  __syncthreads();
  const double *re = s_rec + el * S * kRecStride;

  double Nq[S];
#pragma unroll
  for (int a = 0; a < S; ++a) Nq[a] = shape_N<S>(q, a);

  // dx/dxi (i,j) = sum_a (x+d)_a,i dN_a/dxi_j
  double J[9];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    double p0[S], p1[S], p2[S];
#pragma unroll
    for (int a = 0; a < S; ++a) {
      double Na, d0, d1, d2, gw;
      elem_shape<S>(q, a, Na, d0, d1, d2, gw);
      const double x = re[a * kRecStride + i];
      p0[a] = x * d0;
      p1[a] = x * d1;
      p2[a] = x * d2;
    }
    J[i + 0] = tree_sum<S>(p0);
    J[i + 3] = tree_sum<S>(p1);
    J[i + 6] = tree_sum<S>(p2);
  }
  // get_inverse_3x3 (eval_shape.f90:85-112); column-major Ji(i + 3 j)
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

  // Gauss-point values (e3int_qr)
  double ui[3], umi[3], aci[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    double pu[S], pm[S], pc[S];
#pragma unroll
    for (int a = 0; a < S; ++a) {
      pu[a] = re[a * kRecStride + 3 + i] * Nq[a];
      pc[a] = re[a * kRecStride + 6 + i] * Nq[a];
      pm[a] = re[a * kRecStride + 9 + i] * Nq[a];
    }
    ui[i] = tree_sum<S>(pu);
    umi[i] = tree_sum<S>(pm);
    aci[i] = tree_sum<S>(pc);
  }
  double pri;
  {
    double pp[S];
#pragma unroll
    for (int a = 0; a < S; ++a) pp[a] = re[a * kRecStride + 12] * Nq[a];
    pri = tree_sum<S>(pp);
  }
  const double rho = P.rho;
  double mu = P.mu;
  if (mus) {
    double pm[S];
#pragma unroll
    for (int a = 0; a < S; ++a) pm[a] = re[a * kRecStride + 13] * Nq[a];
    mu = P.mu * tree_sum<S>(pm);
  }
  double uadv[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) uadv[i] = ui[i] - umi[i];

  // global gradients of the S nodes and the gradients (e3int_qr_grad)
  double ga[S][3], gwq = 1.0;
#pragma unroll
  for (int a = 0; a < S; ++a) {
    double Na, d0, d1, d2;
    elem_shape<S>(q, a, Na, d0, d1, d2, gwq);
#pragma unroll
    for (int j = 0; j < 3; ++j)
      ga[a][j] = d0 * Ji[0 + 3 * j] + d1 * Ji[1 + 3 * j] + d2 * Ji[2 + 3 * j];
  }
  if (valid && do_mat && q < S) {
    // J^-1 and uadv: the pair kernel forms dN_a/dx and tmp1_a from them with
    // the expressions above, so the same bits
    double2 *pj = reinterpret_cast<double2 *>(po + q * 12);
    pj[0] = make_double2(Ji[0], Ji[1]);
    pj[1] = make_double2(Ji[2], Ji[3]);
    pj[2] = make_double2(Ji[4], Ji[5]);
    pj[3] = make_double2(Ji[6], Ji[7]);
    pj[4] = make_double2(Ji[8], uadv[0]);
    pj[5] = make_double2(uadv[1], uadv[2]);
  }
  double du[9], dp[3];
#pragma unroll
  for (int j = 0; j < 3; ++j) {
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      double pd[S];
#pragma unroll
      for (int a = 0; a < S; ++a) pd[a] = re[a * kRecStride + 3 + i] * ga[a][j];
      du[i + 3 * j] = tree_sum<S>(pd);
    }
    double pd[S];
#pragma unroll
    for (int a = 0; a < S; ++a) pd[a] = re[a * kRecStride + 12] * ga[a][j];
    dp[j] = tree_sum<S>(pd);
  }
  double G[9];
#pragma unroll
  for (int j = 0; j < 3; ++j)
#pragma unroll
    for (int i = 0; i < 3; ++i)
      G[i + 3 * j] = Ji[0 + 3 * i] * Ji[0 + 3 * j] + Ji[1 + 3 * i] * Ji[1 + 3 * j] +
                     Ji[2 + 3 * i] * Ji[2 + 3 * j];
  const double divu = du[0] + du[4] + du[8];

  // e3STAB_3D
  double taua = 0.0, gg = 0.0;
#pragma unroll
  for (int j = 0; j < 3; ++j)
#pragma unroll
    for (int i = 0; i < 3; ++i) taua += uadv[i] * G[i + 3 * j] * uadv[j];
#pragma unroll
  for (int t = 0; t < 9; ++t) gg += G[t] * G[t];
  const double nu = mu / rho;
  const double taud = nu * nu * 12.0 * gg;
  const double taut = 4.0 / (P.delt * P.delt);
  const double tauM = 1.0 / (rho * sqrt(taut + taua + taud));
  const double tauC = rho * sqrt(taua + taud) / (G[0] + G[4] + G[8]);
  // e3CFL
  double cflq = 0.0;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) cflq += uadv[i] * G[i + 3 * j] * uadv[j];
  cflq = P.delt * sqrt(cflq);
  const double dw = detj * gwq;  // hexes: unit weights; padded prism points 0
  if (valid && q < S && detj < 0.0) atomicAdd(negdet, 1);
  if (valid && do_mat && q < S) {
    const double fact2 = P.alfi * P.gami * P.delt;
    const double cf = fact2 * dw;
    double2 *pf = reinterpret_cast<double2 *>(po + L::F + q * 6);
    pf[0] = make_double2(cf, P.almi * rho * dw);
    pf[1] = make_double2(cf * (mu + P.kdc), cf * tauC);
    pf[2] = make_double2(tauM, 0.0);
  }

  // RHS (e3Rhs_3D_fluid, as in the kernel above): the point's terms t1, t2,
  // split over the S nodes and summed over the S points (the lanes of the
  // element) in the butterfly association; lane q keeps node a = q
  if (do_vec) {
    double own[4] = {0.0, 0.0, 0.0, 0.0};
    if (vec) {
      const double mupkdc = mu + P.kdc;
      const double ptot = pri - tauC * divu;
      const double fi[3] = {P.gx * rho, P.gy * rho, P.gz * rho};
      double t1[3], t2[9], up[3] = {0.0, 0.0, 0.0};
      if (P.vms) {
        double adv[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          const double rLi = rho * (aci[i] + uadv[0] * du[i + 0] + uadv[1] * du[i + 3] +
                                    uadv[2] * du[i + 6]) - fi[i] + dp[i];
          up[i] = -tauM * rLi;
          adv[i] = uadv[i] + up[i];
        }
#pragma unroll
        for (int i = 0; i < 3; ++i)
          t1[i] = rho * (aci[i] + adv[0] * du[i + 0] + adv[1] * du[i + 3] +
                         adv[2] * du[i + 6]) - fi[i];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
          for (int j = 0; j < 3; ++j)
            t2[i + 3 * j] = ((i == j) ? (-ptot + 2.0 * mupkdc * du[i + 3 * i])
                                      : mupkdc * (du[i + 3 * j] + du[j + 3 * i])) -
                            rho * adv[j] * up[i];
      } else {
#pragma unroll
        for (int i = 0; i < 3; ++i)
          t1[i] = rho * (aci[i] + uadv[0] * du[i + 0] + uadv[1] * du[i + 3] +
                         uadv[2] * du[i + 6]) - fi[i];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
          for (int j = 0; j < 3; ++j)
            t2[i + 3 * j] = (i == j) ? (-ptot + 2.0 * mupkdc * du[i + 3 * i])
                                     : mupkdc * (du[i + 3 * j] + du[j + 3 * i]);
      }
#pragma unroll
      for (int a = 0; a < S; ++a) {
        double rc[4];
#pragma unroll
        for (int i = 0; i < 3; ++i)
          rc[i] = -(Nq[a] * t1[i] + (ga[a][0] * t2[i] + ga[a][1] * t2[i + 3] +
                                     ga[a][2] * t2[i + 6])) * dw;
        rc[3] = P.vms ? -(Nq[a] * divu - (ga[a][0] * up[0] + ga[a][1] * up[1] +
                                          ga[a][2] * up[2])) * dw
                      : -(Nq[a] * divu) * dw;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          const double sm = point_sum<S>(rc[c]);
          if (q == a) own[c] = sm;
        }
      }
    }
    if (valid && q < S) {
      double2 *pr = reinterpret_cast<double2 *>(po + L::R + q * 4);
      pr[0] = make_double2(own[0], own[1]);
      pr[1] = make_double2(own[2], own[3]);
    }
  }
  cflq = point_max<S>(cflq);
  if (valid && q == 0) cfl_out[e] = cflq;
}

// The per-element part of the pair kernels.  ns_pair_stage<S> puts a
// wavefront's elements' Dirichlet flags, subdomain flags, point factors, the
// N table and (lane = (slot, q, a)) dN_a/dx and tmp1_a into the wavefront's
// LDS area sw: [slot][q][a] (dN_a/dx, dN_a/dy), (dN_a/dz, tmp1_a), the
// factors from FO, the N table from NO.  ns_pair_acc<S> then forms, for
// lane = (slot, a, b), the 16 entries of block (a, b) over the S points
// (e3LHS_3D_fluid_Old) with the Dirichlet projection.
template <int S>
struct PairLds {
  static constexpr int SS = S * S, EW = 64 / SS;
  static constexpr int FO = 256, NO = FO + EW * S * 6, size = NO + SS;
};

template <int S>
__device__ __forceinline__ void ns_pair_stage(double *sw, uint32_t *sm, int32_t *stg, int t0,
                                              int nv, const double *__restrict__ pts,
                                              const uint8_t *__restrict__ ibc, double rho,
                                              int lane) {
  using L = PtsLayout<S>;
  using W = PairLds<S>;
  constexpr int SS = W::SS, EW = W::EW, FO = W::FO, NO = W::NO;
  if (lane < EW * S) {
    const int sl = lane / S, a = lane % S;
    if (sl < nv) {
      const int32_t *pn = reinterpret_cast<const int32_t *>(
          pts + (size_t)(t0 + sl) * L::stride + L::N);
      sm[lane] = ibc[pn[a]];
      if (a == 0) stg[sl] = pn[S];
    }
  }
  if (lane < EW * S * 3) {
    const int sl = lane / (S * 3), pc = lane % (S * 3);
    if (sl < nv)
      reinterpret_cast<double2 *>(sw + FO + sl * S * 6)[pc] =
          reinterpret_cast<const double2 *>(pts + (size_t)(t0 + sl) * L::stride + L::F)[pc];
  }
  if (lane < SS) sw[NO + lane] = shape_N<S>(lane / S, lane % S);
  // lane = (slot, q, a): dN_a/dx and tmp1_a at point q from J^-1 and uadv
  {
    const int sl = lane / SS, r = lane % SS, q = r / S, a = r % S;
    if (sl < nv) {
      const double2 *pj =
          reinterpret_cast<const double2 *>(pts + (size_t)(t0 + sl) * L::stride + q * 12);
      const double2 j0 = pj[0], j1 = pj[1], j2 = pj[2], j3 = pj[3], j4 = pj[4], j5 = pj[5];
      const double Ji[9] = {j0.x, j0.y, j1.x, j1.y, j2.x, j2.y, j3.x, j3.y, j4.x};
      const double uadv[3] = {j4.y, j5.x, j5.y};
      double Na, d0, d1, d2, gw;
      elem_shape<S>(q, a, Na, d0, d1, d2, gw);
      double g[3];
#pragma unroll
      for (int j = 0; j < 3; ++j)
        g[j] = d0 * Ji[0 + 3 * j] + d1 * Ji[1 + 3 * j] + d2 * Ji[2 + 3 * j];
      const double tmp1a = rho * (uadv[0] * g[0] + uadv[1] * g[1] + uadv[2] * g[2]);
      double2 *sg = reinterpret_cast<double2 *>(sw) + lane * 2;
      sg[0] = make_double2(g[0], g[1]);
      sg[1] = make_double2(g[2], tmp1a);
    }
  }
}

template <int S>
__device__ __forceinline__ void ns_pair_acc(const double *sw, const uint32_t *sm,
                                            const int32_t *stg, int nv, int lane,
                                            double acc[16]) {
  using W = PairLds<S>;
  constexpr int SS = W::SS, FO = W::FO, NO = W::NO;
  const int sl = lane / SS, rr = lane % SS;
  const int pa_ = rr % S;  // row node a
  const int pb_ = rr / S;  // column node b
#pragma unroll
  for (int u = 0; u < 16; ++u) acc[u] = 0.0;
  if (sl < nv && stg[sl]) {
    const double2 *sg = reinterpret_cast<const double2 *>(sw) + sl * SS * 2;  // [q][a][2]
    const double2 *sf = reinterpret_cast<const double2 *>(sw + FO + sl * S * 6);  // [q][3]
    const double *sn = sw + NO;                                                   // [q][a]
    // the accumulation of the kernel above, term for term
#pragma unroll 2
    for (int qq = 0; qq < S; ++qq) {
      const double2 r0 = sg[(qq * S + pa_) * 2], r1 = sg[(qq * S + pa_) * 2 + 1];
      const double2 c0 = sg[(qq * S + pb_) * 2], c1 = sg[(qq * S + pb_) * 2 + 1];
      const double2 f0 = sf[qq * 3], f1 = sf[qq * 3 + 1], f2 = sf[qq * 3 + 2];
      const double Nq_ = sn[qq * S + pa_], g0 = r0.x, g1 = r0.y, g2 = r1.x, t1a = r1.y;
      const double Mb = sn[qq * S + pb_], h0 = c0.x, h1 = c0.y, h2 = c1.x, t1b = c1.y;
      const double cf = f0.x, cB = f0.y, cm = f1.x, ct = f1.y, tM = f2.x;
      const double Bb = cB * Mb + cf * t1b;
      const double hf0 = cf * h0, hf1 = cf * h1, hf2 = cf * h2;
      const double hm0 = cm * h0, hm1 = cm * h1, hm2 = cm * h2;
      const double ht0 = ct * h0, ht1 = ct * h1, ht2 = ct * h2;
      const double Mf = cf * Mb, PB = tM * Bb;
      const double Ta = tM * t1a;
      const double Aa = Nq_ + Ta;
      const double t4 = Aa * Bb + (g0 * hm0 + g1 * hm1 + g2 * hm2);
      acc[0] += t4 + g0 * (hm0 + ht0);
      acc[5] += t4 + g1 * (hm1 + ht1);
      acc[10] += t4 + g2 * (hm2 + ht2);
      acc[1] += g1 * hm0 + g0 * ht1;
      acc[4] += g0 * hm1 + g1 * ht0;
      acc[2] += g2 * hm0 + g0 * ht2;
      acc[8] += g0 * hm2 + g2 * ht0;
      acc[6] += g2 * hm1 + g1 * ht2;
      acc[9] += g1 * hm2 + g2 * ht1;
      acc[3] += Ta * hf0 - g0 * Mf;
      acc[7] += Ta * hf1 - g1 * Mf;
      acc[11] += Ta * hf2 - g2 * Mf;
      acc[12] += Nq_ * hf0 + g0 * PB;
      acc[13] += Nq_ * hf1 + g1 * PB;
      acc[14] += Nq_ * hf2 + g2 * PB;
      acc[15] += tM * (g0 * hf0 + g1 * hf1 + g2 * hf2);
    }
  }
  // Dirichlet projection (BCLhs_3D.f90:105-118), as above
  if (sl < nv) {
    const unsigned ma = sm[sl * S + pa_];
    const unsigned mb = sm[sl * S + pb_];
#pragma unroll
    for (int u = 0; u < 16; ++u) {
      const int r = u >> 2, cc = u & 3;
      const bool colz = ((mb >> cc) & 1u) && !(r == 3 && cc == 3);
      if (((ma >> r) & 1u) || colz) acc[u] = 0.0;
      if (pa_ == pb_ && r == cc && ((ma >> r) & 1u)) acc[u] = 1.0;
    }
  }
}

// One colour's elements at positions first .. first + count - 1 of the
// colour order, 64 / S^2 per wavefront (one hex, four tets).  Everything it
// reads is indexed by position: the point data, the scatter map and the
// first-touch masks.  Every global load is issued before the accumulation
// (the old values of the nonzeros it adds into included), so only the
// point data's latency is in front of it.
template <int S>
__global__ void __launch_bounds__(256)
k_ns_pairs(int first, int count, const int32_t *__restrict__ emap_c,
           const uint64_t *__restrict__ first_c, const double *__restrict__ pts,
           const uint8_t *__restrict__ ibc, double rho, double *__restrict__ A) {
  constexpr int SS = S * S;
  constexpr int EW = 64 / SS;       // elements per wavefront
  // per wavefront: the staged point data (PairLds), then the block
  // transpose [64][18]
  __shared__ __align__(16) double s_w[kWavesPerBlock][64 * 18];
  __shared__ uint32_t s_m[kWavesPerBlock][16];  // [slot][a] Dirichlet flags
  __shared__ int32_t s_tg[kWavesPerBlock][4];   // [slot] in the selected subdomains
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const int t0 = first + (blockIdx.x * kWavesPerBlock + w) * EW;
  if (t0 >= first + count) return;  // whole wavefronts; no block barrier below
  const int nv = (first + count - t0 < EW) ? first + count - t0 : EW;  // live slots
  const int piece = lane & 7;
  // the scatter plan: round rnd moves block jj = rnd * 8 + lane / 8, i.e.
  // pair (aj, bj) of slot jj / S^2
  int kk[8];
  bool fst[8];
#pragma unroll
  for (int rnd = 0; rnd < 8; ++rnd) {
    const int jj = rnd * 8 + (lane >> 3);
    const int js = jj / SS, rj = jj % SS;
    const int aj = rj % S, bj = rj / S;
    const int tj = t0 + (js < nv ? js : 0);
    kk[rnd] = emap_c[(size_t)tj * 64 + aj + 8 * bj];
    fst[rnd] = (first_c[tj] >> (aj + 8 * bj)) & 1ull;
  }
  double *sw = s_w[w];
  ns_pair_stage<S>(sw, s_m[w], s_tg[w], t0, nv, pts, ibc, rho, lane);
  // the old values of the nonzeros this lane adds into (not first touches)
  double2 old[8];
#pragma unroll
  for (int rnd = 0; rnd < 8; ++rnd) {
    const int js = (rnd * 8 + (lane >> 3)) / SS;
    old[rnd] = (fst[rnd] || js >= nv)
                   ? make_double2(0.0, 0.0)
                   : reinterpret_cast<const double2 *>(A + (size_t)kk[rnd] * 16)[piece];
  }
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  __builtin_amdgcn_wave_barrier();

  double acc[16];
  ns_pair_acc<S>(sw, s_m[w], s_tg[w], nv, lane, acc);
  // every lane has read the point data: the transpose may overwrite them
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  __builtin_amdgcn_wave_barrier();
  double2 *mine = reinterpret_cast<double2 *>(sw + lane * 18);
#pragma unroll
  for (int u = 0; u < 8; ++u) mine[u] = make_double2(acc[2 * u], acc[2 * u + 1]);
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  __builtin_amdgcn_wave_barrier();
  // coalesced scatter: 8 lanes per 128-byte block; the first element in
  // colour order to reach a nonzero stores, the later ones add to the old
  // value read above
#pragma unroll
  for (int rnd = 0; rnd < 8; ++rnd) {
    const int jj = rnd * 8 + (lane >> 3);
    if (jj / SS >= nv) continue;
    const double2 v = reinterpret_cast<const double2 *>(sw + jj * 18)[piece];
    double2 *dst = reinterpret_cast<double2 *>(A + (size_t)kk[rnd] * 16) + piece;
    if (fst[rnd]) {
      *dst = v;
    } else {
      double2 o = old[rnd];
      o.x += v.x;
      o.y += v.y;
      *dst = o;
    }
  }
}

// Hexes by clusters: one workgroup per cluster of up to kNsClElems elements
// that share a node (one wavefront per element).  Each wavefront forms its
// element's 64 blocks as k_ns_pairs<8> does; the wavefronts then add them,
// in member order, into LDS slots, one per distinct nonzero of the cluster
// (343 for 8 hexes around a node), and the slots go to the matrix once: the
// first cluster in cluster-colour order to reach a nonzero stores, later
// ones add to its old value.  Clusters of one colour share no node, so no
// nonzero.  Per cluster ci: cl_elem[ci][m] the members' positions in the
// element colour order (-1 past the last), cl_map[ci][m][a + 8 b] the slot
// of block (a, b) of member m, cl_kf[ci][s] the nonzero of slot s with bit
// 31 set on a first touch, cl_nslot[ci] the number of slots.
__global__ void __launch_bounds__(64 * kNsClElems) __attribute__((amdgpu_waves_per_eu(4)))
k_ns_pairs_cl(int first, const int32_t *__restrict__ cl_elem,
              const uint16_t *__restrict__ cl_map, const uint32_t *__restrict__ cl_kf,
              const int32_t *__restrict__ cl_nslot, const double *__restrict__ pts,
              const uint8_t *__restrict__ ibc, double rho, double *__restrict__ A) {
  constexpr int kStride = 18;  // doubles per LDS slot: 16 entries, 16-byte aligned rows
  constexpr int kThreads = 64 * kNsClElems;
  constexpr int kRounds = (kNsClSlots * 8 + kThreads - 1) / kThreads;
  __shared__ __align__(16) double s_slot[kNsClSlots * kStride];
  __shared__ __align__(16) double s_w[kNsClElems][PairLds<8>::size];
  __shared__ uint32_t s_m[kNsClElems][8];
  __shared__ int32_t s_tg[kNsClElems][1];
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int w = tid >> 6;
  const size_t ci = (size_t)first + blockIdx.x;
  const int nslot = cl_nslot[ci];
  const int t = cl_elem[ci * kNsClElems + w];
  const int nv = t >= 0 ? 1 : 0;
  // the write plan: piece p = tid + kThreads * r of slot p / 8
  uint32_t kf[kRounds];
#pragma unroll
  for (int r = 0; r < kRounds; ++r) {
    const int sl = (tid + kThreads * r) >> 3;
    kf[r] = sl < nslot ? cl_kf[ci * kNsClSlots + sl] : 0x80000000u;
  }
  const int my_slot = nv ? cl_map[(ci * kNsClElems + w) * 64 + lane] : 0;
  double *sw = s_w[w];
  if (nv) ns_pair_stage<8>(sw, s_m[w], s_tg[w], t, 1, pts, ibc, rho, lane);
  for (int i = tid; i < nslot * (kStride / 2); i += kThreads)
    reinterpret_cast<double2 *>(s_slot)[i] = make_double2(0.0, 0.0);
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  __builtin_amdgcn_wave_barrier();
  double acc[16];
  ns_pair_acc<8>(sw, s_m[w], s_tg[w], nv, lane, acc);
  // the old values of the nonzeros that are not first touches, in flight
  // over the ordered adds
  double2 old[kRounds];
#pragma unroll
  for (int r = 0; r < kRounds; ++r)
    old[r] = (kf[r] >> 31) ? make_double2(0.0, 0.0)
                           : reinterpret_cast<const double2 *>(
                                 A + (size_t)(kf[r] & 0x7fffffffu) * 16)[(tid + kThreads * r) & 7];
  __syncthreads();  // the slots are zero
  // the members' blocks into the slots, in member order (the 64 blocks of
  // one element are 64 distinct nonzeros)
  for (int m = 0; m < kNsClElems; ++m) {
    if (w == m && nv) {
      double2 *d = reinterpret_cast<double2 *>(s_slot + my_slot * kStride);
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        double2 o = d[u];
        o.x += acc[2 * u];
        o.y += acc[2 * u + 1];
        d[u] = o;
      }
    }
    __syncthreads();
  }
  // one write per nonzero, 8 threads per 128-byte block
#pragma unroll
  for (int r = 0; r < kRounds; ++r) {
    const int p = tid + kThreads * r;
    const int sl = p >> 3;
    if (sl >= nslot) continue;
    const double2 v = reinterpret_cast<const double2 *>(s_slot + sl * kStride)[p & 7];
    double2 *dst = reinterpret_cast<double2 *>(A + (size_t)(kf[r] & 0x7fffffffu) * 16) + (p & 7);
    if (kf[r] >> 31) {
      *dst = v;
    } else {
      double2 o = old[r];
      o.x += v.x;
      o.y += v.y;
      *dst = o;
    }
  }
}

// rhs[n] = sum of the node's element contributions in colour order
// (inc_off/inc: entries t * 8 + a, t the element's position in colour
// order), with its Dirichlet components zeroed
__global__ void __launch_bounds__(256)
k_ns_rhs_gather(int N, const int32_t *__restrict__ inc_off, const int32_t *__restrict__ inc,
                const double *__restrict__ pts, int stride, int roff,
                const uint8_t *__restrict__ ibc, double *__restrict__ rhs) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  double s[4] = {0.0, 0.0, 0.0, 0.0};
  const unsigned m0 = ibc[n];
  for (int t = inc_off[n]; t < inc_off[n + 1]; ++t) {
    const int ea = inc[t];
    const double2 *pr =
        reinterpret_cast<const double2 *>(pts + (size_t)(ea >> 3) * stride + roff +
                                          (ea & 7) * 4);
    const double2 r0 = pr[0], r1 = pr[1];
    const double v[4] = {r0.x, r0.y, r1.x, r1.y};
#pragma unroll
    for (int c = 0; c < 4; ++c) s[c] += ((m0 >> c) & 1u) ? 0.0 : v[c];
  }
  double2 *rr = reinterpret_cast<double2 *>(rhs + (size_t)n * 4);
  rr[0] = make_double2(s[0], s[1]);
  rr[1] = make_double2(s[2], s[3]);
}

__global__ void k_zero(double *p, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = 0.0;
}

// Maximum of v[0..n) (n >= 0; 0 for n = 0) in two launches: each block of
// the first writes the maximum of its grid-stride share to part[block], the
// single block of the second reduces the parts.  The maximum does not depend
// on the order, so the result is the same bits on every run.
__global__ void __launch_bounds__(256) k_max_part(const double *__restrict__ v, int n,
                                                  double *__restrict__ part) {
  __shared__ double s[256];
  double m = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    m = fmax(m, v[i]);
  s[threadIdx.x] = m;
  __syncthreads();
  for (int h = blockDim.x / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) s[threadIdx.x] = fmax(s[threadIdx.x], s[threadIdx.x + h]);
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

__global__ void __launch_bounds__(256) k_max_reduce(const double *__restrict__ v, int n,
                                                    double *__restrict__ out) {
  __shared__ double s[256];
  double m = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) m = fmax(m, v[i]);
  s[threadIdx.x] = m;
  __syncthreads();
  for (int h = blockDim.x / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) s[threadIdx.x] = fmax(s[threadIdx.x], s[threadIdx.x + h]);
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = s[0];
}

// the record's last slot holds rT (temperature equation) or, with rphi set,
// the VOF rate rphi (coupled NS + VOF system)
hipError_t launch_alpha_stage(Ctx &c, bool rphi) {
  const int N = c.nnode;
  const double almi = c.prm[OFD_P_ALMI], alfi = c.prm[OFD_P_ALFI];
  hipLaunchKernelGGL(k_alpha_stage, dim3((N + 255) / 256), dim3(256), 0, c.stream, N,
                     c.d_field[OFD_XG], c.d_field[OFD_DG], c.d_field[OFD_DGOLD],
                     c.d_field[OFD_UG], c.d_field[OFD_UGOLD], c.d_field[OFD_UGM],
                     c.d_field[OFD_UGMOLD], c.d_field[OFD_ACG],
                     c.d_field[OFD_ACGOLD], c.d_field[OFD_PG],
                     c.quench ? c.d_field[OFD_PHIG] : nullptr, c.d_field[OFD_PHIGOLD],
                     c.d_field[OFD_TG], c.d_field[OFD_TGOLD],
                     c.d_field[rphi ? OFD_RPHIG : OFD_RTG],
                     c.d_field[rphi ? OFD_RPHIGOLD : OFD_RTGOLD], almi, alfi, c.d_noderec);
  return hipGetLastError();
}

static AsmParams asm_params(const Ctx &c) {
  AsmParams P;
  P.almi = c.prm[OFD_P_ALMI];
  P.alfi = c.prm[OFD_P_ALFI];
  P.gami = c.prm[OFD_P_GAMI];
  P.delt = c.prm[OFD_P_DELT];
  P.rho = c.prm[OFD_P_RHO];
  P.mu = c.prm[OFD_P_MU];
  P.kdc = c.prm[OFD_P_KDC];
  P.gx = c.prm[OFD_P_GX];
  P.gy = c.prm[OFD_P_GY];
  P.gz = c.prm[OFD_P_GZ];
  P.vms = c.res_form;
  return P;
}

// The volume loop of assembleNavSto (IntElmAss_NS_overset): hexes through
// the point, pair and gather kernels (c.d_pts sized by the caller), tets and
// prisms by the one-wavefront-per-element kernel over zeroed RHS.
hipError_t launch_assemble_ns(Ctx &c, int tensor_flag, int subdomain_flag) {
  const AsmParams P = asm_params(c);
  const int do_mat = (tensor_flag & OFD_ASSEMBLE_MAT) ? 1 : 0;
  const int do_vec = (tensor_flag & OFD_ASSEMBLE_VEC) ? 1 : 0;
  const double *mus = c.mus_on ? c.d_field[OFD_MUSCALE] : nullptr;
  hipError_t e;
  {
    const bool hex = c.nshl == 8;
    const int S = c.nshl;
    const int epw = 64 / (S == 4 ? 4 : 8);  // point kernel: elements per wavefront
    const int epb = 64 / (S * S) * kWavesPerBlock;  // pair kernel: per block
    hipLaunchKernelGGL((S == 8 ? k_ns_points<8> : S == 6 ? k_ns_points<6> : k_ns_points<4>),
                       dim3((c.nelem + epw - 1) / epw), dim3(64), 0, c.stream, c.nelem,
                       c.d_ien, c.d_elm_id, c.d_cpos, c.d_noderec, mus, P, c.elm_id_set[0],
                       c.elm_id_set[1], subdomain_flag, do_mat, do_vec, c.d_pts, c.d_cfl,
                       c.d_negdet);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if (do_mat && hex && !c.cl_off.empty()) {
      for (int col = 0; col + 1 < (int)c.cl_off.size(); ++col) {
        const int first = c.cl_off[col];
        const int count = c.cl_off[col + 1] - first;
        if (count <= 0) continue;
        hipLaunchKernelGGL(k_ns_pairs_cl, dim3(count), dim3(64 * kNsClElems), 0, c.stream,
                           first, c.d_cl_elem, c.d_cl_map, c.d_cl_kf, c.d_cl_nslot, c.d_pts,
                           c.d_ibc, P.rho, c.d_A64);
        if ((e = hipGetLastError()) != hipSuccess) return e;
      }
    } else if (do_mat) {
      for (int col = 0; col + 1 < (int)c.color_off.size(); ++col) {
        const int first = c.color_off[col];
        const int count = c.color_off[col + 1] - first;
        if (count <= 0) continue;
        hipLaunchKernelGGL((S == 8 ? k_ns_pairs<8> : S == 6 ? k_ns_pairs<6> : k_ns_pairs<4>),
                           dim3((count + epb - 1) / epb), dim3(64 * kWavesPerBlock), 0,
                           c.stream, first, count, c.d_emap_c, c.d_first_c, c.d_pts, c.d_ibc,
                           P.rho, c.d_A64);
        if ((e = hipGetLastError()) != hipSuccess) return e;
      }
    }
    if (do_vec) {
      hipLaunchKernelGGL(k_ns_rhs_gather, dim3((c.nnode + 255) / 256), dim3(256), 0, c.stream,
                         c.nnode, c.d_inc_off, c.d_inc, c.d_pts, pts_stride(c.nshl),
                         S == 8 ? PtsLayout<8>::R : S == 6 ? PtsLayout<6>::R : PtsLayout<4>::R,
                         c.d_ibc, c.d_rhs);
      if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    return hipSuccess;
  }
}

hipError_t launch_zero(Ctx &c, double *p, size_t n) {
  if (n == 0) return hipSuccess;
  size_t blocks = (n + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(k_zero, dim3((unsigned)blocks), dim3(256), 0, c.stream, p, n);
  return hipGetLastError();
}

// out_dev[0] = max(v[0..n)); out_dev[1 .. 1 + 512) holds the per-block parts
hipError_t launch_max_reduce(Ctx &c, const double *v, int n, double *out_dev) {
  int blocks = (n + 2047) / 2048;  // at least 8 values per thread
  blocks = blocks < 1 ? 1 : (blocks > 512 ? 512 : blocks);
  hipLaunchKernelGGL(k_max_part, dim3(blocks), dim3(256), 0, c.stream, v, n, out_dev + 1);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_max_reduce, dim3(1), dim3(256), 0, c.stream, out_dev + 1, blocks,
                     out_dev);
  return hipGetLastError();
}

}  // namespace ofd
