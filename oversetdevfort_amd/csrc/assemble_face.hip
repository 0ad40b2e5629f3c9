// Boundary-face terms of the Navier-Stokes system on gfx950, hexes and tets:
// FaceAssembly_NS_weak (FaceAssembly_weak.f90:4-193) with e3bSTAB_weak
// (e3STAB.f90:256-287), e3bRHS_weak (e3RHS.f90:105-170) and e3bLHS_weak
// (e3LHS.f90:167-294), then FaceAssembly_NS_outflow
// (FaceAssembly_weak.f90:197-409), which the reference runs on the same
// (BCugType 2) faces; each face's element arrays go through BCLhs_3D and
// FillSparseMat_3D as in the volume loop.  assembleNavSto calls both after
// IntElmAss_NS_overset (assembleNavSto.f90:50-60), and so does
// ofd_assemble_ns.
//
// The reference has face maps for tets and prisms only.  Hex faces (an
// extension): face fo (1..6) holds xi_c = -1 (fo odd) or +1 (fo even),
// c = (fo-1)/2; 2x2 points at the volume rule's abscissa 0.57735026918963
// over the other two coordinates (p, q) = ((c+1)%3, (c+2)%3), unit weights;
// n dA = sc (dx/dxi_p x dx/dxi_q), which points out of an element with
// positive Jacobian.  The wall velocity g is the mesh velocity (gi = umi).
//
// One 64-lane wavefront per element with faces (its faces as a bitmask):
//   phase A  per face, lane = (point q < 4, node a): shape functions,
//            Jacobian and its inverse, face tangents and normal,
//            interpolated u, um, p, grad u (8-lane butterflies); tauB;
//            the weak and outflow RHS of node a, summed over q and faces
//   phase B  lane = (a, b): the 4x4 block (a, b) summed over points and
//            faces: e3bLHS_weak's K11, G, D1 and the outflow K11 diagonal
//   phase C  BCLhs_3D as row/column masks; each face passes through it
//            twice (once per routine), so a constrained diagonal receives
//            2 per face; then the read-modify-write at emap[e][a + 8 b],
//            one element colour per launch (no two elements of a colour
//            share a node).
// One wave per block: the face loop's trip count differs between elements.
#include "ofd_internal.h"
#include "wave_dpp.h"

namespace ofd {

namespace {

__constant__ int8_t c_fsign[8][3] = {{-1, 1, -1}, {1, 1, -1}, {1, 1, 1},  {-1, 1, 1},
                                     {-1, -1, -1}, {1, -1, -1}, {1, -1, 1}, {-1, -1, 1}};

__device__ __forceinline__ double fred8(double v) { return wsum8(v); }

}  // namespace

struct FaceParams {
  double rho, mu, fact2;
};

__global__ void __launch_bounds__(64)
k_face_ns_hex(int E, const int32_t *__restrict__ ien, const int32_t *__restrict__ elm_id,
              const int32_t *__restrict__ fe_elem, const int32_t *__restrict__ fe_mask,
              int first, int count, const int32_t *__restrict__ emap,
              const double *__restrict__ rec, const uint8_t *__restrict__ ibc, FaceParams P,
              int id1, int id2, int subdomain_flag, int do_mat, int do_vec,
              double *__restrict__ A, double *__restrict__ rhs) {
  __shared__ double s_sh[4][8][5];  // [q][a]: N, dN/dx (3), dN/dx . n
  __shared__ double s_q[4][6];      // [q]: n (3), DetJb * w, tauB, u_neg
  __shared__ int32_t s_node[8];

  const int lane = threadIdx.x;
  const int idx = first + blockIdx.x;
  if (idx >= first + count) return;
  const int e = fe_elem[idx];
  const int fmask = fe_mask[idx];
  const int id = elm_id[e];
  bool tgt = false;
  if (subdomain_flag & 1) tgt = tgt || id == id1;
  if (subdomain_flag & 2) tgt = tgt || id == id2;

  // ---------------- node data, lane = (q, a) ----------------
  const int q = lane >> 3;
  const int a = lane & 7;
  const int na = ien[(size_t)a * E + e];
  if (lane < 8) s_node[lane] = na;
  const double2 *r2 = reinterpret_cast<const double2 *>(rec + (size_t)na * 16);
  double xr[13];
#pragma unroll
  for (int t = 0; t < 6; ++t) {
    const double2 v = r2[t];
    xr[2 * t] = v.x;
    xr[2 * t + 1] = v.y;
  }
  xr[12] = r2[6].x;
  const double *xd = xr;       // x + d
  const double *ua = xr + 3;   // u
  const double *uma = xr + 9;  // um
  const double pa = xr[12];
  const double sx = c_fsign[a][0], sy = c_fsign[a][1], sz = c_fsign[a][2];

  double ru[3] = {0.0, 0.0, 0.0}, rp = 0.0;  // RHS of node a, lanes q < 4
  double acc[16];
#pragma unroll
  for (int t = 0; t < 16; ++t) acc[t] = 0.0;
  const int pa_ = lane & 7;   // phase B row node
  const int pb_ = lane >> 3;  // phase B column node
  const double gpt = 0.57735026918963;
  int nface = 0;

  for (int fo = 1; fo <= 6; ++fo) {
    if (!((fmask >> (fo - 1)) & 1)) continue;  // the same for the whole wave
    ++nface;
    // ---------------- phase A ----------------
    const int c = (fo - 1) / 2, p = (c + 1) % 3, pq = (c + 2) % 3;
    const double sc = (fo & 1) ? -1.0 : 1.0;
    const int qq = q & 3;  // lanes q >= 4 repeat points 0..3 and keep nothing
    double xi3[3];
    xi3[c] = sc;
    xi3[p] = (qq & 1) ? gpt : -gpt;
    xi3[pq] = (qq & 2) ? gpt : -gpt;
    const double fx = (sx > 0) ? (xi3[0] + 1.0) : (1.0 - xi3[0]);
    const double fy = (sy > 0) ? (xi3[1] + 1.0) : (1.0 - xi3[1]);
    const double fz = (sz > 0) ? (xi3[2] + 1.0) : (1.0 - xi3[2]);
    const double Na = fx * fy * fz * 0.125;
    double dl[3];
    dl[0] = sx * (fy * fz * 0.125);
    dl[1] = sy * (fx * fz * 0.125);
    dl[2] = sz * (fx * fy * 0.125);
    double J[9];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      J[i + 0] = fred8(xd[i] * dl[0]);
      J[i + 3] = fred8(xd[i] * dl[1]);
      J[i + 6] = fred8(xd[i] * dl[2]);
    }
    // get_inverse_3x3 (eval_shape.f90:85-112); Ji(i + 3 j) = dxi_i/dx_j
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
    double ga[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) ga[j] = dl[0] * Ji[0 + 3 * j] + dl[1] * Ji[1 + 3 * j] + dl[2] * Ji[2 + 3 * j];
    // face normal: sc (dx/dxi_p x dx/dxi_q), columns p and q of J
    double t1[3], t2[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      t1[i] = sc * J[i + 3 * p];
      t2[i] = J[i + 3 * pq];
    }
    double nor[3];
    nor[0] = t1[1] * t2[2] - t1[2] * t2[1];
    nor[1] = t1[2] * t2[0] - t1[0] * t2[2];
    nor[2] = t1[0] * t2[1] - t1[1] * t2[0];
    const double detjb = sqrt(nor[0] * nor[0] + nor[1] * nor[1] + nor[2] * nor[2]);
    nor[0] /= detjb;
    nor[1] /= detjb;
    nor[2] /= detjb;
    const double wq = detjb;  // unit weights
    // e3int_qr / e3int_qr_grad
    double ui[3], umi[3], du[9];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      ui[i] = fred8(ua[i] * Na);
      umi[i] = fred8(uma[i] * Na);
#pragma unroll
      for (int j = 0; j < 3; ++j) du[i + 3 * j] = fred8(ua[i] * ga[j]);
    }
    const double pri = fred8(pa * Na);
    // e3bSTAB_weak: tauB = 4 mu / h_n, h_n = 2 / |dxidx n|; tauNor = tauB
    double tt = 0.0;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const double s = Ji[i] * nor[0] + Ji[i + 3] * nor[1] + Ji[i + 6] * nor[2];
      tt += s * s;
    }
    const double tauB = 4.0 * P.mu / (2.0 / sqrt(tt));
    const double *gi = umi;
    const double unor_rel = (ui[0] - umi[0]) * nor[0] + (ui[1] - umi[1]) * nor[1] +
                            (ui[2] - umi[2]) * nor[2];
    const double uneg = 0.5 * (unor_rel - fabs(unor_rel));
    const double gnor = gi[0] * nor[0] + gi[1] * nor[1] + gi[2] * nor[2];
    const double unor = ui[0] * nor[0] + ui[1] * nor[1] + ui[2] * nor[2];
    if (do_vec && q < 4) {
      if (tgt) {
        // e3bRHS_weak, gmul = 1 (tauNor - tauB = 0)
        double ti[3], tv[3];
#pragma unroll
        for (int k = 0; k < 3; ++k)
          ti[k] = -pri * nor[k] + P.mu * (du[k] * nor[0] + du[k + 3] * nor[1] + du[k + 6] * nor[2] +
                                          du[3 * k] * nor[0] + du[1 + 3 * k] * nor[1] +
                                          du[2 + 3 * k] * nor[2]);
#pragma unroll
        for (int k = 0; k < 3; ++k)
          tv[k] = -ti[k] + tauB * (ui[k] - gi[k]) - uneg * P.rho * (ui[k] - gi[k]);
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          double s = 0.0;
#pragma unroll
          for (int j = 0; j < 3; ++j)
            s += ga[j] * -(P.mu * ((ui[i] - gi[i]) * nor[j] + (ui[j] - gi[j]) * nor[i]));
          ru[i] -= (Na * tv[i] + s) * wq;
        }
        rp += Na * (unor - gnor) * wq;
      }
      // FaceAssembly_NS_outflow's RHS (every face, no subdomain test)
#pragma unroll
      for (int i = 0; i < 3; ++i) ru[i] += Na * P.rho * uneg * (ui[i] - gi[i]) * wq;
    }
    if (q < 4) {
      s_sh[q][a][0] = Na;
      s_sh[q][a][1] = ga[0];
      s_sh[q][a][2] = ga[1];
      s_sh[q][a][3] = ga[2];
      s_sh[q][a][4] = ga[0] * nor[0] + ga[1] * nor[1] + ga[2] * nor[2];
      if (a == 0) {
        s_q[q][0] = nor[0];
        s_q[q][1] = nor[1];
        s_q[q][2] = nor[2];
        s_q[q][3] = wq;
        s_q[q][4] = tauB;
        s_q[q][5] = uneg;
      }
    }
    __syncthreads();
    // ---------------- phase B: lane = (a, b) ----------------
    if (do_mat) {
      for (int k = 0; k < 4; ++k) {
        const double Nq = s_sh[k][pa_][0], Nb = s_sh[k][pb_][0];
        const double gA[3] = {s_sh[k][pa_][1], s_sh[k][pa_][2], s_sh[k][pa_][3]};
        const double gB[3] = {s_sh[k][pb_][1], s_sh[k][pb_][2], s_sh[k][pb_][3]};
        const double tnA = s_sh[k][pa_][4], tnB = s_sh[k][pb_][4];
        const double n3[3] = {s_q[k][0], s_q[k][1], s_q[k][2]};
        const double w = s_q[k][3], tB = s_q[k][4], un = s_q[k][5];
        if (tgt) {
          // e3bLHS_weak, gmul = 1; the (tauNor - tauB) n n term is zero
          const double t2 = -Nq * P.mu * tnB - tnA * P.mu * Nb + Nq * tB * Nb - Nq * un * P.rho * Nb;
#pragma unroll
          for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int cc = 0; cc < 3; ++cc) {
              const double Av = Nq * P.mu * gB[r] * n3[cc];
              const double Bv = gA[cc] * n3[r] * P.mu * Nb;
              const double v = (r == cc) ? (t2 - Av - Bv) : (-Av - Bv);
              acc[r * 4 + cc] += P.fact2 * v * w;
            }
#pragma unroll
          for (int d = 0; d < 3; ++d) {
            const double v = P.fact2 * Nq * Nb * n3[d] * w;
            acc[d * 4 + 3] += v;   // G
            acc[12 + d] -= v;      // D1
          }
        }
        // FaceAssembly_NS_outflow's K11 diagonal
        const double to = -(Nq * un * P.rho * Nb);
#pragma unroll
        for (int r = 0; r < 3; ++r) acc[r * 5] += P.fact2 * to * w;
      }
    }
    __syncthreads();
  }

  // RHS of node a summed over the 4 points (lanes q >= 4 hold zeros)
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    ru[i] += __shfl_xor(ru[i], 8);
    ru[i] += __shfl_xor(ru[i], 16);
    ru[i] += __shfl_xor(ru[i], 32);
  }
  rp += __shfl_xor(rp, 8);
  rp += __shfl_xor(rp, 16);
  rp += __shfl_xor(rp, 32);

  // ---------------- phase C: BCLhs_3D per face and routine, scatter ----------------
  const double diag = 2.0 * nface;
  if (do_mat) {
    const unsigned ma = ibc[s_node[pa_]];
    const unsigned mb = ibc[s_node[pb_]];
    const int k = emap[(size_t)e * 64 + lane];
    double2 *blk = reinterpret_cast<double2 *>(A + (size_t)k * 16);
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      double v[2] = {acc[2 * t], acc[2 * t + 1]};
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int r = (2 * t + h) >> 2, cc = (2 * t + h) & 3;
        const bool colz = ((mb >> cc) & 1u) && !(r == 3 && cc == 3);
        if (((ma >> r) & 1u) || colz) v[h] = 0.0;
        if (pa_ == pb_ && r == cc && ((ma >> r) & 1u)) v[h] = diag;
      }
      double2 old = blk[t];
      old.x += v[0];
      old.y += v[1];
      blk[t] = old;
    }
  }
  if (do_vec && lane < 8) {
    const unsigned m0 = ibc[na];
    double o[4] = {ru[0], ru[1], ru[2], rp};
#pragma unroll
    for (int cc = 0; cc < 4; ++cc)
      if ((m0 >> cc) & 1u) o[cc] = 0.0;
    double2 *rr = reinterpret_cast<double2 *>(rhs + (size_t)na * 4);
    double2 r0 = rr[0], r1 = rr[1];
    r0.x += o[0];
    r0.y += o[1];
    r1.x += o[2];
    r1.y += o[3];
    rr[0] = r0;
    rr[1] = r1;
  }
}

// Tet faces: the reference's own face map.  eval_faceshape_tet
// (eval_shape_FEM.f90:66-143) takes the face's three displaced nodes in the
// order of FACE_OR (1: 2 3 1, 2: 4 3 2, 3: 1 3 4, 4: 2 1 4) with the
// lintrishl gradients (1, 0), (0, 1), (-1, -1) (eval_shape_FEM.f90:261-282):
// n dA = dx/dxi_1 x dx/dxi_2, constant over the face.  The points are
// genGPtritet's images of the 3-point triangle rule in the tet (genGP_GW.f90:
// 478-572), weights 1/6 (NGAUSSB = NSHLB = 3, input.f90:160-161); the volume
// shape functions lintetshl at them (eval_shape_FEM.f90:201-225), whose
// gradients, and so grad u, are constant over the element.
//
// One wavefront per element with faces: lanes (a, b) < 16 sum block (a, b)
// over the element's faces and points, lanes 16 + a the RHS of node a; each
// lane forms the element's Jacobian, gradients and the points' values
// itself (4 nodes).  BCLhs_3D and the scatter as for hexes (2 per face on a
// constrained diagonal).
__constant__ int8_t c_tface[4][3] = {{1, 2, 0}, {3, 2, 1}, {0, 2, 3}, {1, 0, 3}};
__constant__ double c_tgp[4][3][3] = {
    {{0.166666666666666666667, 0.666666666666666666667, 0.166666666666666666667},
     {0.166666666666666666667, 0.166666666666666666667, 0.666666666666666666667},
     {0.666666666666666666667, 0.166666666666666666667, 0.166666666666666666667}},
    {{0.0, 0.166666666666666666667, 0.166666666666666666667},
     {0.0, 0.166666666666666666667, 0.666666666666666666667},
     {0.0, 0.666666666666666666667, 0.166666666666666666667}},
    {{0.666666666666666666667, 0.0, 0.166666666666666666667},
     {0.166666666666666666667, 0.0, 0.666666666666666666667},
     {0.166666666666666666667, 0.0, 0.166666666666666666667}},
    {{0.166666666666666666667, 0.666666666666666666667, 0.0},
     {0.666666666666666666667, 0.166666666666666666667, 0.0},
     {0.166666666666666666667, 0.166666666666666666667, 0.0}}};

__global__ void __launch_bounds__(64)
k_face_ns_tet(int E, const int32_t *__restrict__ ien, const int32_t *__restrict__ elm_id,
              const int32_t *__restrict__ fe_elem, const int32_t *__restrict__ fe_mask,
              int first, int count, const int32_t *__restrict__ emap,
              const double *__restrict__ rec, const uint8_t *__restrict__ ibc, FaceParams P,
              int id1, int id2, int subdomain_flag, int do_mat, int do_vec,
              double *__restrict__ A, double *__restrict__ rhs) {
  const int lane = threadIdx.x;
  const int idx = first + blockIdx.x;
  if (idx >= first + count || lane >= 20) return;
  const int e = fe_elem[idx];
  const int fmask = fe_mask[idx];
  const int id = elm_id[e];
  bool tgt = false;
  if (subdomain_flag & 1) tgt = tgt || id == id1;
  if (subdomain_flag & 2) tgt = tgt || id == id2;
  const bool blk = lane < 16;
  const int pa_ = blk ? (lane & 3) : lane - 16;  // block row node, or the RHS node
  const int pb_ = blk ? (lane >> 2) : 0;         // block column node
  // node records: x + d, u, um, p
  int nd[4];
  double xl[4][3], ul[4][3], uml[4][3], pl[4];
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    nd[a] = ien[(size_t)a * E + e];
    const double *r = rec + (size_t)nd[a] * 16;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      xl[a][i] = r[i];
      ul[a][i] = r[3 + i];
      uml[a][i] = r[9 + i];
    }
    pl[a] = r[12];
  }
  // lintetshl gradients in xi: node 1..3 the unit vectors, node 4 -(1,1,1)
  const double dl[4][3] = {{1.0, 0.0, 0.0}, {0.0, 1.0, 0.0}, {0.0, 0.0, 1.0},
                           {-1.0, -1.0, -1.0}};
  double J[9];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      double s = 0.0;
#pragma unroll
      for (int a = 0; a < 4; ++a) s = s + xl[a][i] * dl[a][j];
      J[i + 3 * j] = s;
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
  double g[4][3];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      g[a][j] = dl[a][0] * Ji[0 + 3 * j] + dl[a][1] * Ji[1 + 3 * j] + dl[a][2] * Ji[2 + 3 * j];
  double du[9];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      double s = 0.0;
#pragma unroll
      for (int a = 0; a < 4; ++a) s = s + ul[a][i] * g[a][j];
      du[i + 3 * j] = s;
    }

  double acc[16];
#pragma unroll
  for (int t = 0; t < 16; ++t) acc[t] = 0.0;
  double ru[3] = {0.0, 0.0, 0.0}, rp = 0.0;
  int nface = 0;
  for (int fo = 1; fo <= 4; ++fo) {
    if (!((fmask >> (fo - 1)) & 1)) continue;
    ++nface;
    double t1[3] = {0.0, 0.0, 0.0}, t2[3] = {0.0, 0.0, 0.0};
    const double gr[3][2] = {{1.0, 0.0}, {0.0, 1.0}, {-1.0, -1.0}};
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const int a = c_tface[fo - 1][i];
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        t1[d] = t1[d] + xl[a][d] * gr[i][0];
        t2[d] = t2[d] + xl[a][d] * gr[i][1];
      }
    }
    double nor[3];
    nor[0] = t1[1] * t2[2] - t1[2] * t2[1];
    nor[1] = t1[2] * t2[0] - t1[0] * t2[2];
    nor[2] = t1[0] * t2[1] - t1[1] * t2[0];
    const double detjb = sqrt(nor[0] * nor[0] + nor[1] * nor[1] + nor[2] * nor[2]);
    nor[0] /= detjb;
    nor[1] /= detjb;
    nor[2] /= detjb;
    double tt = 0.0;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const double s = Ji[i] * nor[0] + Ji[i + 3] * nor[1] + Ji[i + 6] * nor[2];
      tt += s * s;
    }
    const double tauB = 4.0 * P.mu / (2.0 / sqrt(tt));
    double tn[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) tn[a] = g[a][0] * nor[0] + g[a][1] * nor[1] + g[a][2] * nor[2];
    for (int q = 0; q < 3; ++q) {
      const double xi = c_tgp[fo - 1][q][0], eta = c_tgp[fo - 1][q][1],
                   zeta = c_tgp[fo - 1][q][2];
      const double Nv[4] = {xi, eta, zeta, 1.0 - xi - eta - zeta};
      const double w = detjb * 0.166666666666666666667;
      double ui[3], umi[3], pri = 0.0;
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        double su = 0.0, sm = 0.0;
#pragma unroll
        for (int a = 0; a < 4; ++a) {
          su = su + ul[a][i] * Nv[a];
          sm = sm + uml[a][i] * Nv[a];
        }
        ui[i] = su;
        umi[i] = sm;
      }
#pragma unroll
      for (int a = 0; a < 4; ++a) pri = pri + pl[a] * Nv[a];
      const double *gi = umi;  // the wall moves with the mesh
      const double unor_rel = (ui[0] - umi[0]) * nor[0] + (ui[1] - umi[1]) * nor[1] +
                              (ui[2] - umi[2]) * nor[2];
      const double uneg = 0.5 * (unor_rel - fabs(unor_rel));
      if (blk) {
        if (!do_mat) continue;
        const double Na = Nv[pa_], Nb = Nv[pb_];
        if (tgt) {
          // e3bLHS_weak, gmul = 1; the (tauNor - tauB) n n term is zero
          const double t2v = -Na * P.mu * tn[pb_] - tn[pa_] * P.mu * Nb + Na * tauB * Nb -
                             Na * uneg * P.rho * Nb;
#pragma unroll
          for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int cc = 0; cc < 3; ++cc) {
              const double Av = Na * P.mu * g[pb_][r] * nor[cc];
              const double Bv = g[pa_][cc] * nor[r] * P.mu * Nb;
              const double v = (r == cc) ? (t2v - Av - Bv) : (-Av - Bv);
              acc[r * 4 + cc] += P.fact2 * v * w;
            }
#pragma unroll
          for (int d = 0; d < 3; ++d) {
            const double v = P.fact2 * Na * Nb * nor[d] * w;
            acc[d * 4 + 3] += v;   // G
            acc[12 + d] -= v;      // D1
          }
        }
        // FaceAssembly_NS_outflow's K11 diagonal
        const double to = -(Na * uneg * P.rho * Nb);
#pragma unroll
        for (int r = 0; r < 3; ++r) acc[r * 5] += P.fact2 * to * w;
      } else if (do_vec) {
        const double Na = Nv[pa_];
        if (tgt) {
          // e3bRHS_weak, gmul = 1
          const double gnor = gi[0] * nor[0] + gi[1] * nor[1] + gi[2] * nor[2];
          const double unor = ui[0] * nor[0] + ui[1] * nor[1] + ui[2] * nor[2];
          double ti[3], tv[3];
#pragma unroll
          for (int k = 0; k < 3; ++k)
            ti[k] = -pri * nor[k] + P.mu * (du[k] * nor[0] + du[k + 3] * nor[1] +
                                            du[k + 6] * nor[2] + du[3 * k] * nor[0] +
                                            du[1 + 3 * k] * nor[1] + du[2 + 3 * k] * nor[2]);
#pragma unroll
          for (int k = 0; k < 3; ++k)
            tv[k] = -ti[k] + tauB * (ui[k] - gi[k]) - uneg * P.rho * (ui[k] - gi[k]);
#pragma unroll
          for (int i = 0; i < 3; ++i) {
            double s = 0.0;
#pragma unroll
            for (int j = 0; j < 3; ++j)
              s += g[pa_][j] * -(P.mu * ((ui[i] - gi[i]) * nor[j] + (ui[j] - gi[j]) * nor[i]));
            ru[i] -= (Na * tv[i] + s) * w;
          }
          rp += Na * (unor - gnor) * w;
        }
#pragma unroll
        for (int i = 0; i < 3; ++i) ru[i] += Na * P.rho * uneg * (ui[i] - gi[i]) * w;
      }
    }
  }

  // BCLhs_3D per face and routine, scatter (this colour's elements share no node)
  if (blk && do_mat) {
    const double diag = 2.0 * nface;
    const unsigned ma = ibc[nd[pa_]];
    const unsigned mb = ibc[nd[pb_]];
    const int k = emap[(size_t)e * 64 + pa_ + 8 * pb_];
    double *b16 = A + (size_t)k * 16;
#pragma unroll
    for (int t = 0; t < 16; ++t) {
      const int r = t >> 2, cc = t & 3;
      double v = acc[t];
      const bool colz = ((mb >> cc) & 1u) && !(r == 3 && cc == 3);
      if (((ma >> r) & 1u) || colz) v = 0.0;
      if (pa_ == pb_ && r == cc && ((ma >> r) & 1u)) v = diag;
      b16[t] += v;
    }
  }
  if (!blk && do_vec) {
    const unsigned m0 = ibc[nd[pa_]];
    const double o[4] = {ru[0], ru[1], ru[2], rp};
    double *rr = rhs + (size_t)nd[pa_] * 4;
#pragma unroll
    for (int cc = 0; cc < 4; ++cc) rr[cc] += ((m0 >> cc) & 1u) ? 0.0 : o[cc];
  }
}

hipError_t launch_face_ns(Ctx &c, int tensor_flag, int subdomain_flag, int color) {
  const int first = c.fe_color_off[color];
  const int count = c.fe_color_off[color + 1] - first;
  if (count <= 0) return hipSuccess;
  FaceParams P;
  P.rho = c.prm[OFD_P_RHO];
  P.mu = c.prm[OFD_P_MU];
  P.fact2 = c.prm[OFD_P_ALFI] * c.prm[OFD_P_GAMI] * c.prm[OFD_P_DELT];
  hipLaunchKernelGGL((c.nshl == 4 ? k_face_ns_tet : k_face_ns_hex), dim3(count), dim3(64), 0,
                     c.stream, c.nelem, c.d_ien,
                     c.d_elm_id, c.d_fe_elem, c.d_fe_mask, first, count, c.d_emap,
                     c.d_noderec, c.d_ibc, P, c.elm_id_set[0], c.elm_id_set[1],
                     subdomain_flag, (tensor_flag & OFD_ASSEMBLE_MAT) ? 1 : 0,
                     (tensor_flag & OFD_ASSEMBLE_VEC) ? 1 : 0, c.d_A64, c.d_rhs);
  return hipGetLastError();
}

}  // namespace ofd
