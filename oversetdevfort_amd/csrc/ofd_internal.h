// Internal definitions of the gfx950 OversetDevFort hot-path library.
//
// Device data layout (HBM):
//   node fields        SoA [3][N] doubles (the reference's column-major layout)
//   node record        [N][16] doubles: x+d (3), u (3), ac (3), um (3), p,
//                      pad (3) -- the alpha-stage gather unit of the element
//                      kernels, one 128-byte line per node
//   system matrix      block CSR in the reference's nonzero order k:
//                      A[k][16] row-major 4x4 = [K11 G; D1 M]
//                      (double, plus a float copy for the mixed solver)
//   preconditioner     [N][16] 4x4 block-diagonal inverse [Binv 0; 0 1/M]
//   vectors            [N][4] interleaved (u, v, w, p) doubles
//   scatter map        emap[e][64]: nonzero index of pair (a, b), lane = a + 8 b
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/ofd.h"

namespace ofd {

void set_error(const std::string &msg);

#define OFD_HIP(call)                                                    \
  do {                                                                   \
    hipError_t _e = (call);                                              \
    if (_e != hipSuccess) {                                              \
      ::ofd::set_error(std::string(#call) + ": " + hipGetErrorString(_e)); \
      return OFD_ERR_HIP;                                                \
    }                                                                    \
  } while (0)

#define OFD_CHECK(cond, code, msg) \
  do {                             \
    if (!(cond)) {                 \
      ::ofd::set_error(msg);       \
      return code;                 \
    }                              \
  } while (0)

constexpr int kMaxKspace = 64;     // GMRES restart length upper bound
constexpr int kNsClElems = 8;      // hex pair kernel: elements per cluster (one node's)
constexpr int kNsClSlots = 352;    // distinct nonzeros per cluster (343 for 8 hexes)
constexpr int kRedBlock = 256;     // threads per reduction block
constexpr int kMaxRedBlocks = 4096;  // node blocks of 256: N <= 1,048,576

// per-solve scalars living on the device (one GMRES cycle)
struct GmresState {
  double H[(kMaxKspace + 1) * kMaxKspace];  // column-major (K+1) x K
  double e[kMaxKspace + 1];
  double ec[kMaxKspace + 1];  // e_k as the close of step k-1 left it: the close of
                             // step k reads it here, never in e (which it rotates)
  double rc[kMaxKspace], rs[kMaxKspace];
  double s[kMaxKspace + 1];   // scale of stored basis vector j (v_j = s_j W_j)
  double h1[kMaxKspace + 1], h2[kMaxKspace + 1];
  double y[kMaxKspace];
  double beta;                // running MGS projection coefficient
  double eps;                 // tol * |P^-1 b|
  double tol;
  double rnorm0;
  double resid;               // |e_{k+1}| at exit
  double scal[4];             // misc reductions (norms)
  double outer_thr;           // restarted solve: stop once |b - A x| <= outer_thr
  int32_t iters;
  int32_t done;
  int32_t kmin;
  int32_t kspace;
  int32_t outer_done;         // restarted solve reached outer_thr: later cycles are no-ops
  int32_t pad;
};

// What the host reads after each cycle of a restarted solve, written by the
// outer residual's norm kernel into coherent pinned memory: the fields
// first, then `seq` (the cycle's number) with a system-scope release.
struct CycleRecord {
  double resid;   // GmresState::resid
  double rnorm0;  // GmresState::rnorm0
  double rnorm;   // |b - A x| after the cycle
  int32_t iters;  // GmresState::iters
  int32_t seq;
};

struct Ctx {
  int device = 0;
  int n_cu = 256;                     // compute units of the device
  hipStream_t stream = nullptr;
  // mesh
  int nnode = 0, nelem = 0, nshl = 0;
  std::vector<int32_t> h_ien;   // [nshl][nelem] 0-based
  std::vector<int32_t> h_elm_id;
  int32_t *d_ien = nullptr;
  int32_t *d_elm_id = nullptr;
  // pattern
  int64_t icnt = 0;
  std::vector<int32_t> h_col, h_row;  // 0-based
  int32_t *d_col = nullptr, *d_row = nullptr;
  int32_t *d_emap = nullptr;          // [E][64]
  int32_t *d_color_elems = nullptr;   // elements grouped by colour
  int32_t *d_inc_off = nullptr;       // [N+1] node -> its (element, node) entries
  int32_t *d_inc = nullptr;           // t * 8 + a: position t in colour order, node a
  double *d_pts = nullptr;            // hex/tet NS assembly: point data per element,
  size_t pts_len = 0;                 // by position in colour order (PtsLayout); doubles
  int32_t *d_emap_c = nullptr;        // hexes, tets: emap and first-touch masks by colour
  uint64_t *d_first_c = nullptr;      // position, and each element's position
  int32_t *d_cpos = nullptr;
  uint64_t *d_first = nullptr;        // [E] bit a + 8 b: the element is the first (in
                                      // colour order) to add into pair (a, b)'s nonzero
  std::vector<int32_t> color_off;     // ncolor+1
  // hexes: clusters of the pair kernel (k_ns_pairs_cl), by cluster colour;
  // empty when the mesh has none
  std::vector<int32_t> cl_off;        // ncolour+1 over the clusters
  int32_t *d_cl_elem = nullptr;       // [C][kNsClElems] positions in colour order, -1 none
  uint16_t *d_cl_map = nullptr;       // [C][kNsClElems][64] slot of block (a, b)
  uint32_t *d_cl_kf = nullptr;        // [C][kNsClSlots] nonzero | first-touch bit 31
  int32_t *d_cl_nslot = nullptr;      // [C]
  std::vector<int32_t> h_elem_color;  // colour of each element
  // weak boundary faces (BCugType 2): registered list, and the elements
  // that hold them grouped by colour with their faces as a bitmask
  std::vector<int32_t> h_face_e, h_face_or;  // 0-based element, FACE_OR 1..6
  int32_t *d_fe_elem = nullptr, *d_fe_mask = nullptr;
  std::vector<int32_t> fe_color_off;  // ncolor+1 (empty: no faces)
  bool fe_dirty = false;
  int32_t *d_diag = nullptr;          // nonzero index of the diagonal block per row
  // parameters
  double prm[OFD_NPARAMS] = {0};
  int32_t elm_id_set[2] = {1, 2};
  // state
  double *d_field[OFD_NFIELDS] = {nullptr};
  uint8_t *d_ibc = nullptr;           // bit c set: IBC(node, c+1) == 1
  double *d_bcval = nullptr;          // [3][N] Dirichlet velocity values
  double *d_noderec = nullptr;        // [N][16]
  // system
  double *d_A64 = nullptr;            // [icnt][16]
  float *d_A32 = nullptr;
  double *d_P64 = nullptr;            // [N][16]
  double *d_Ap64 = nullptr;           // [icnt][16] block-Jacobi-preconditioned P^-1 A
  float *d_Ap32 = nullptr;
  bool mat_valid = false, prec_valid = false, a32_valid = false, ap32_valid = false;
  double *d_rhs = nullptr;            // [N][4]
  double *d_b = nullptr;              // [N][4] right-hand side of a solve
  double *d_x = nullptr;              // [N][4] solution accumulator
  double *d_r = nullptr;              // [N][4] outer residual
  double *d_cfl = nullptr;            // [E] per-element max CFL
  int32_t *d_negdet = nullptr;        // count of negative Jacobians
  // GMRES workspace
  int kalloc = 0;
  double *d_W = nullptr;              // [(kalloc+1)][N][4] FP64 Krylov basis
  float *d_W32 = nullptr;             // [(kalloc32+1)][N][4] FP32 basis (mixed CGS solver)
  int kalloc32 = 0;
  bool basis32 = false;               // the running cycle uses d_W32
  double *d_part = nullptr;           // 4 x [kMaxRedBlocks][kMaxKspace + 1] partial-sum buffers
  GmresState *d_gs = nullptr;
  GmresState *h_gs = nullptr;         // pinned mirror
  int part_cur = 0;                   // partial buffer holding the latest sums
  int32_t *h_flags = nullptr;         // coherent pinned: 2 x [kMaxKspace] GmresState::done
                                      // per step (-1 unset), one set per cycle parity
  int32_t *cur_flags = nullptr;       // the set of the cycle being enqueued
  CycleRecord *h_rec = nullptr;       // coherent pinned: one record per cycle parity
  int32_t rec_seq = 0;                // number of the last cycle record asked for
  double *h_pin = nullptr;            // pinned scratch (host transfers)
  size_t h_pin_bytes = 0;
  int64_t dev_bytes = 0;
  std::vector<double> solve_hist;     // relres after each cycle of the last solve
  // quenching temperature system (LHSTem/RHSGTem): scalar per nonzero / node;
  // vectors zero-padded to whole 4-tuples for the GMRES kernels
  double qprm[OFD_NQPARAMS] = {0};
  bool quench = false;                // ofd_set_quench_params was called
  double *d_AT = nullptr;             // [icnt] LHSTem
  float *d_AT32 = nullptr;            // its FP32 copy (plain SpMV)
  double *d_PT = nullptr;             // [N] inverse diagonal, 0 -> 1
  double *d_ApT = nullptr;            // [icnt] diag(P)^-1 LHSTem
  float *d_ApT32 = nullptr;
  bool matT_valid = false, precT_valid = false, aT32_valid = false, apT32_valid = false;
  double *d_rhsT = nullptr, *d_xT = nullptr, *d_rT = nullptr;  // [4 * ceil(N/4)]
  double *d_yT = nullptr;             // [4 * ceil(N/4)] product input scratch
  int64_t temT_icnt = 0;              // pattern size LHSTem was allocated for
  int temT_n4 = 0;                    // 4-tuples the T vectors were allocated for
  // NSVOF system: [icnt][24] per nonzero, the 4x4 block [K11 G; D1 M] (the
  // NS layout) then the phi couplings (ULS(r), LSU(r)) r < 3, (PLS, LS);
  // node vectors [N][4] (u, p) then [N] phi, zero-padded to whole 4-tuples
  double *d_AV = nullptr;
  float *d_AV32 = nullptr;
  double *d_PV = nullptr;             // [N][16] 4x4 block inverses, then [N] 1/LS
  double *d_ApV = nullptr;            // P^-1 A, same two-array layout
  float *d_ApV32 = nullptr;
  bool matV_valid = false, precV_valid = false, aV32_valid = false, apV32_valid = false;
  bool pV_valid = false;              // the NSVOF block inverses alone (right form)
  // right-preconditioned cycle (NS or NSVOF): the Arnoldi product reads
  // T = P^-1 W_k (d_T, or d_T32 with the FP32 basis) and the plain matrix
  bool right = false;
  double *d_T = nullptr;
  float *d_T32 = nullptr;
  int t_n = 0;                        // 4-tuples d_T / d_T32 hold
  double *d_rhsV = nullptr, *d_xV = nullptr, *d_rV = nullptr, *d_yV = nullptr;
  int64_t nsvV_icnt = 0;
  int nsvV_n4 = 0;
  // the system the GMRES kernels run on: 0 the NS node vectors [N][4]
  // (gm_n = N), 1 the T vector as ceil(N/4) 4-tuples, 2 the NSVOF vectors
  // as ceil(5N/4) 4-tuples
  int gm_n = 0;
  int gm_kind = 0;
  int kalloc_n = 0;                   // tuples per Krylov vector allocated
  // which increment SOLU / SOLP (/ SOLPHI) show: 0 the NS d_x, 2 the NSVOF d_xV
  int sol_src = 0;
  // per-node viscosity scale (OFD_MUSCALE) in use by the NS assembly
  bool mus_on = false;
  // NS residual form: 0 the reference's (uprime = 0), 1 with the fine-scale
  // velocity uprime = -tauM rLi
  int res_form = 0;
  // adjoint: the mirrored nonzero of each nonzero (the pattern is symmetric),
  // the transposed LHS and its solver copies (swapped in for the stored-
  // matrix kernels by a view), bumped / recorded versions of the LHS
  int32_t *d_tmap = nullptr;          // [icnt]
  double *t_A64 = nullptr, *t_P64 = nullptr, *t_Ap64 = nullptr;
  float *t_A32 = nullptr, *t_Ap32 = nullptr;
  bool t_mat_valid = false, t_prec_valid = false, t_a32_valid = false, t_ap32_valid = false;
  int64_t mat_version = 0, t_version = -1;
  double *d_jt = nullptr;             // [N][4] J_bc^T output, inner-solve residual
  // [N][4] adjoint solution, residual, right-hand side and correction
  double *d_lam = nullptr, *d_lamr = nullptr, *d_lamb = nullptr, *d_lamd = nullptr;
  // flexible GMRES on J_bc or J_bc^T: basis V [(m+1)][N][4], preconditioned
  // directions Z [m][N][4], small device vectors of coefficients
  int fg_m = 0;
  size_t fg_n = 0;
  double *d_fv = nullptr, *d_fz = nullptr, *d_fh = nullptr;
  double *d_vjp = nullptr;            // [8N] raw gradients u, ac, p, scale
  // temperature adjoint: per-node conductivity scale (OFD_KSCALE) in use,
  // LHSTem transposed and its solver copies (swapped in by a view), the
  // bumped / recorded versions of LHSTem, and the adjoint vectors
  // (4 * ceil(N/4) each, zero padded): lambda, right-hand side, inner-solve
  // residual, operator output; raw gradients [6N] T, rT, u, scale
  bool ks_on = false;
  double *tT_A = nullptr, *tT_P = nullptr, *tT_Ap = nullptr;
  float *tT_A32 = nullptr, *tT_Ap32 = nullptr;
  bool tT_mat_valid = false, tT_prec_valid = false, tT_a32_valid = false,
       tT_ap32_valid = false;
  int64_t matT_version = 0, tT_version = -1;
  double *d_lamT = nullptr, *d_lamTb = nullptr, *d_lamTd = nullptr, *d_jtT = nullptr;
  double *d_vjpT = nullptr;
  int tadj_n4 = 0;
  // NS + VOF adjoint: the stored matrix transposed (tmap mirror; its (phi
  // row, p column) entries dropped) with its solver copies, swapped in by a
  // view for the inner solves; adjoint vectors (4 nsvV_n4 each) lambda,
  // right-hand side, inner residual, operator output; raw gradients [10N]
  double *t_AV = nullptr, *t_PV = nullptr, *t_ApV = nullptr;
  float *t_AV32 = nullptr, *t_ApV32 = nullptr;
  bool t_matV_valid = false, t_precV_valid = false, t_pV_valid = false,
       t_aV32_valid = false, t_apV32_valid = false;
  int64_t t_nsv_icnt = 0;
  double *d_lamV = nullptr, *d_lamVb = nullptr, *d_lamVd = nullptr, *d_jtV = nullptr;
  double *d_vjpV = nullptr;
  int vadj_n4 = 0;
  // overset fringe: per-element boxes, source flags and geometry; receivers
  int fr_alloc_e = 0, fr_alloc_r = 0;
  double *d_fr_box = nullptr;         // [E][6] widened bounding box
  int8_t *d_fr_ok = nullptr;          // [E] element is a source element
  double *d_fr_geo = nullptr;         // [E][24] tet inverse edge matrix / hex corners
  double *d_fr_xr = nullptr, *d_fr_dr = nullptr, *d_fr_val = nullptr;  // (nrecv,3)
  int32_t *d_fr_donor = nullptr, *d_fr_nodes = nullptr;                // (nrecv)
  // sampled timing of the Arnoldi product kernel inside real solves: every
  // probe_every-th iteration is bracketed by a pair of timing events, which
  // are read back at the next host synchronisation of the solve
  int probe_every = 0;
  std::vector<hipEvent_t> probe_ev;   // 2 per slot
  int probe_used = 0;                 // slots recorded since the last read-back
  int64_t probe_tick = 0;             // Arnoldi product launches since ofd_probe_set
  int64_t probe_count = 0;
  double probe_ms = 0.0;
};

// ---- launchers (defined in the .hip files) ----
// assembly
hipError_t launch_alpha_stage(Ctx &c, bool rphi = false);
hipError_t launch_assemble_ns(Ctx &c, int tensor_flag, int subdomain_flag);
// adjoint (adjoint.hip)
hipError_t launch_block_transpose(Ctx &c);  // t_A64 = blocks of A64 mirrored and transposed
// gradients of lambda . F (lam [N][4]) into d_vjp (zeroed by the caller)
hipError_t launch_vjp_ns(Ctx &c, int subdomain_flag, const double *lam);
// y = J_bc^T lam
hipError_t launch_jt_ns(Ctx &c, int subdomain_flag, const double *lam, double *y);
// y = J_bc v, v a Newton increment
hipError_t launch_jv_ns(Ctx &c, int subdomain_flag, const double *v, double *y);
// temperature adjoint (adjoint_tem.hip): tT_A = LHSTem mirrored through d_tmap
hipError_t launch_scalar_transpose(Ctx &c);
// exact derivatives by dual numbers (quench_dual.hip), hexes, one colour per
// launch: the projected NS + VOF Jacobian into d_AV; the VJP of the NS + VOF
// residual (lam4 [N][4], lamphi [N]) into out [10][N]; the VJP of the
// temperature residual (lam [N]) into out [7][N] (outputs zeroed by the
// caller; alpha stage with rphi for NS + VOF, with rT for T)
hipError_t launch_nsvof_jac(Ctx &c, int color);
hipError_t launch_nsvof_vjp(Ctx &c, int color, const double *lam4, const double *lamphi,
                            double *out);
hipError_t launch_tem_dual_vjp(Ctx &c, int color, const double *ks, const double *lam,
                               double *out, const uint8_t *tmask);
// the temperature adjoint's operators by dual numbers: raw gradients [7][N]
// (T, rT, u (3), phi, scale; zeroed here), y = J_T,bc^T lambda (scratch [7N])
hipError_t launch_vjp_tem_dual(Ctx &c, const double *lam, double *out);
hipError_t launch_jt_tem_dual(Ctx &c, const double *lam, double *y, double *scratch);
// At = the stored NS + VOF matrix transposed without its (phi, p) slot;
// y = A^T x exactly (FP64)
hipError_t launch_nsvof_transpose(Ctx &c, double *At);
hipError_t launch_spmv_v_t(Ctx &c, const double *x, double *y);
// gradients of lambda . R_T (lam [N]) into out [6N] (zeroed by the caller)
hipError_t launch_vjp_tem(Ctx &c, const double *lam, double *out);
// y = J_T,bc^T lam (y: 4 ceil(N/4), zeroed here)
hipError_t launch_jt_tem(Ctx &c, const double *lam, double *y);
// out = x + a y
hipError_t launch_axpby(Ctx &c, const double *x, const double *y, double a, double *out,
                        size_t n);
// out[j] = <w, V_j>, j < m <= 128 (deterministic; uses d_part)
hipError_t launch_multidot(Ctx &c, size_t n, const double *w, const double *V, size_t vs,
                           int m, double *out);
// w += sign sum_j h_j V_j (h on the device)
hipError_t launch_multi_axpy(Ctx &c, size_t n, double *w, const double *V, size_t vs,
                             const double *h, int m, double sign);
hipError_t launch_scale(Ctx &c, size_t n, const double *x, double a, double *y);
hipError_t launch_zero(Ctx &c, double *p, size_t n);
hipError_t launch_max_reduce(Ctx &c, const double *v, int n, double *out_dev);
// matrix
hipError_t launch_build_prec(Ctx &c);
hipError_t launch_precondition_matrix(Ctx &c, bool want32);
hipError_t launch_convert_a32(Ctx &c);
hipError_t launch_spmv(Ctx &c, int precision, const double *x, double *y,
                       const double *b, int mode);
hipError_t launch_spmv_arnoldi(Ctx &c, int precision, int k, hipEvent_t ev0 = nullptr,
                               hipEvent_t ev1 = nullptr);
hipError_t launch_cgs_dots(Ctx &c, int k);
hipError_t launch_face_ns(Ctx &c, int tensor_flag, int subdomain_flag, int color);
// genSparStruc on the device (pattern.hip)
hipError_t launch_node_elem_count(Ctx &c, int32_t *cnt);
hipError_t launch_node_elem_fill(Ctx &c, const int32_t *off, int32_t *pos, int32_t *inc);
hipError_t launch_adjacency(Ctx &c, const int32_t *off, const int32_t *inc, const int32_t *col,
                            int32_t *out, int32_t *diag, int32_t *err);
hipError_t launch_tmap_emap(Ctx &c, int32_t *err);
// d_first from d_emap and the colouring (touched: [icnt] zeroed bytes)
hipError_t launch_first_touch(Ctx &c, uint8_t *touched);
hipError_t launch_colour_maps(Ctx &c);
hipError_t scan_offsets(Ctx &c, const int32_t *in, int32_t *out, int n, void *tmp,
                        size_t &tmp_bytes);
hipError_t launch_cgs_update(Ctx &c, int k, int mode);
// vectors
hipError_t launch_soa_to_node4(Ctx &c, const double *u, const double *p,
                               double *out);
hipError_t launch_node4_to_soa(Ctx &c, const double *in, double *u, double *p);
hipError_t launch_axpy_fields(Ctx &c, int which);
hipError_t launch_fields_s(Ctx &c, int which, const double *x, int fa, int faold, int fra,
                           int fraold);
hipError_t launch_fields_v(Ctx &c);
hipError_t launch_nsvof_to_soa(Ctx &c, const double *in);
hipError_t launch_norm2_v(Ctx &c, const double *v);
hipError_t launch_apply_dirichlet(Ctx &c);
hipError_t launch_copy(Ctx &c, const void *src, void *dst, size_t nbytes);
constexpr int kReadBlocks = 8192;  // blocks of the read-only stream kernel
// NS point data per element of S nodes and S Gauss points (hexes, tets), in
// doubles: [q] (J^-1 column-major, uadv) 12 S, [q] (cf, cB, cm, ct, tauM, 0)
// 6 S, [a] RHS contributions 4 S, then the S node numbers and the subdomain
// flag as int32; 1536 bytes per hex, 1088 per prism, 768 per tet
template <int S>
struct PtsLayout {
  static constexpr int F = 12 * S, R = 18 * S, N = 22 * S;
  static constexpr int stride = S == 8 ? 192 : S == 6 ? 136 : 96;
  // the S + 1 int32 after N fit, and rows stay 16-byte aligned
  static_assert(2 * N + S + 1 <= 2 * stride && stride % 2 == 0, "point data layout");
};
constexpr int pts_stride(int nshl) { return nshl == 8 ? 192 : nshl == 6 ? 136 : 96; }
hipError_t launch_read(Ctx &c, const void *src, size_t nbytes, double *out);
// overset fringe
hipError_t launch_fringe(Ctx &c, int32_t src_id, int nrecv, const double *xr,
                         const double *dr, int32_t *donor, double *val, double *box,
                         int8_t *ok, double *geo);
// static scratch of ofd_fringe_interp_noalloc (fringe.hip): its size, the
// once-per-device lookup of its address (context creation) and the bytes a
// search over nelem elements and nrecv receivers needs
constexpr size_t kFringeScratchBytes = size_t(64) << 20;
int fringe_scratch_setup(int device);
size_t fringe_scratch_need(int nshl, int nelem, int nrecv);
hipError_t launch_fringe_apply(Ctx &c, int nrecv, const int32_t *nodes, const double *val);
// gmres pieces
hipError_t launch_gmres_init(Ctx &c, const double *b, double tol, int kspace, int kmin);
hipError_t launch_mgs_step(Ctx &c, int k, int j);
hipError_t launch_gmres_close(Ctx &c, int k, int ortho);
hipError_t launch_gmres_finish(Ctx &c, double *x);
hipError_t launch_norm2(Ctx &c, const double *v, int slot, int outer = 0,
                        CycleRecord *rec = nullptr, int32_t seq = 0);
hipError_t launch_outer_reset(Ctx &c, double thr);
hipError_t launch_norm2_up(Ctx &c, const double *v, int slot);
int nblocks_gm(const Ctx &c);
// temperature system
hipError_t launch_assemble_tem(Ctx &c, int tensor_flag);
hipError_t launch_spmv_s(Ctx &c, int precision, const double *x, double *y,
                         const double *b, int mode);
hipError_t launch_build_prec_s(Ctx &c, bool want32);
hipError_t launch_convert_aT32(Ctx &c);
// NSVOF system
hipError_t launch_assemble_nsvof_hex(Ctx &c, int tensor_flag, int color);
hipError_t launch_spmv_v(Ctx &c, int precision, const double *x, double *y,
                         const double *b, int mode);
hipError_t launch_build_prec_v(Ctx &c, bool want32, bool ap);
hipError_t launch_convert_aV32(Ctx &c);

int nblocks_nodes(const Ctx &c);

}  // namespace ofd

struct ofd_ctx {
  ofd::Ctx c;
};
