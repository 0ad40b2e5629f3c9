/*
 * ofd.h -- C ABI of the MI355X (gfx950) implementation of the OversetDevFort
 * hot path: VMS/SUPG element assembly, CSR scatter, block SpMV, block-Jacobi
 * GMRES and the overset fringe interpolation.
 *
 * Plain pointers and sizes only.  Every entry returns an int status
 * (OFD_OK = 0); ofd_last_error() gives the message of the last failure on
 * the calling thread.  A context owns its device buffers and one HIP stream;
 * calls on a context are serialised on that stream and are synchronous on
 * return (host outputs are valid when the call returns).  A context is not
 * thread-safe; use one per thread.
 *
 * Array conventions are the reference's (Fortran column-major, 1-based
 * indices), so a Fortran caller passes its module arrays unchanged:
 *   IEN(NELEM,NSHL), ELM_ID(NELEM), IBC(NNODE,ncol), xg/ug/...(NNODE,3),
 *   pg(NNODE), col(NNODE+1), row(icnt), LHSK11(9,icnt), LHSG(3,icnt),
 *   LHSD1(3,icnt), LHSM(icnt), RHSGu(NNODE,3), RHSGp(NNODE).
 * Citations name the reference routine each entry point replaces
 * (paths relative to the reference's src/).
 */
#ifndef OFD_H
#define OFD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ofd_ctx ofd_ctx;

enum {
  OFD_OK = 0,
  OFD_ERR_ARG = 1,        /* bad argument / size / call order */
  OFD_ERR_HIP = 2,        /* HIP runtime error */
  OFD_ERR_NEGDETJ = 3,    /* assembly met a negative Jacobian (warning in the reference) */
  OFD_ERR_NONFINITE = 4,  /* non-finite norm */
  OFD_ERR_NOCONV = 5,     /* solver stopped before the tolerance */
  OFD_ERR_ALLOC = 6,
  OFD_ERR_UNSUPPORTED = 7
};

/* assemble_tensor_flag bits: modules.f90:225-228 */
enum { OFD_ASSEMBLE_VEC = 2, OFD_ASSEMBLE_MAT = 4 };

/* precision of the stored matrix / preconditioner used by SpMV and GMRES */
enum { OFD_FP64 = 0, OFD_FP32 = 1 };

/* Gram-Schmidt variant of the Arnoldi loop */
enum {
  OFD_ORTHO_MGS = 0,  /* SparseGMRES_NS.f90:245-287 sequence, k+1 sweeps */
  OFD_ORTHO_CGS2 = 1, /* classical GS with one re-orthogonalisation */
  OFD_ORTHO_CGS1 = 2  /* classical GS, one projection (restart cycles are short) */
};

/* OR'ed into the ortho argument of ofd_solve_nsvof or ofd_solve with
 * OFD_ORTHO_CGS1 or OFD_ORTHO_CGS2: the same block-Jacobi preconditioner
 * applied on the right, GMRES on A P^-1 (x = P^-1 y), so each cycle
 * minimises the true residual.  No counterpart in the reference, which
 * applies it on the left (SparseGMRES_NS.f90:53-121,
 * SparseGMRES_NS_conv.f:53-150); on the quench tank from 32^3 nodes up the
 * left form of the NS + VOF solve stagnates well above 1e-8.  The other
 * solve entry points reject it. */
enum { OFD_PREC_RIGHT = 16 };

/* node fields held by a context: (NNODE,3) vectors from OFD_XG to OFD_ACGMOLD,
 * OFD_SOLU and OFD_LAMU; (NNODE) scalars otherwise.  The quenching fields
 * (VOF phig, its rate rphig, temperature Tg, its rate rTg and the T
 * increment SOLT) follow modules.f90:291-300.  OFD_LAMU/OFD_LAMP hold the
 * last adjoint solution; OFD_MUSCALE, once set, scales the viscosity of the
 * NS assembly per node (mu at a Gauss point = mu * sum_a N_a s_a), the
 * parameter of the viscosity gradient (no reference counterpart).
 * OFD_KSCALE does the same for the conductivity of the temperature
 * assembly (fluid and solid elements) and OFD_LAMT holds the last
 * temperature adjoint solution. */
enum {
  OFD_XG = 0, OFD_DG, OFD_DGOLD, OFD_UG, OFD_UGOLD, OFD_UGM, OFD_UGMOLD,
  OFD_ACG, OFD_ACGOLD, OFD_ACGM, OFD_ACGMOLD, OFD_PG, OFD_PGOLD,
  OFD_SOLU, OFD_SOLP,
  OFD_PHIG, OFD_PHIGOLD, OFD_RPHIG, OFD_RPHIGOLD, OFD_TG, OFD_TGOLD,
  OFD_RTG, OFD_RTGOLD, OFD_SOLT, OFD_SOLPHI, OFD_LAMU, OFD_LAMP, OFD_MUSCALE,
  OFD_KSCALE, OFD_LAMT,
  OFD_NFIELDS
};

/* parameter vector of ofd_set_params */
enum {
  OFD_P_ALMI = 0, OFD_P_ALFI, OFD_P_GAMI, OFD_P_BETI, OFD_P_DELT, OFD_P_RHO,
  OFD_P_MU, OFD_P_KDC, OFD_P_GX, OFD_P_GY, OFD_P_GZ, OFD_NPARAMS
};

/* material vector of ofd_set_quench_params (param.f90:222-242, 405-406):
 * air/vapour (a), water (w) and solid (s) properties, saturation
 * temperature, latent heat, phase-change rate coefficients, the T
 * discontinuity-capturing coefficient and reference temperature, the
 * ELM_ID of fluid elements (102 in the reference, solveQuenching.f90:693)
 * and the VOF discontinuity-capturing coefficient LSC_kdc */
enum {
  OFD_Q_RHOA = 0, OFD_Q_RHOW, OFD_Q_MUA, OFD_Q_MUW, OFD_Q_CPA, OFD_Q_CPW,
  OFD_Q_KAPPAA, OFD_Q_KAPPAW, OFD_Q_RHOS, OFD_Q_MUS, OFD_Q_CPS, OFD_Q_KAPPAS,
  OFD_Q_TS, OFD_Q_LH, OFD_Q_CCOND, OFD_Q_CEVAP, OFD_Q_CDMDOT, OFD_Q_TEMKDC,
  OFD_Q_TREF, OFD_Q_FLUID_ID, OFD_Q_LSCKDC, OFD_NQPARAMS
};

const char *ofd_last_error(void);
int ofd_version(void);

/* context lifetime; device = HIP ordinal */
int ofd_ctx_create(int device, ofd_ctx **out);
int ofd_ctx_destroy(ofd_ctx *ctx);

/* mesh: replaces input_mesh (input.f90:15-135) for the in-memory case */
int ofd_upload_mesh(ofd_ctx *ctx, int nnode, int nelem, int nshl,
                    const int32_t *ien, const int32_t *elm_id, const double *xg);

/* genSparStruc (genSparStruc.f90:4-47): builds col/row, colours the
 * elements and the (element,a,b) -> nonzero map used by the scatter */
int ofd_build_pattern(ofd_ctx *ctx, int64_t *icnt);
/* the pattern in the reference's 1-based form */
int ofd_get_pattern(ofd_ctx *ctx, int32_t *col, int32_t *row);
/* Checks a caller's pattern against the context's: the col(NNODZu+1),
 * row(icntu) arguments of SparseProdUP_3D (SparseProd.f90:4-7),
 * SparseGMRES_up (SparseGMRES_NS.f90:4-8), SparseProd_LS (SparseProd.f90:
 * 128-132), SparseGMRES_tem (SparseGMRES_tem.f90:1-3), SparseProd_NS_conv
 * (SparseProd.f90:167-172) and SparseGMRES_ns_conv (SparseGMRES_NS_conv.f:
 * 1-6), which index the caller's matrices with them.  OFD_ERR_ARG unless
 * nnode, icnt and every entry of col and row (1-based) equal the context's;
 * the drop-ins call it before they read a matrix of icnt blocks. */
int ofd_check_pattern(ofd_ctx *ctx, int nnode, int64_t icnt, const int32_t *col,
                      const int32_t *row);

/* physics and time-integration parameters (getparam, param.f90:205-490):
 * prm[OFD_NPARAMS] = almi alfi gami beti Delt rho mu NS_kdc gravvec(3) */
int ofd_set_params(ofd_ctx *ctx, const double *prm);
/* ELM_ID_SET(1:2) (param.f90:246-247) */
int ofd_set_subdomains(ofd_ctx *ctx, int32_t id1, int32_t id2);
/* Dirichlet flags IBC(NNODE,ncol), 4 <= ncol <= 8: columns 1..4 = u,v,w,p
 * (BCLhs_3D.f90:4-122), column 6 = T (BCLhs_tem, BCLhs_3D.f90:247-274) */
int ofd_set_ibc(ofd_ctx *ctx, const int32_t *ibc, int ncol);

/* Dirichlet velocity values BCugValu per node, (NNODE,3); used by
 * ofd_apply_dirichlet on the nodes flagged in IBC columns 1..3 */
int ofd_set_bcvals(ofd_ctx *ctx, const double *vals);
/* setBCs_CFD value part (setBCs.f90:84-91): ug = value and the consistent
 * acg = ((ug - ugold)/Delt + (gami-1) acgold)/gami on Dirichlet dofs */
int ofd_apply_dirichlet(ofd_ctx *ctx);

/* The faces of the weak-Dirichlet boundaries (BCugType = 2): bound(b)%F2E
 * and bound(b)%FACE_OR of input.f90:80-92, nface faces, f2e 1-based
 * elements.  Linear tets: the reference's face map (eval_faceshape_tet,
 * FACE_OR 1..4, genGPtritet's 3 points per face).  The reference has no hex
 * face map; here FACE_OR 1..6 of a hex is the face at xi = -1, xi = +1,
 * eta = -1, eta = +1, zeta = -1, zeta = +1 in linhexshl's node order, 2x2
 * Gauss points.  Prisms: OFD_ERR_UNSUPPORTED.  The weak velocity is the mesh
 * velocity (gi = umi, FaceAssembly_weak.f90:147).  Replaces the faces given
 * before; nface = 0 clears them.  The adjoint entries refuse a context with
 * faces (they differentiate the volume residual only). */
int ofd_set_weak_faces(ofd_ctx *ctx, int nface, const int32_t *f2e, const int32_t *face_or);

/* copy a node field between host and the context */
int ofd_set_field(ofd_ctx *ctx, int field, const double *host);
int ofd_get_field(ofd_ctx *ctx, int field, double *host);

/* assembleNavSto(assemble_tensor_flag, subdomain_flag)
 * (assembleNavSto.f90:4-72 -> IntElmAss_NS_overset :76-336, BCLhs_3D,
 * FillSparseMat_3D): alpha stage from the context fields, zeroing, element
 * loop, Dirichlet projection, scatter; then, over the faces registered with
 * ofd_set_weak_faces, FaceAssembly_NS_weak and FaceAssembly_NS_outflow
 * (FaceAssembly_weak.f90:4-409) with their own projection and scatter, as
 * assembleNavSto calls them (:54-60).  LHS/RHS stay on the device.
 * cfl (may be NULL) receives the maximum Gauss-point CFL number. */
int ofd_assemble_ns(ofd_ctx *ctx, int tensor_flag, int subdomain_flag,
                    double *cfl);

/* assembled system in the reference layout */
int ofd_get_rhs(ofd_ctx *ctx, double *rhsu, double *rhsp);
int ofd_get_lhs(ofd_ctx *ctx, double *lhsk11, double *lhsg, double *lhsd1,
                double *lhsm);
/* load a system given in the reference layout (for the drop-in wrappers of
 * SparseProdUP_3D / SparseGMRES_up that receive host matrices) */
int ofd_set_lhs(ofd_ctx *ctx, const double *lhsk11, const double *lhsg,
                const double *lhsd1, const double *lhsm);
int ofd_set_rhs(ofd_ctx *ctx, const double *rhsu, const double *rhsp);

/* SparseProdUP_3D (SparseProd.f90:4-60) on the context matrix:
 * y = A x with x, y in the reference layout (NNODE,3)+(NNODE) */
int ofd_spmv_up(ofd_ctx *ctx, int precision, const double *xu,
                const double *xp, double *yu, double *yp);

/* SparseGMRES_up (SparseGMRES_NS.f90:4-369): one cycle of left
 * block-Jacobi-preconditioned GMRES on the context system, <= kspace
 * iterations, stop when |e_{k+1}| <= tol*|P^-1 b| and k >= kspace_min.
 * sol (reference layout, may be NULL) receives the increment; it is also
 * kept in the context as OFD_SOLU/OFD_SOLP. */
int ofd_gmres_up(ofd_ctx *ctx, int precision, int ortho, double tol,
                 int kspace, int kspace_min, double *solu, double *solp,
                 int *iters, double *resid);

/* Restarted GMRES with an FP64 outer residual (iterative refinement):
 * repeat GMRES(kspace) cycles in `precision` on r = b - A64 x until
 * |b - A64 x|_2 <= rtol*|b|_2 or max_cycles.  A cycle stops early once its
 * preconditioned residual has fallen by inner_tol; a loose inner_tol (1e-2)
 * restarts often and can need several times the cycles of 1e-4.  x stays in
 * the context (OFD_SOLU/OFD_SOLP).  iters = total Arnoldi iterations. */
int ofd_solve(ofd_ctx *ctx, int precision, int ortho, double rtol,
              double inner_tol, int kspace, int max_cycles, int *iters,
              int *cycles, double *relres);

/* relative FP64 residual after each cycle of the last ofd_solve; n gets the
 * number of cycles (at most maxn values are written) */
int ofd_solve_history(ofd_ctx *ctx, double *relres, int maxn, int *n);

/* Newton update (solveOverset.f90:76-78): acg += du; ug += gami*Delt*du;
 * pg += alfi*gami*Delt*dp with du, dp the context solution */
int ofd_newton_update(ofd_ctx *ctx);
/* predictor (driver.f90:83-85) and old <- new (driver.f90:98-103); once
 * quench parameters are set they also predict (Tg = Tgold, rTg =
 * (gami-1)/gami rTgold, likewise phig, rphig) and advance the quench fields */
int ofd_predictor(ofd_ctx *ctx);
int ofd_advance(ofd_ctx *ctx);
/* |RHSGu|_2, |RHSGp|_2 over all nodes (calculate_residual_overset,
 * solveOverset.f90:131-173, one subdomain) */
int ofd_rhs_norms(ofd_ctx *ctx, double *norms2);

/* ---- temperature equation of the quenching solver ---- */

/* material data, q[OFD_NQPARAMS] in the order of the OFD_Q_* enum */
int ofd_set_quench_params(ofd_ctx *ctx, const double *q);

/* assembleQuenching(tensor_flag, ASSEMBLE_FIELD_TEM) (solveQuenching.f90:4-
 * 117) -> IntElmAss_Tem_Quenching_STAB (:579-882) with BCLhs_tem (IBC
 * column 6, BCLhs_3D.f90:247-274) and FillSparseMat_tem (FillSparseMat.f90:
 * 175-201), from the alpha stage of the context's fields (:25-37).  Hex
 * elements; the face term FaceAssembly_ConvectiveH is not part of it (the
 * reference defines it nowhere). */
int ofd_assemble_tem(ofd_ctx *ctx, int tensor_flag);

/* LHSTem(icnt) and RHSGTem(NNODE) in the reference layout */
int ofd_get_tem_system(ofd_ctx *ctx, double *lhs, double *rhs);
int ofd_set_tem_system(ofd_ctx *ctx, const double *lhs, const double *rhs);

/* SparseProd_LS (SparseProd.f90:128-161): y = LHSTem x */
int ofd_spmv_tem(ofd_ctx *ctx, int precision, const double *x, double *y);

/* SparseGMRES_tem (SparseGMRES_tem.f90:1-256): one cycle with the diagonal
 * preconditioner (zero diagonal -> 1) on LHSTem sol = RHSGTem.  sol (may be
 * NULL) receives the increment, 0 when |RHSGTem| = 0; it also stays in the
 * context as OFD_SOLT. */
int ofd_gmres_tem(ofd_ctx *ctx, int precision, int ortho, double tol, int kspace,
                  int kspace_min, double *sol, int *iters, double *resid);

/* restarted SparseGMRES_tem with an FP64 outer residual, as ofd_solve */
int ofd_solve_tem(ofd_ctx *ctx, int precision, int ortho, double rtol, double inner_tol,
                  int kspace, int max_cycles, int *iters, int *cycles, double *relres);

/* temperature Newton update (solveFlow.f90:139-140): rTg += dT,
 * Tg += gami*Delt*dT with dT = OFD_SOLT */
int ofd_tem_update(ofd_ctx *ctx);

/* |RHSGTem|_2 */
int ofd_tem_rhs_norm(ofd_ctx *ctx, double *norm2);

/* ---- coupled NS + VOF system of the quenching solver ---- */

/* assembleQuenching(tensor_flag, ASSEMBLE_FIELD_NS + ASSEMBLE_FIELD_VOF)
 * (solveQuenching.f90:4-117) -> IntElmAss_NSVOF_Quenching_STAB (:122-573)
 * with the 5-field projection BCLhs_NS_conv (BCLhs_3D.f90:278-428; IBC
 * columns 1-3 = u, 7 = p, 8 = phi) and scatter FillSparseMat_NS_conv /
 * LocaltoGlobalNSVOF_3D, from the alpha stage of the context's fields.  Hex
 * elements; the NS discontinuity-capturing coefficient is 0; the face terms
 * of assembleQuenching are not part of it.  cfl (may be NULL) gets the
 * largest Gauss-point CFL number. */
int ofd_assemble_nsvof(ofd_ctx *ctx, int tensor_flag, double *cfl);

/* the reference arrays LHSK11(9,icnt), LHSG, LHSD1, LHSM, LHSls, LHSlsu(3,
 * icnt), LHSPls, LHSUls(3,icnt) and RHSGu(NNODE,3), RHSGp, RHSGls (any
 * pointer may be NULL) */
int ofd_get_nsvof_system(ofd_ctx *ctx, double *k11, double *g, double *d1, double *m,
                         double *ls, double *lsu, double *pls, double *uls, double *rhsu,
                         double *rhsp, double *rhsls);
int ofd_set_nsvof_system(ofd_ctx *ctx, const double *k11, const double *g, const double *d1,
                         const double *m, const double *ls, const double *lsu,
                         const double *pls, const double *uls, const double *rhsu,
                         const double *rhsp, const double *rhsls);

/* SparseProd_NS_conv (SparseProd.f90:167-242) on host vectors u(NNODE,3),
 * p(NNODE), phi(NNODE) */
int ofd_spmv_nsvof(ofd_ctx *ctx, int precision, const double *xu, const double *xp,
                   const double *xl, double *yu, double *yp, double *yl);

/* SparseGMRES_ns_conv (SparseGMRES_NS_conv.f:1-406): one cycle with the
 * 5-field block-Jacobi preconditioner (3x3 momentum inverse, 1/M with
 * 0 -> 1, 1/LS with 0 -> 0).  solu/solp/solls (may be NULL) receive the
 * increment, also kept as OFD_SOLU/OFD_SOLP/OFD_SOLPHI. */
int ofd_gmres_nsvof(ofd_ctx *ctx, int precision, int ortho, double tol, int kspace,
                    int kspace_min, double *solu, double *solp, double *solls, int *iters,
                    double *resid);

/* restarted SparseGMRES_ns_conv with an FP64 outer residual, as ofd_solve;
 * ortho may carry OFD_PREC_RIGHT */
int ofd_solve_nsvof(ofd_ctx *ctx, int precision, int ortho, double rtol, double inner_tol,
                    int kspace, int max_cycles, int *iters, int *cycles, double *relres);

/* Newton update (solveFlow.f90:99-104): acg += du; ug += gami*Delt*du;
 * pg += alfi*gami*Delt*dp; rphig += dphi; phig += gami*Delt*dphi */
int ofd_nsvof_update(ofd_ctx *ctx);

/* |RHSGu|_2, |RHSGp|_2, |RHSGls|_2 */
int ofd_nsvof_rhs_norms(ofd_ctx *ctx, double *norms3);

/* interpolation_tet (interpolation.f90:5-76) on the device.  Receivers are
 * the points of a boundary mesh, xr and dr (nrecv,3) column-major (its xg
 * and dg).  Donors are the context's elements with ELM_ID == src_id in
 * element index order, in the current configuration xg + dg; the first
 * element that contains a receiver wins and ug is interpolated there.
 * Linear tets: the reference's barycentric test (all four >= 0, no
 * tolerance) with its receiver offset xr - x_1 + dr.  Trilinear hexes (no
 * reference counterpart; the reference skips them): local coordinates from
 * 6 Newton steps with the linhexshl shape functions, all in
 * [-1-1e-10, 1+1e-10], among elements whose bounding box, widened by 1e-8
 * of its largest extent, holds xr + dr.  donor[j] = 1-based element index
 * (0: no element holds the receiver); val (nrecv,3) column-major, 0 where
 * there is no donor. */
int ofd_fringe_interp(ofd_ctx *ctx, int32_t src_id, int nrecv, const double *xr,
                      const double *dr, int32_t *donor, double *val);

/* ofd_fringe_interp without allocation.  No counterpart in the reference:
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, serving loops, harnesses that forbid hipMalloc).  Same arguments,
 * results and errors as ofd_fringe_interp; it calls no hipMalloc, hipFree,
 * hipHostMalloc or allocating library routine.  Its scratch is a static
 * device array of 64 MiB per GPU, held until the process ends: per element
 * 48 bytes of box, 192 (hex) or 120 (tet) of geometry and 1 of flag, per
 * receiver 76 bytes, each region rounded up to 256 bytes.  A search that
 * needs more (a hex mesh of more than about 270,000 elements) is refused
 * with OFD_ERR_ARG.  All contexts on one GPU share the array: calls that
 * overlap in time on one GPU race; the array keeps the last call's
 * contents. */
int ofd_fringe_interp_noalloc(ofd_ctx *ctx, int32_t src_id, int nrecv, const double *xr,
                              const double *dr, int32_t *donor, double *val);

/* setBCs_overset step 3 (setBCs.f90:158-165): for j < nrecv, node
 * k = nodes[j] (1-based, distinct) gets IBC(k,1:3) = 1, ug(k,:) = val(j,:)
 * and acg(k,:) = ((val(j,:) - ugold(k,:))/Delt + (gami-1) acgold(k,:))/gami.
 * val is (nrecv,3) column-major. */
int ofd_apply_fringe(ofd_ctx *ctx, int nrecv, const int32_t *nodes, const double *val);

/* kernel timing on the context stream (HIP events around `reps`
 * back-to-back launches); ms = average duration of one launch.
 * ofd_time_assemble times the device work of one ofd_assemble_ns (elements
 * of subdomain flags 1 | 2): the alpha stage, the RHS zeroing, every colour
 * launch of the element kernel and of the weak faces, and the CFL maximum;
 * it leaves out only the entry's host synchronisation and read-back. */
int ofd_time_spmv(ofd_ctx *ctx, int precision, int reps, double *ms);
int ofd_time_assemble(ofd_ctx *ctx, int tensor_flag, int reps, double *ms);
/* the weak-face launches of ofd_assemble_ns alone (k_face_ns_hex, every
 * colour), `reps` times into the assembled arrays: the context holds no
 * valid system afterwards.  Measurement only; no reference counterpart. */
int ofd_time_faces(ofd_ctx *ctx, int tensor_flag, int reps, double *ms);

/* plain device-to-device copy of nbytes (multiple of 16) by a grid-stride
 * kernel; ms = average of `reps` launches.  The bandwidth reference for the
 * HBM-bound kernels (and the calibration of hardware byte counters).
 * Allocates and frees its two buffers.  No reference counterpart. */
int ofd_time_copy(ofd_ctx *ctx, int64_t nbytes, int reps, double *ms);
/* read-only stream of nbytes (multiple of 16, at most 32 GiB) by a
 * grid-stride kernel with 8 loads in flight per thread, one sum stored per
 * block; ms = average of `reps` launches.  The achievable-read reference
 * for the SpMV, whose matrix stream is read-only.  Allocates and frees its
 * buffers.  No reference counterpart. */
int ofd_time_read(ofd_ctx *ctx, int64_t nbytes, int reps, double *ms);

/* ---- adjoint of the NS Newton step (no reference code) ----
 *
 * The residual is the assembled NS right-hand side as the reference computes
 * it (IntElmAss_NS_overset, assembleNavSto.f90:76-336: e3Rhs_3D_fluid with
 * uprime = 0), F(u, ac, p; s) at the alpha stage, before the Dirichlet rows
 * are cleared.  Its exact Jacobian with respect to the Newton increments
 * (acg += d, ug += gami*Delt d, pg += alfi*gami*Delt dp) is
 * J = -dF/d(d, dp); J_bc is J with the Dirichlet rows and columns of IBC
 * replaced as BCLhs_3D does (identity per element sharing the node).  J
 * differs from the assembled LHS (SUPG/PSPG blocks, Picard convection).
 * Trilinear hexes, linear tets and linear prisms. */

/* Transposed product of the assembled LHS: y = LHS^T x, (N,3)+(N) vectors. */
int ofd_spmv_up_t(ofd_ctx *ctx, int precision, const double *xu, const double *xp,
                  double *yu, double *yp);

/* Restarted GMRES on LHS^T lambda = g (FP64 outer residual, as ofd_solve);
 * lambda goes to OFD_LAMU / OFD_LAMP. */
int ofd_solve_t(ofd_ctx *ctx, int precision, int ortho, double rtol, double inner_tol,
                int kspace, int max_cycles, const double *gu, const double *gp, int *iters,
                int *cycles, double *relres);

/* Element VJP of the residual: for lambda (lu (N,3), lp (N)) the gradients
 * of lambda . F with respect to the alpha-stage u, ac, p and to the
 * viscosity scale OFD_MUSCALE (constant 1 when unset), before any Dirichlet
 * projection; elements of the subdomain flag as ofd_assemble_ns.  Any
 * output may be NULL. */
int ofd_vjp_ns(ofd_ctx *ctx, int subdomain_flag, const double *lu, const double *lp,
               double *gu, double *gac, double *gp, double *gmus);

/* y = J_bc^T lambda, matrix-free (element VJP with the Dirichlet projection). */
int ofd_jt_apply(ofd_ctx *ctx, int subdomain_flag, const double *lu, const double *lp,
                 double *yu, double *yp);

/* NS residual form for ofd_assemble_ns and the adjoint: 0 (default) the
 * reference's live call, e3Rhs_3D_fluid with uprime = 0; 1 the residual with
 * the fine-scale velocity uprime = -tauM rLi (e3RHS.f90:31, tauBar = 0),
 * whose continuity rows carry the PSPG term.  The LHS is the same for both.
 * With the reference's form the pressure rows of an enclosed flow may have
 * no exact root (the equal-order element without PSPG); the fine-scale form
 * is stabilized. */
int ofd_set_residual_form(ofd_ctx *ctx, int form);

/* y = J_bc v, v a Newton increment (forward tangent of the residual). */
int ofd_j_apply(ofd_ctx *ctx, int subdomain_flag, const double *vu, const double *vp,
                double *yu, double *yp);

/* J_bc^T lambda = g: restarted flexible GMRES (kspace outer steps, at most
 * max_cycles restarts) on the exact operator, right-preconditioned by LHS^T
 * solved to inner_tol with the restarted device GMRES in `precision` (FP64
 * or mixed); the outer residual and basis are FP64.  iters counts the inner
 * GMRES iterations, cycles the restarts.  lambda goes to OFD_LAMU /
 * OFD_LAMP.  Needs an assembled LHS.  With the reference's residual form
 * J_bc has no pressure-pressure block, so g must not load the pressure modes
 * it leaves free (velocity objectives have g_p = 0). */
int ofd_solve_adjoint(ofd_ctx *ctx, int subdomain_flag, int precision, int ortho, double rtol,
                      double inner_tol, int kspace, int max_cycles, const double *gu,
                      const double *gp, int *iters, int *cycles, double *relres);

/* The consistent Newton correction: J_bc d = RHS (the assembled right-hand
 * side) by the same flexible GMRES, preconditioned by the assembled LHS;
 * the increment goes where ofd_solve puts it, for ofd_newton_update.  With
 * it the Newton iteration converges to the residual's own root, which the
 * LHS of the reference approaches only linearly.  Needs an assembled LHS
 * and RHS. */
int ofd_solve_newton(ofd_ctx *ctx, int subdomain_flag, int precision, int ortho, double rtol,
                     double inner_tol, int kspace, int max_cycles, int *iters, int *cycles,
                     double *relres);

/* ---- adjoint of the quenching temperature equation (no reference code) ----
 *
 * The residual is the assembled RHSGTem of IntElmAss_Tem_Quenching_STAB
 * (solveQuenching.f90:579-882), R(T, rT, u; s) at the alpha stage, before
 * the Dirichlet rows of IBC column 6 are cleared; s is OFD_KSCALE (constant
 * 1 when unset) and the VOF field and mesh velocity are held.  The Newton
 * increments (rTg += d, Tg += gami*Delt*d, solveFlow.f90:139-140) give
 * J_T = -dR/dd; J_T,bc carries the BCLhs_tem projection (BCLhs_3D.f90:
 * 247-274).  J_T equals the assembled LHSTem when TEM_KDC = 0 and
 * C_DMDOT = 1: the LHS leaves out the derivative of the discontinuity-
 * capturing conductivity and scales that of the phase-change source by
 * C_DMDOT.  Vectors are (NNODE).  Trilinear hexes take the hand-written
 * kernel, linear tets and prisms the dual-number kernel of ofd_vjp_tem_dual
 * (the same residual differentiated by forward mode). */

/* y = LHSTem^T x. */
int ofd_spmv_tem_t(ofd_ctx *ctx, int precision, const double *x, double *y);

/* Restarted GMRES on LHSTem^T lambda = g (FP64 outer residual, as
 * ofd_solve_tem); lambda goes to OFD_LAMT. */
int ofd_solve_tem_t(ofd_ctx *ctx, int precision, int ortho, double rtol, double inner_tol,
                    int kspace, int max_cycles, const double *g, int *iters, int *cycles,
                    double *relres);

/* Element VJP of the temperature residual: gradients of lambda . R with
 * respect to the alpha-stage T, rT and u (N,3) and to OFD_KSCALE, before any
 * Dirichlet projection.  Any output may be NULL. */
int ofd_vjp_tem(ofd_ctx *ctx, const double *lam, double *gT, double *grT, double *gu,
                double *gks);

/* y = J_T,bc^T lambda, matrix-free. */
int ofd_jt_apply_tem(ofd_ctx *ctx, const double *lam, double *y);

/* J_T,bc^T lambda = g: restarted flexible GMRES on the exact operator,
 * right-preconditioned by LHSTem^T solved to inner_tol; arguments as
 * ofd_solve_adjoint.  lambda goes to OFD_LAMT.  Needs an assembled LHSTem. */
int ofd_solve_tem_adjoint(ofd_ctx *ctx, int precision, int ortho, double rtol,
                          double inner_tol, int kspace, int max_cycles, const double *g,
                          int *iters, int *cycles, double *relres);

/* ---- exact NS + VOF Jacobian and the coupled quench adjoint (no reference
 * code) ----
 *
 * The residual is the assembled RHSGu/RHSGp/RHSGls of
 * IntElmAss_NSVOF_Quenching_STAB (solveQuenching.f90:122-573), R(u, ac, p,
 * phi, rphi, T) at the alpha stage, fluid elements only, before the
 * Dirichlet rows of IBC columns 1-3, 7, 8 are cleared.  The Newton
 * increments (acg += d, ug += gami*Delt*d, pg += alfi*gami*Delt*dp, rphig +=
 * dphi, phig += gami*Delt*dphi; solveFlow.f90:99-104) give J = -dR/d(d, dp,
 * dphi).  Derivatives come from forward-mode dual numbers through the whole
 * element residual (no hand differentiation); trilinear hexes, linear
 * tets and linear prisms. */

/* Replace the stored NS + VOF matrix (the layout ofd_get_nsvof_system
 * reads) by J_bc: J with the rows and columns of constrained dofs cleared
 * and 1 on their diagonal once per element.  The (phi row, p column) entry,
 * which the layout has no slot for, is zero in J.  ofd_solve_nsvof then
 * solves the exact Newton correction. */
int ofd_assemble_nsvof_jac(ofd_ctx *ctx);

/* Gradients of lambda . R for lambda = (lu (N,3), lp, lphi) with respect to
 * the alpha-stage fields: g is [10][N], rows u (3), ac (3), p, phi, rphi, T;
 * no Dirichlet projection. */
int ofd_vjp_nsvof(ofd_ctx *ctx, const double *lu, const double *lp, const double *lphi,
                  double *g);

/* Gradients of lambda . R_T (the temperature residual of ofd_vjp_tem) by
 * dual numbers, including the VOF field: g is [7][N], rows T, rT, u (3),
 * phi, OFD_KSCALE. */
int ofd_vjp_tem_dual(ofd_ctx *ctx, const double *lam, double *g);

/* y = A^T x for the stored NS + VOF matrix (FP64; (N,3), (N), (N)). */
int ofd_spmv_nsvof_t(ofd_ctx *ctx, const double *xu, const double *xp, const double *xl,
                     double *yu, double *yp, double *yl);

/* A^T lambda = g for the stored NS + VOF matrix: restarted flexible GMRES
 * (FP64 basis) on the exact transposed product, right-preconditioned by the
 * restarted device GMRES on the stored transpose without its (phi row, p
 * column) coupling, solved to inner_tol in `precision` (ortho as
 * ofd_solve_nsvof, OFD_PREC_RIGHT allowed).  lambda goes to lu, lp, ll. */
int ofd_solve_nsvof_t(ofd_ctx *ctx, int precision, int ortho, double rtol, double inner_tol,
                      int kspace, int max_cycles, const double *gu, const double *gp,
                      const double *gl, double *lu, double *lp, double *ll, int *iters,
                      int *cycles, double *relres);

/* Mean device time (HIP events on the context stream) of `reps` back-to-back
 * launches of one adjoint kernel on the current state, all colour launches
 * and the zeroing of its output included: what = OFD_TIME_VJP_NS (raw
 * gradients), OFD_TIME_JT_NS (J_bc^T lambda), OFD_TIME_VJP_TEM,
 * OFD_TIME_JT_TEM.  lambda is the one the last ofd_vjp_ns / ofd_jt_apply
 * (ofd_vjp_tem / ofd_jt_apply_tem) call uploaded.  Measurement only. */
enum { OFD_TIME_VJP_NS = 0, OFD_TIME_JT_NS = 1, OFD_TIME_VJP_TEM = 2, OFD_TIME_JT_TEM = 3 };
int ofd_time_adjoint(ofd_ctx *ctx, int what, int reps, double *ms);

/* Mean device time (HIP events on the context stream) of `reps` back-to-back
 * launches of one quenching-solver kernel: what = OFD_TIME_ASM_NSVOF or
 * OFD_TIME_ASM_TEM (arg = tensor flag; alpha stage, zeroing and the colour
 * launches), OFD_TIME_SPMV_NSVOF or OFD_TIME_SPMV_TEM (arg = precision; needs
 * an assembled system).  Measurement only, no reference counterpart. */
enum { OFD_TIME_ASM_NSVOF = 0, OFD_TIME_ASM_TEM = 1, OFD_TIME_SPMV_NSVOF = 2,
       OFD_TIME_SPMV_TEM = 3 };
int ofd_time_quench(ofd_ctx *ctx, int what, int arg, int reps, double *ms);

/* sampled timing of the Arnoldi product kernel (the preconditioned SpMV of
 * each GMRES iteration) inside ofd_solve / ofd_gmres_up: every `every`-th
 * such launch (counted across cycles and solves) records the kernel's own
 * start and end in HIP timing events on the context stream (0 switches
 * sampling off).  ofd_probe_read returns the number of
 * sampled launches and their summed duration since the last read, and
 * resets both.  Measurement only; no reference counterpart. */
int ofd_probe_set(ofd_ctx *ctx, int every);
int ofd_probe_read(ofd_ctx *ctx, int64_t *count, double *total_ms);

/* device memory footprint of the context in bytes */
int ofd_ctx_bytes(ofd_ctx *ctx, int64_t *bytes);

#ifdef __cplusplus
}
#endif
#endif
