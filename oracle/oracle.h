/*
 * CPU oracle for the OversetDevFort hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * This library is a plain-C restatement of the reference Fortran routines on
 * the data-parallel path (shape functions, quadrature, VMS/SUPG element
 * assembly, Dirichlet projection, CSR pattern and scatter, block SpMV,
 * block-Jacobi GMRES, tet fringe interpolation).  It exists so that tests,
 * smoke() and the CPU-baseline leg of bench.py can check and time the HIP
 * library against it.  The product library never links or calls it.
 *
 * Conventions follow the reference's in-memory layouts, but with 0-based C
 * indices:
 *   node fields     u(NNODE,3) column-major  -> u[c*nnode + i]
 *   connectivity    IEN(NELEM,NSHL)          -> ien[a*nelem + e]  (0-based node ids)
 *   IBC(NNODE,ncol)                          -> ibc[c*nnode + i]
 *   CSR pattern     col(NNODE+1), row(icnt)  -> col[i]..col[i+1]-1, row[k] (0-based)
 *   LHSK11(9,icnt)  -> k11[k*9 + r*3 + c]   (row-major 3x3, index (r-1)*3+c)
 *   LHSG(3,icnt), LHSD1(3,icnt), LHSM(icnt)
 *   element locals  xKebe11(9,NSHL,NSHL) -> ke[(b*nshl + a)*9 + i]
 */
#ifndef OFD_ORACLE_H
#define OFD_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* generalized-alpha constants: out = {almi, alfi, gami, beti} */
void orc_genalpha(double rhoinf, double out[4]);

/* volume quadrature, gp(ngp,3) column-major, gw(ngp); returns 0 or -1 */
int orc_gauss_volume(int ngp, double *gp, double *gw);

/* reference-element shape functions; shgradl(NSHL,3) column-major */
void orc_linhexshl(const double gp[3], double *shl, double *shgradl);
void orc_lintetshl(const double gp[3], double *shl, double *shgradl);
void orc_linprishl(const double gp[3], double *shl, double *shgradl);

/* 3x3 inverse by cofactors; returns DetJ = 1/(1/det) as the reference does */
void orc_inverse3x3(const double *A, double *Ainv, double *detj);

/* eval_shape_fem: xl,dl (NSHL,3) col-major; outputs col-major 3x3 */
int orc_eval_shape_fem(int nshl, const double gp[3], const double *xl,
                       const double *dl, double *shlu, double *shgradgu,
                       double *dxidx, double *Gij, double *Ginv, double *detj);

/* stabilisation parameters */
void orc_stab_ns(const double *Gij, double dt, const double *uadv, double rho,
                 double mu, double *tauM, double *tauP, double *tauC);

/* genSparStruc: pass col=NULL,row=NULL to get icnt only */
long orc_sparsity(int nnode, int nelem, int nshl, const int *ien, int *col,
                  int *row);

/* binary search of SparseMatLoc_3D over row[lo:hi) */
int orc_sparse_loc(const int *row, int lo, int hi, int goal);

/*
 * IntElmAss_NS_overset + BCLhs_3D + FillSparseMat_3D + RHS scatter.
 * prm = {almi, alfi, gami, beti, delt, rho, mu, kdc, gx, gy, gz}; mus (N) a
 * per-node viscosity scale (mu at a Gauss point = mu * sum_a N_a mus_a) or
 * NULL for the reference's constant mu; vms 0 is the live call's residual
 * (uprime = 0), 1 the residual with uprime = -tauM rLi (an extension)
 * Caller zeroes (or not) the outputs, exactly like assembleNavSto does.
 * cfl_out receives the maximum Gauss-point CFL number.
 */
int orc_assemble_ns(int nnode, int nelem, int nshl, int ngauss,
                    const int *ien, const int *elm_id, const int *elm_id_set,
                    int subdomain_flag, int tensor_flag, const double *xg,
                    const int *ibc, int ibc_ncol, const double *dgA,
                    const double *ugA, const double *ugmA, const double *acgA,
                    const double *acgmA, const double *pgA, const double *mus,
                    const double *prm, int vms, const int *col, const int *row,
                    double *k11, double *g, double *d1, double *m, double *rhsu,
                    double *rhsp, double *cfl_out);

/*
 * FaceAssembly_NS_weak then FaceAssembly_NS_outflow (FaceAssembly_weak.f90)
 * over the nface faces (f2e 0-based elements, face_or their orientations,
 * ngb points per face), each with BCLhs_3D and the scatter, accumulated into
 * the outputs (the reference adds them after IntElmAss_NS_overset).  Tets:
 * the reference's face rule (ngb 1 or 3, face_or 1..4); hexes (an
 * extension): face_or 1..6 = xi_1 -1/+1, xi_2 -1/+1, xi_3 -1/+1, ngb 4.
 */
int orc_face_ns(int nnode, int nelem, int nshl, const int *ien, const int *elm_id,
                const int *elm_id_set, int subdomain_flag, int tensor_flag,
                const double *xg, const int *ibc, const double *dgA, const double *ugA,
                const double *ugmA, const double *pgA, const double *prm, int nface,
                const int *f2e, const int *face_or, int ngb, const int *col,
                const int *row, double *k11, double *g, double *d1, double *m,
                double *rhsu, double *rhsp);

/* one element only, before BC (for kernel-level tests) */
int orc_element_ns(int nshl, int ngauss, const double *xl, const double *dl,
                   const double *ul, const double *uml, const double *acl,
                   const double *acml, const double *pl, const double *musl,
                   const double *prm, int vms, int tensor_flag, double *ke, double *ge,
                   double *de, double *me, double *ru, double *rp, double *cfl);

/* BCLhs_3D on element locals; ibcl(NSHL,4) col-major */
void orc_bclhs_3d(int nshl, const int *ibcl, double *ke, double *ge, double *de,
                  double *me, double *ru, double *rp);

/* SparseProdUP_3D; x,y as u(N,3) SoA + p(N) */
void orc_spmv_up(int nnode, const int *col, const int *row, const double *k11,
                 const double *g, const double *d1, const double *m,
                 const double *xu, const double *xp, double *yu, double *yp);

/*
 * SparseGMRES_up, single MGS cycle.  sol is accumulated into (caller zeroes).
 * Returns the number of Arnoldi iterations; final |e_{k+1}| in *resid,
 * the preconditioned initial norm in *rnorm0.
 */
int orc_gmres_up(int nnode, const int *col, const int *row, const double *k11,
                 const double *g, const double *d1, const double *m,
                 const double *rhsu, const double *rhsp, double *solu,
                 double *solp, double tol, int kspace, int kspace_min,
                 double *resid, double *rnorm0);

/*
 * interpolation_tet: donor search on tet source elements with
 * elm_id == src_id, for nrecv receiver points (xr,dr: (nrecv,3) SoA).
 * donor[j] = first element index that contains point j (or -1), val(nrecv,3).
 */
void orc_interp_tet(int nnode, int nelem, int nshl, const int *ien,
                    const int *elm_id, int src_id, const double *xg,
                    const double *dg, const double *ug, int nrecv,
                    const double *xr, const double *dr, int *donor,
                    double *val);

/*
 * Extension to trilinear hexes (the reference searches tets only): donor =
 * first element (index order) with elm_id == src_id whose widened bounding
 * box holds p = xr + dr and whose 6-step Newton local coordinates lie in
 * [-1-1e-10, 1+1e-10]^3; val = trilinear interpolation of ug.
 */
void orc_interp_hex(int nnode, int nelem, const int *ien, const int *elm_id,
                    int src_id, const double *xg, const double *dg,
                    const double *ug, int nrecv, const double *xr,
                    const double *dr, int *donor, double *val);
int orc_hex_local(const double X[8][3], const double p[3], double xi[3]);

/*
 * Temperature equation of the quenching solver (solveQuenching.f90:579-882):
 * IntElmAss_Tem_Quenching_STAB + BCLhs_tem + FillSparseMat_tem.
 * prm = the NS parameter vector (almi, alfi, gami, beti, delt, ...);
 * qp = {rhoa, rhow, mua, muw, cpa, cpw, kappaa, kappaw, rhos, mus, cps,
 *       kappas, Ts, lh, c_cond, c_evap, c_dmdot, tem_kdc, tref, fluid_id};
 * ibc_t(NNODE) = IBC(:,6).  ksA (NULL: none) scales the conductivity per
 * node, kappa at a Gauss point times sum_a N_a ks_a (an extension of the
 * reference, for the conductivity gradient).  Returns the number of
 * negative Jacobians.
 */
int orc_assemble_tem(int nnode, int nelem, int nshl, int ngauss, const int *ien,
                     const int *elm_id, int tensor_flag, const double *xg,
                     const int *ibc_t, const double *dgA, const double *ugA,
                     const double *ugmA, const double *phigA, const double *TgA,
                     const double *rTgA, const double *ksA, const double *prm,
                     const double *qp, const int *col, const int *row, double *lhs,
                     double *rhs);
int orc_element_tem(int nshl, int ngauss, const double *xl, const double *dl,
                    const double *ul, const double *uml, const double *phil,
                    const double *Tl, const double *rTl, const double *ksl,
                    const double *prm, const double *qp, int is_fluid, int tensor_flag,
                    double *te, double *rt);
/* SparseProd_LS: y = A x, scalar CSR */
void orc_spmv_s(int nnode, const int *col, const int *row, const double *A, const double *x,
                double *y);
/* SparseGMRES_tem: one MGS cycle, diagonal preconditioner */
int orc_gmres_tem(int nnode, const int *col, const int *row, const double *A,
                  const double *b, double *sol, double tol, int kspace, int kspace_min,
                  double *resid);

/*
 * Coupled NS + VOF assembly (solveQuenching.f90:122-573) with the 5-field
 * projection BCLhs_NS_conv and scatter FillSparseMat_NS_conv;
 * ibc8(NNODE,8) = IBC; ex = {gx, gy, gz, lsc_kdc}; outputs in the reference
 * layouts (k11 9, g/d1/lsu/uls 3, m/ls/pls 1 per nonzero; rhsu (N,3),
 * rhsp, rhsls (N)).  cfl_out may be NULL.
 */
int orc_assemble_nsvof(int nnode, int nelem, int nshl, int ngauss, const int *ien,
                       const int *elm_id, int tensor_flag, const double *xg, const int *ibc8,
                       const double *dgA, const double *ugA, const double *ugmA,
                       const double *acgA, const double *pgA, const double *phigA,
                       const double *rphigA, const double *TgA, const double *prm,
                       const double *qp, const double *ex, const int *col, const int *row,
                       double *k11, double *g, double *d1, double *m, double *ls, double *lsu,
                       double *pls, double *uls, double *rhsu, double *rhsp, double *rhsls,
                       double *cfl_out);
/* SparseProd_NS_conv */
void orc_spmv_ns_conv(int nnode, const int *col, const int *row, const double *k11,
                      const double *g, const double *d1, const double *m, const double *ls,
                      const double *lsu, const double *pls, const double *uls,
                      const double *xu, const double *xp, const double *xl, double *yu,
                      double *yp, double *yl);
/* SparseGMRES_ns_conv: one MGS cycle, sol accumulates */
int orc_gmres_ns_conv(int nnode, const int *col, const int *row, const double *k11,
                      const double *g, const double *d1, const double *m, const double *ls,
                      const double *lsu, const double *pls, const double *uls,
                      const double *bu, const double *bp, const double *bl, double *solu,
                      double *solp, double *soll, double tol, int kspace, int kspace_min,
                      double *resid);

/* host baseline: threads of the OpenMP build (n > 0 sets them; returns the
 * count in use, 1 for the parity build) */
int orc_set_threads(int n);
/* the NS element loop colour by colour, the elements of a colour in parallel
 * in the OpenMP build (constant viscosity, the reference's residual form) */
int orc_assemble_ns_col(int nnode, int nelem, int nshl, int ngauss, const int *ien,
                        const int *elm_id, const int *elm_id_set, int subdomain_flag,
                        int tensor_flag, const double *xg, const int *ibc, const double *dgA,
                        const double *ugA, const double *ugmA, const double *acgA,
                        const double *acgmA, const double *pgA, const double *prm,
                        const int *col, const int *row, int ncolor, const int *color_off,
                        const int *color_elems, double *k11, double *g, double *d1,
                        double *m, double *rhsu, double *rhsp, double *cfl_out);

#ifdef __cplusplus
}
#endif
#endif
