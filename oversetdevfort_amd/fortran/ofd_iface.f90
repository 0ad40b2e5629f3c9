!======================================================================
! ISO_C_BINDING interfaces to libofd.so (include/ofd.h).
!
! Every entry returns an integer status (OFD_OK = 0).  Arrays are passed
! in the reference's own layouts: IEN(NELEM,NSHL), IBC(NNODE,ncol),
! node fields (NNODE,3), pg(NNODE), col(NNODE+1), row(icnt),
! LHSK11(9,icnt), LHSG(3,icnt), LHSD1(3,icnt), LHSM(icnt).
!======================================================================
module ofd_iface
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: OFD_OK = 0, OFD_ERR_ARG = 1, OFD_ERR_HIP = 2, &
                               OFD_ERR_NEGDETJ = 3, OFD_ERR_NONFINITE = 4, &
                               OFD_ERR_NOCONV = 5
  integer(c_int), parameter :: OFD_ASSEMBLE_VEC = 2, OFD_ASSEMBLE_MAT = 4
  integer(c_int), parameter :: OFD_FP64 = 0, OFD_FP32 = 1
  integer(c_int), parameter :: OFD_ORTHO_MGS = 0, OFD_ORTHO_CGS2 = 1, OFD_ORTHO_CGS1 = 2
  integer(c_int), parameter :: OFD_PREC_RIGHT = 16  ! ior() into ortho of ofd_solve_nsvof
  integer(c_int), parameter :: OFD_XG = 0, OFD_DG = 1, OFD_DGOLD = 2, &
                               OFD_UG = 3, OFD_UGOLD = 4, OFD_UGM = 5, &
                               OFD_UGMOLD = 6, OFD_ACG = 7, OFD_ACGOLD = 8, &
                               OFD_ACGM = 9, OFD_ACGMOLD = 10, OFD_PG = 11, &
                               OFD_PGOLD = 12, OFD_SOLU = 13, OFD_SOLP = 14, &
                               OFD_PHIG = 15, OFD_PHIGOLD = 16, OFD_RPHIG = 17, &
                               OFD_RPHIGOLD = 18, OFD_TG = 19, OFD_TGOLD = 20, &
                               OFD_RTG = 21, OFD_RTGOLD = 22, OFD_SOLT = 23, &
                               OFD_SOLPHI = 24, OFD_LAMU = 25, OFD_LAMP = 26, &
                               OFD_MUSCALE = 27, OFD_KSCALE = 28, OFD_LAMT = 29
  integer(c_int), parameter :: OFD_NPARAMS = 11, OFD_NQPARAMS = 21

  interface
    function ofd_last_error() bind(C, name='ofd_last_error') result(p)
      import :: c_ptr
      type(c_ptr) :: p
    end function

    function ofd_ctx_create(device, ctx) bind(C, name='ofd_ctx_create') result(rc)
      import :: c_int, c_ptr
      integer(c_int), value :: device
      type(c_ptr), intent(out) :: ctx
      integer(c_int) :: rc
    end function

    function ofd_ctx_destroy(ctx) bind(C, name='ofd_ctx_destroy') result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int) :: rc
    end function

    function ofd_upload_mesh(ctx, nnode, nelem, nshl, ien, elm_id, xg) &
        bind(C, name='ofd_upload_mesh') result(rc)
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: nnode, nelem, nshl
      integer(c_int32_t), intent(in) :: ien(*), elm_id(*)
      real(c_double), intent(in) :: xg(*)
      integer(c_int) :: rc
    end function

    function ofd_build_pattern(ctx, icnt) bind(C, name='ofd_build_pattern') result(rc)
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int64_t), intent(out) :: icnt
      integer(c_int) :: rc
    end function

    function ofd_get_pattern(ctx, col, row) bind(C, name='ofd_get_pattern') result(rc)
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), intent(out) :: col(*), row(*)
      integer(c_int) :: rc
    end function

    function ofd_check_pattern(ctx, nnode, icnt, col, row) &
        bind(C, name='ofd_check_pattern') result(rc)
      import :: c_int, c_ptr, c_int32_t, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int), value :: nnode
      integer(c_int64_t), value :: icnt
      integer(c_int32_t), intent(in) :: col(*), row(*)
      integer(c_int) :: rc
    end function

    function ofd_set_params(ctx, prm) bind(C, name='ofd_set_params') result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: prm(*)
      integer(c_int) :: rc
    end function

    function ofd_set_subdomains(ctx, id1, id2) bind(C, name='ofd_set_subdomains') &
        result(rc)
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: id1, id2
      integer(c_int) :: rc
    end function

    function ofd_set_ibc(ctx, ibc, ncol) bind(C, name='ofd_set_ibc') result(rc)
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), intent(in) :: ibc(*)
      integer(c_int), value :: ncol
      integer(c_int) :: rc
    end function

    function ofd_set_bcvals(ctx, vals) bind(C, name='ofd_set_bcvals') result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: vals(*)
      integer(c_int) :: rc
    end function

    function ofd_apply_dirichlet(ctx) bind(C, name='ofd_apply_dirichlet') result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int) :: rc
    end function

    function ofd_set_field(ctx, field, host) bind(C, name='ofd_set_field') result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: field
      real(c_double), intent(in) :: host(*)
      integer(c_int) :: rc
    end function

    function ofd_get_field(ctx, field, host) bind(C, name='ofd_get_field') result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: field
      real(c_double), intent(out) :: host(*)
      integer(c_int) :: rc
    end function

    function ofd_assemble_ns(ctx, tensor_flag, subdomain_flag, cfl) &
        bind(C, name='ofd_assemble_ns') result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: tensor_flag, subdomain_flag
      real(c_double), intent(out) :: cfl
      integer(c_int) :: rc
    end function

    function ofd_get_rhs(ctx, rhsu, rhsp) bind(C, name='ofd_get_rhs') result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: rhsu(*), rhsp(*)
      integer(c_int) :: rc
    end function

    function ofd_set_rhs(ctx, rhsu, rhsp) bind(C, name='ofd_set_rhs') result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: rhsu(*), rhsp(*)
      integer(c_int) :: rc
    end function

    function ofd_get_lhs(ctx, k11, g, d1, m) bind(C, name='ofd_get_lhs') result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: k11(*), g(*), d1(*), m(*)
      integer(c_int) :: rc
    end function

    function ofd_set_lhs(ctx, k11, g, d1, m) bind(C, name='ofd_set_lhs') result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: k11(*), g(*), d1(*), m(*)
      integer(c_int) :: rc
    end function

    function ofd_spmv_up(ctx, precision, xu, xp, yu, yp) bind(C, name='ofd_spmv_up') &
        result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: precision
      real(c_double), intent(in) :: xu(*), xp(*)
      real(c_double), intent(out) :: yu(*), yp(*)
      integer(c_int) :: rc
    end function

    function ofd_gmres_up(ctx, precision, ortho, tol, kspace, kspace_min, solu, solp, &
                          iters, resid) bind(C, name='ofd_gmres_up') result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: precision, ortho, kspace, kspace_min
      real(c_double), value :: tol
      real(c_double), intent(out) :: solu(*), solp(*)
      integer(c_int), intent(out) :: iters
      real(c_double), intent(out) :: resid
      integer(c_int) :: rc
    end function

    function ofd_solve(ctx, precision, ortho, rtol, inner_tol, kspace, max_cycles, &
                       iters, cycles, relres) bind(C, name='ofd_solve') result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: precision, ortho, kspace, max_cycles
      real(c_double), value :: rtol, inner_tol
      integer(c_int), intent(out) :: iters, cycles
      real(c_double), intent(out) :: relres
      integer(c_int) :: rc
    end function

    function ofd_newton_update(ctx) bind(C, name='ofd_newton_update') result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int) :: rc
    end function

    function ofd_predictor(ctx) bind(C, name='ofd_predictor') result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int) :: rc
    end function

    function ofd_advance(ctx) bind(C, name='ofd_advance') result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int) :: rc
    end function

    function ofd_set_quench_params(ctx, q) bind(C, name='ofd_set_quench_params') result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: q(*)
      integer(c_int) :: rc
    end function

    function ofd_assemble_tem(ctx, tensor_flag) bind(C, name='ofd_assemble_tem') result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: tensor_flag
      integer(c_int) :: rc
    end function

    function ofd_get_tem_system(ctx, lhs, rhs) bind(C, name='ofd_get_tem_system') result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out), optional :: lhs(*), rhs(*)
      integer(c_int) :: rc
    end function

    ! an absent optional array reaches the library as a null pointer: that
    ! part of the system is left as it is
    function ofd_set_tem_system(ctx, lhs, rhs) bind(C, name='ofd_set_tem_system') result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in), optional :: lhs(*), rhs(*)
      integer(c_int) :: rc
    end function

    function ofd_spmv_tem(ctx, precision, x, y) bind(C, name='ofd_spmv_tem') result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: precision
      real(c_double), intent(in) :: x(*)
      real(c_double), intent(out) :: y(*)
      integer(c_int) :: rc
    end function

    function ofd_gmres_tem(ctx, precision, ortho, tol, kspace, kspace_min, sol, iters, &
                           resid) bind(C, name='ofd_gmres_tem') result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: precision, ortho, kspace, kspace_min
      real(c_double), value :: tol
      real(c_double), intent(out) :: sol(*)
      integer(c_int), intent(out) :: iters
      real(c_double), intent(out) :: resid
      integer(c_int) :: rc
    end function

    function ofd_solve_tem(ctx, precision, ortho, rtol, inner_tol, kspace, max_cycles, &
                           iters, cycles, relres) bind(C, name='ofd_solve_tem') result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: precision, ortho, kspace, max_cycles
      real(c_double), value :: rtol, inner_tol
      integer(c_int), intent(out) :: iters, cycles
      real(c_double), intent(out) :: relres
      integer(c_int) :: rc
    end function

    function ofd_tem_update(ctx) bind(C, name='ofd_tem_update') result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int) :: rc
    end function

    function ofd_tem_rhs_norm(ctx, norm2) bind(C, name='ofd_tem_rhs_norm') result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: norm2
      integer(c_int) :: rc
    end function

    function ofd_assemble_nsvof(ctx, tensor_flag, cfl) bind(C, name='ofd_assemble_nsvof') &
        result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: tensor_flag
      real(c_double), intent(out) :: cfl
      integer(c_int) :: rc
    end function

    function ofd_get_nsvof_system(ctx, k11, g, d1, m, ls, lsu, pls, uls, rhsu, rhsp, &
                                  rhsls) bind(C, name='ofd_get_nsvof_system') result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out), optional :: k11(*), g(*), d1(*), m(*), ls(*), lsu(*), &
                                               pls(*), uls(*), rhsu(*), rhsp(*), rhsls(*)
      integer(c_int) :: rc
    end function

    function ofd_set_nsvof_system(ctx, k11, g, d1, m, ls, lsu, pls, uls, rhsu, rhsp, &
                                  rhsls) bind(C, name='ofd_set_nsvof_system') result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in), optional :: k11(*), g(*), d1(*), m(*), ls(*), lsu(*), &
                                              pls(*), uls(*), rhsu(*), rhsp(*), rhsls(*)
      integer(c_int) :: rc
    end function

    function ofd_spmv_nsvof(ctx, precision, xu, xp, xl, yu, yp, yl) &
        bind(C, name='ofd_spmv_nsvof') result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: precision
      real(c_double), intent(in) :: xu(*), xp(*), xl(*)
      real(c_double), intent(out) :: yu(*), yp(*), yl(*)
      integer(c_int) :: rc
    end function

    function ofd_gmres_nsvof(ctx, precision, ortho, tol, kspace, kspace_min, solu, solp, &
                             solls, iters, resid) bind(C, name='ofd_gmres_nsvof') result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: precision, ortho, kspace, kspace_min
      real(c_double), value :: tol
      real(c_double), intent(out) :: solu(*), solp(*), solls(*)
      integer(c_int), intent(out) :: iters
      real(c_double), intent(out) :: resid
      integer(c_int) :: rc
    end function

    function ofd_solve_nsvof(ctx, precision, ortho, rtol, inner_tol, kspace, max_cycles, &
                             iters, cycles, relres) bind(C, name='ofd_solve_nsvof') result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: precision, ortho, kspace, max_cycles
      real(c_double), value :: rtol, inner_tol
      integer(c_int), intent(out) :: iters, cycles
      real(c_double), intent(out) :: relres
      integer(c_int) :: rc
    end function

    function ofd_nsvof_update(ctx) bind(C, name='ofd_nsvof_update') result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int) :: rc
    end function

    function ofd_nsvof_rhs_norms(ctx, norms3) bind(C, name='ofd_nsvof_rhs_norms') result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: norms3(3)
      integer(c_int) :: rc
    end function

    function ofd_fringe_interp(ctx, src_id, nrecv, xr, dr, donor, val) &
        bind(C, name='ofd_fringe_interp') result(rc)
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: src_id
      integer(c_int), value :: nrecv
      real(c_double), intent(in) :: xr(*), dr(*)
      integer(c_int32_t), intent(out) :: donor(*)
      real(c_double), intent(out) :: val(*)
      integer(c_int) :: rc
    end function

    ! bound(b)%F2E / %FACE_OR of the weak-Dirichlet (BCugType 2) boundaries
    function ofd_set_weak_faces(ctx, nface, f2e, face_or) &
        bind(C, name='ofd_set_weak_faces') result(rc)
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int), value :: nface
      integer(c_int32_t), intent(in) :: f2e(*), face_or(*)
      integer(c_int) :: rc
    end function
    function ofd_apply_fringe(ctx, nrecv, nodes, val) bind(C, name='ofd_apply_fringe') &
        result(rc)
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: nrecv
      integer(c_int32_t), intent(in) :: nodes(*)
      real(c_double), intent(in) :: val(*)
      integer(c_int) :: rc
    end function

    function ofd_rhs_norms(ctx, norms2) bind(C, name='ofd_rhs_norms') result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: norms2(2)
      integer(c_int) :: rc
    end function

    ! ---- adjoint (no reference counterpart); (NNODE,3) + (NNODE) vectors ----
    function ofd_spmv_up_t(ctx, precision, xu, xp, yu, yp) bind(C, name='ofd_spmv_up_t') &
        result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: precision
      real(c_double), intent(in) :: xu(*), xp(*)
      real(c_double), intent(out) :: yu(*), yp(*)
      integer(c_int) :: rc
    end function

    function ofd_solve_t(ctx, precision, ortho, rtol, inner_tol, kspace, max_cycles, gu, gp, &
                         iters, cycles, relres) bind(C, name='ofd_solve_t') result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: precision, ortho, kspace, max_cycles
      real(c_double), value :: rtol, inner_tol
      real(c_double), intent(in) :: gu(*), gp(*)
      integer(c_int), intent(out) :: iters, cycles
      real(c_double), intent(out) :: relres
      integer(c_int) :: rc
    end function

    function ofd_vjp_ns(ctx, subdomain_flag, lu, lp, gu, gac, gp, gmus) &
        bind(C, name='ofd_vjp_ns') result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: subdomain_flag
      real(c_double), intent(in) :: lu(*), lp(*)
      real(c_double), intent(out) :: gu(*), gac(*), gp(*), gmus(*)
      integer(c_int) :: rc
    end function

    function ofd_jt_apply(ctx, subdomain_flag, lu, lp, yu, yp) bind(C, name='ofd_jt_apply') &
        result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: subdomain_flag
      real(c_double), intent(in) :: lu(*), lp(*)
      real(c_double), intent(out) :: yu(*), yp(*)
      integer(c_int) :: rc
    end function

    function ofd_j_apply(ctx, subdomain_flag, vu, vp, yu, yp) bind(C, name='ofd_j_apply') &
        result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: subdomain_flag
      real(c_double), intent(in) :: vu(*), vp(*)
      real(c_double), intent(out) :: yu(*), yp(*)
      integer(c_int) :: rc
    end function

    function ofd_set_residual_form(ctx, form) bind(C, name='ofd_set_residual_form') &
        result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: form
      integer(c_int) :: rc
    end function

    function ofd_solve_adjoint(ctx, subdomain_flag, precision, ortho, rtol, inner_tol, &
                               kspace, max_cycles, gu, gp, iters, cycles, relres) &
        bind(C, name='ofd_solve_adjoint') result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: subdomain_flag, precision, ortho, kspace, max_cycles
      real(c_double), value :: rtol, inner_tol
      real(c_double), intent(in) :: gu(*), gp(*)
      integer(c_int), intent(out) :: iters, cycles
      real(c_double), intent(out) :: relres
      integer(c_int) :: rc
    end function

    function ofd_solve_newton(ctx, subdomain_flag, precision, ortho, rtol, inner_tol, &
                              kspace, max_cycles, iters, cycles, relres) &
        bind(C, name='ofd_solve_newton') result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: subdomain_flag, precision, ortho, kspace, max_cycles
      real(c_double), value :: rtol, inner_tol
      integer(c_int), intent(out) :: iters, cycles
      real(c_double), intent(out) :: relres
      integer(c_int) :: rc
    end function

    function ofd_spmv_tem_t(ctx, precision, x, y) bind(C, name='ofd_spmv_tem_t') result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: precision
      real(c_double), intent(in) :: x(*)
      real(c_double), intent(out) :: y(*)
      integer(c_int) :: rc
    end function

    function ofd_solve_tem_t(ctx, precision, ortho, rtol, inner_tol, kspace, max_cycles, g, &
                             iters, cycles, relres) bind(C, name='ofd_solve_tem_t') result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: precision, ortho, kspace, max_cycles
      real(c_double), value :: rtol, inner_tol
      real(c_double), intent(in) :: g(*)
      integer(c_int), intent(out) :: iters, cycles
      real(c_double), intent(out) :: relres
      integer(c_int) :: rc
    end function

    function ofd_vjp_tem(ctx, lam, gT, grT, gu, gks) bind(C, name='ofd_vjp_tem') result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: lam(*)
      real(c_double), intent(out) :: gT(*), grT(*), gu(*), gks(*)
      integer(c_int) :: rc
    end function

    function ofd_jt_apply_tem(ctx, lam, y) bind(C, name='ofd_jt_apply_tem') result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: lam(*)
      real(c_double), intent(out) :: y(*)
      integer(c_int) :: rc
    end function

    function ofd_solve_tem_adjoint(ctx, precision, ortho, rtol, inner_tol, kspace, &
                                   max_cycles, g, iters, cycles, relres) &
        bind(C, name='ofd_solve_tem_adjoint') result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: precision, ortho, kspace, max_cycles
      real(c_double), value :: rtol, inner_tol
      real(c_double), intent(in) :: g(*)
      integer(c_int), intent(out) :: iters, cycles
      real(c_double), intent(out) :: relres
      integer(c_int) :: rc
    end function
  end interface

contains

  ! status check: print the library message and stop, as the reference
  ! stops on fatal errors (eval_shape.f90:35-36, param.f90:595-598)
  subroutine ofd_check(rc, where)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: where
    character(kind=c_char), pointer :: msg(:)
    character(len=512) :: text
    integer :: i
    if (rc == OFD_OK) return
    text = ''
    call c_f_pointer(ofd_last_error(), msg, [512])
    do i = 1, 512
      if (msg(i) == c_null_char) exit
      text(i:i) = msg(i)
    end do
    write (*, '(a,a,a,i0,a,a)') 'ofd error in ', where, ' (', rc, '): ', trim(text)
    error stop 911
  end subroutine ofd_check

end module ofd_iface
