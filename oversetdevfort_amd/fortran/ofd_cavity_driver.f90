!======================================================================
! Fortran time-stepping driver for the lid-driven cavity on the GPU.
!
! Mirrors driver.f90:67-110 / solveOverset_stag (solveOverset.f90:4-126)
! for one block: predictor, Dirichlet values (setBCs_CFD), Newton loop of
! assembleNavSto_resident(VEC+MAT) -> linear solve -> update -> assembleNavSto(VEC)
! -> residual norms, then old <- new.  Everything runs through the
! reference-named wrappers and ISO_C_BINDING on libofd.so.
!
! usage: ofd_cavity_driver [n=8] [nsteps=2] [nnewton=2] [precision=0|1]
! Prints "NEWTON istep inewt |Ru| |Rp| iters relres", the drop-in checks
! "SPMVNORM" / "GMRESNORM", "FIELDNORMS |ug| |acg| |pg|", and then the
! adjoint entries on the final state (no reference counterpart):
! "ADJOINT |lam_u| |lam_p| |dF/du| |dF/dac| |dF/dp| restarts" for the
! velocity load on the free dofs with the fine-scale residual form, and
! "SPMVTNORM" for the transposed product of the assembled LHS.
!======================================================================
program ofd_cavity_driver
  use ofd_state
  implicit none
  integer :: n, nsteps, nnewton, prec, nn, NNODE, NELEM, i, j, k, e, a, istep, inewt
  integer :: iters, kk
  integer(c_int64_t) :: icnt
  integer, allocatable :: IEN(:, :), ELM_ID(:), IBC(:, :), col(:), row(:), iper(:), &
                          dflag(:), pflag(:)
  real(8), allocatable :: xg(:, :), zero3(:, :), zero1(:), bcval(:, :), ug(:, :), &
                          acg(:, :), pg(:), k11(:, :), g(:, :), d1(:, :), m(:), &
                          ru(:, :), rp(:), yu(:, :), yp(:), su(:, :), sp(:), &
                          lu(:, :), lp(:), gu(:, :), gac(:, :), gp(:), gm(:)
  real(8) :: prm(OFD_NPARAMS), rhoinf, almi, alfi, gami, beti, relres, res(2), h
  character(len=32) :: arg
  integer :: di(8), dj(8), dk(8)
  ! linhexshl node order (eval_shape_FEM.f90:161-168) in (x, y, z)
  data di/0, 1, 1, 0, 0, 1, 1, 0/
  data dj/1, 1, 1, 1, 0, 0, 0, 0/
  data dk/0, 0, 1, 1, 0, 0, 1, 1/

  n = 8; nsteps = 2; nnewton = 2; prec = 0
  if (command_argument_count() >= 1) then
    call get_command_argument(1, arg); read (arg, *) n
  end if
  if (command_argument_count() >= 2) then
    call get_command_argument(2, arg); read (arg, *) nsteps
  end if
  if (command_argument_count() >= 3) then
    call get_command_argument(3, arg); read (arg, *) nnewton
  end if
  if (command_argument_count() >= 4) then
    call get_command_argument(4, arg); read (arg, *) prec
  end if

  ! ---- structured unit-cube hex mesh, x fastest ----
  nn = n + 1
  NNODE = nn**3
  NELEM = n**3
  h = 1.0d0/n
  allocate (xg(NNODE, 3), IEN(NELEM, 8), ELM_ID(NELEM), IBC(NNODE, 8), &
            bcval(NNODE, 3), zero3(NNODE, 3), zero1(NNODE))
  do k = 0, n
    do j = 0, n
      do i = 0, n
        a = 1 + i + nn*(j + nn*k)
        xg(a, :) = (/i*h, j*h, k*h/)
      end do
    end do
  end do
  do k = 0, n - 1
    do j = 0, n - 1
      do i = 0, n - 1
        e = 1 + i + n*(j + n*k)
        do a = 1, 8
          IEN(e, a) = 1 + (i + di(a)) + nn*((j + dj(a)) + nn*(k + dk(a)))
        end do
      end do
    end do
  end do
  ELM_ID = 1
  ! no-slip walls, lid u = (1,0,0) on z = 1 (the lid wins on its edges),
  ! pressure fixed at node 1
  IBC = 0
  bcval = 0.0d0
  do a = 1, NNODE
    if (any(abs(xg(a, :)) < 1.0d-12) .or. any(abs(xg(a, :) - 1.0d0) < 1.0d-12)) &
      IBC(a, 1:3) = 1
    if (abs(xg(a, 3) - 1.0d0) < 1.0d-12) bcval(a, 1) = 1.0d0
  end do
  IBC(1, 4) = 1

  ! ---- generalized alpha, rho_inf = 0.5 (param.f90:469-480) ----
  rhoinf = 0.5d0
  almi = (3.0d0 - rhoinf)/(1.0d0 + rhoinf)/2.0d0
  alfi = 1.0d0/(1.0d0 + rhoinf)
  gami = 0.5d0 + almi - alfi
  beti = 0.25d0*(1.0d0 + almi - alfi)*(1.0d0 + almi - alfi)
  prm = (/almi, alfi, gami, beti, 0.01d0, 1.0d0, 0.01d0, 0.0d0, 0.0d0, 0.0d0, -9.81d0/)

  call ofd_state_init(0)
  ofd_precision = int(prec, c_int)
  call ofd_check(ofd_upload_mesh(ofd_ctx, NNODE, NELEM, 8, IEN, ELM_ID, xg), 'upload')
  call ofd_check(ofd_build_pattern(ofd_ctx, icnt), 'genSparStruc')
  allocate (col(NNODE + 1), row(icnt))
  call ofd_check(ofd_get_pattern(ofd_ctx, col, row), 'genSparStruc')
  call ofd_check(ofd_set_params(ofd_ctx, prm), 'getparam')
  call ofd_check(ofd_set_ibc(ofd_ctx, IBC, 8), 'setBCs')
  call ofd_check(ofd_set_bcvals(ofd_ctx, bcval), 'setBCs')
  ! fluid at rest with the wall values applied
  zero3 = 0.0d0
  zero1 = 0.0d0
  allocate (ug(NNODE, 3), acg(NNODE, 3), pg(NNODE))
  ug = 0.0d0
  where (IBC(:, 1:3) == 1) ug = bcval
  call ofd_check(ofd_set_field(ofd_ctx, OFD_UG, ug), 'IC')
  call ofd_check(ofd_set_field(ofd_ctx, OFD_UGOLD, ug), 'IC')

  ! ---- drop-in wrappers on the first assembled system ----
  call ofd_check(ofd_predictor(ofd_ctx), 'predictor')
  call ofd_check(ofd_apply_dirichlet(ofd_ctx), 'setBCs_CFD')
  call assembleNavSto_resident(OFD_ASSEMBLE_VEC + OFD_ASSEMBLE_MAT, 1)
  allocate (k11(9, icnt), g(3, icnt), d1(3, icnt), m(icnt), ru(NNODE, 3), rp(NNODE), &
            yu(NNODE, 3), yp(NNODE), su(NNODE, 3), sp(NNODE), iper(NNODE), &
            dflag(NNODE), pflag(NNODE))
  call ofd_check(ofd_get_lhs(ofd_ctx, k11, g, d1, m), 'get_lhs')
  call ofd_check(ofd_get_rhs(ofd_ctx, ru, rp), 'get_rhs')
  iper = (/(i, i=1, NNODE)/)
  dflag = 0; pflag = 0
  call SparseProdUP_3D(col, row, k11, g, d1, m, ru, rp, yu, yp, dflag, pflag, &
                       NNODE, 8, int(icnt), 3)
  write (*, '(a,2es25.16)') 'SPMVNORM ', sqrt(sum(yu*yu)), sqrt(sum(yp*yp))
  su = 0.0d0; sp = 0.0d0
  call SparseGMRES_up(col, row, IBC, iper, dflag, pflag, ru, rp, su, sp, &
                      k11, g, d1, m, int(icnt), 1.0d-6, 20, 2, NNODE, 8, 3)
  write (*, '(a,2es25.16)') 'GMRESNORM ', sqrt(sum(su*su)), sqrt(sum(sp*sp))

  ! ---- time loop (driver.f90:67-110, one block) ----
  call ofd_check(ofd_set_field(ofd_ctx, OFD_UG, ug), 'IC')
  call ofd_check(ofd_set_field(ofd_ctx, OFD_ACG, zero3), 'IC')
  call ofd_check(ofd_set_field(ofd_ctx, OFD_PG, zero1), 'IC')
  do istep = 1, nsteps
    call ofd_check(ofd_predictor(ofd_ctx), 'predictor')
    call ofd_check(ofd_apply_dirichlet(ofd_ctx), 'setBCs_CFD')
    do inewt = 1, nnewton
      call assembleNavSto_resident(OFD_ASSEMBLE_VEC + OFD_ASSEMBLE_MAT, 1)
      call solveNS_resident(1.0d-10, 30, iters, relres)
      call ofd_check(ofd_newton_update(ofd_ctx), 'update')
      call assembleNavSto_resident(OFD_ASSEMBLE_VEC, 1)
      call ofd_check(ofd_rhs_norms(ofd_ctx, res), 'residual')
      write (*, '(a,2i4,2es25.16,i6,es12.4)') 'NEWTON ', istep, inewt, res(1), res(2), &
        iters, relres
    end do
    call ofd_check(ofd_advance(ofd_ctx), 'advance')
  end do
  call ofd_check(ofd_get_field(ofd_ctx, OFD_UG, ug), 'get ug')
  call ofd_check(ofd_get_field(ofd_ctx, OFD_ACG, acg), 'get acg')
  call ofd_check(ofd_get_field(ofd_ctx, OFD_PG, pg), 'get pg')
  write (*, '(a,3es25.16)') 'FIELDNORMS ', sqrt(sum(ug*ug)), sqrt(sum(acg*acg)), &
    sqrt(sum(pg*pg))

  ! ---- adjoint on the final state (no reference counterpart) ----
  allocate (lu(NNODE, 3), lp(NNODE), gu(NNODE, 3), gac(NNODE, 3), gp(NNODE), gm(NNODE))
  call ofd_check(ofd_set_residual_form(ofd_ctx, 1), 'residual form')
  call assembleNavSto_resident(OFD_ASSEMBLE_VEC + OFD_ASSEMBLE_MAT, 1)
  su = ug
  where (IBC(:, 1:3) == 1) su = 0.0d0
  sp = 0.0d0
  call ofd_check(ofd_solve_adjoint(ofd_ctx, 1, OFD_FP64, OFD_ORTHO_CGS1, 1.0d-10, 1.0d-3, &
                                   30, 20, su, sp, iters, kk, relres), 'adjoint')
  call ofd_check(ofd_get_field(ofd_ctx, OFD_LAMU, lu), 'lambda')
  call ofd_check(ofd_get_field(ofd_ctx, OFD_LAMP, lp), 'lambda')
  call ofd_check(ofd_vjp_ns(ofd_ctx, 1, lu, lp, gu, gac, gp, gm), 'VJP')
  write (*, '(a,5es25.16,i4)') 'ADJOINT ', sqrt(sum(lu*lu)), sqrt(sum(lp*lp)), &
    sqrt(sum(gu*gu)), sqrt(sum(gac*gac)), sqrt(sum(gp*gp)), kk
  call ofd_check(ofd_spmv_up_t(ofd_ctx, OFD_FP64, su, lp, yu, yp), 'transposed product')
  write (*, '(a,2es25.16)') 'SPMVTNORM ', sqrt(sum(yu*yu)), sqrt(sum(yp*yp))
  kk = 0
  call ofd_state_final()
end program ofd_cavity_driver
