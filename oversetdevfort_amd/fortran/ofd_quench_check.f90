!======================================================================
! The quenching solver's linear-algebra drop-ins called as the reference
! calls them: SparseProd_LS, SparseGMRES_tem (temperature system),
! SparseProd_NS_conv, SparseGMRES_ns_conv (coupled NS + VOF system), all
! through ISO_C_BINDING on libofd.so.
!
! usage: ofd_quench_check in.bin out.bin
! in.bin (stream, Fortran order): NNODE NELEM icnt kspace kmin (int32),
!   tol (real8), IEN(NELEM,8) (1-based), ELM_ID(NELEM), xg(NNODE,3),
!   lhsT(icnt), rhsT(NNODE), xT(NNODE),
!   K11(9,icnt), G(3,icnt), D1(3,icnt), M(icnt), LS(icnt), LSU(3,icnt),
!   PLS(icnt), ULS(3,icnt), rhsu(NNODE,3), rhsp(NNODE), rhsls(NNODE),
!   xu(NNODE,3), xp(NNODE), xl(NNODE)
! out.bin: prodT, solT, produ, prodp, prodl, solu, solp, solls
!======================================================================
program ofd_quench_check
  use ofd_state
  implicit none
  integer :: NNODE, NELEM, icnt, kspace, kmin, u
  integer(c_int64_t) :: icnt_dev
  real(8) :: tol
  integer, allocatable :: IEN(:, :), ELM_ID(:), col(:), row(:), ibc(:, :), iper(:), &
                          dflag(:), pflag(:)
  real(8), allocatable :: xg(:, :), lhsT(:), rhsT(:), xT(:), prodT(:), solT(:), &
                          k11(:, :), g(:, :), d1(:, :), m(:), ls(:), lsu(:, :), pls(:), &
                          uls(:, :), rhsu(:, :), rhsp(:), rhsls(:), xu(:, :), xp(:), xl(:), &
                          pu(:, :), pp(:), pl(:), su(:, :), sp(:), sl(:)
  character(len=512) :: fin, fout

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open (newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
  read (u) NNODE, NELEM, icnt, kspace, kmin
  read (u) tol
  allocate (IEN(NELEM, 8), ELM_ID(NELEM), xg(NNODE, 3), lhsT(icnt), rhsT(NNODE), &
            xT(NNODE), k11(9, icnt), g(3, icnt), d1(3, icnt), m(icnt), ls(icnt), &
            lsu(3, icnt), pls(icnt), uls(3, icnt), rhsu(NNODE, 3), rhsp(NNODE), &
            rhsls(NNODE), xu(NNODE, 3), xp(NNODE), xl(NNODE))
  read (u) IEN, ELM_ID, xg, lhsT, rhsT, xT, k11, g, d1, m, ls, lsu, pls, uls, rhsu, rhsp, &
    rhsls, xu, xp, xl
  close (u)

  call ofd_state_init(0)
  call ofd_check(ofd_upload_mesh(ofd_ctx, NNODE, NELEM, 8, IEN, ELM_ID, xg), 'input')
  call ofd_check(ofd_build_pattern(ofd_ctx, icnt_dev), 'genSparStruc')
  if (icnt_dev /= icnt) error stop 'pattern size'
  allocate (col(NNODE + 1), row(icnt), ibc(NNODE, 7), iper(NNODE), dflag(NNODE), &
            pflag(NNODE))
  call ofd_check(ofd_get_pattern(ofd_ctx, col, row), 'genSparStruc')
  ibc = 0; iper = 0; dflag = 0; pflag = 0

  allocate (prodT(NNODE), solT(NNODE), pu(NNODE, 3), pp(NNODE), pl(NNODE), &
            su(NNODE, 3), sp(NNODE), sl(NNODE))
  call SparseProd_LS(lhsT, col, row, xT, prodT, NNODE, 8, icnt, 3)
  solT = 0.0d0
  call SparseGMRES_tem(lhsT, tol, col, row, rhsT, solT, kspace, kmin, NNODE, 8, icnt, 3)
  call SparseProd_NS_conv(col, row, k11, g, d1, m, ls, lsu, pls, uls, xu, xp, xl, &
                          pu, pp, pl, dflag, pflag, NNODE, 8, icnt, 3)
  su = 0.0d0; sp = 0.0d0; sl = 0.0d0
  call SparseGMRES_ns_conv(col, row, ibc, iper, dflag, pflag, rhsu, rhsp, rhsls, &
                           su, sp, sl, k11, g, d1, m, ls, lsu, pls, uls, icnt, tol, &
                           kspace, kmin, NNODE, 8, 3)

  open (newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  write (u) prodT, solT, pu, pp, pl, su, sp, sl
  close (u)
  call ofd_state_final()
end program ofd_quench_check
