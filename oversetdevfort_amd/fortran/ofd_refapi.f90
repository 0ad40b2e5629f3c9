!======================================================================
! Drop-in replacements for the reference's hot-path routines, keeping
! their names and argument lists (implicit interfaces, explicit-shape
! dummies), implemented over libofd.so through ISO_C_BINDING.
!
! A driver creates the context once (ofd_state_init) and uploads the mesh
! and the Dirichlet flags; after that:
!   * SparseProdUP_3D / SparseGMRES_up take host matrices exactly as the
!     reference (SparseProd.f90:4-7, SparseGMRES_NS.f90:4-8) and upload them
!     on every call (drop-in, correct, not the fast path);
!   * SparseProd_LS, SparseGMRES_tem, SparseProd_NS_conv and
!     SparseGMRES_ns_conv do the same for the quenching solver's temperature
!     and coupled NS + VOF systems;
!   * assembleNavSto_resident(flag, sub) (assembleNavSto.f90:4) assembles
!     from the state held on the device; solveNS_resident solves that system
!     without leaving the device -- the fast path a GPU driver uses.
!======================================================================
module ofd_state
  use ofd_iface
  implicit none
  type(c_ptr), save :: ofd_ctx = c_null_ptr
  integer(c_int), save :: ofd_precision = OFD_FP64
  integer(c_int), save :: ofd_ortho = OFD_ORTHO_MGS
contains
  subroutine ofd_state_init(device)
    integer, intent(in) :: device
    if (c_associated(ofd_ctx)) return
    call ofd_check(ofd_ctx_create(int(device, c_int), ofd_ctx), 'ofd_ctx_create')
  end subroutine ofd_state_init

  subroutine ofd_state_final()
    integer(c_int) :: rc
    if (.not. c_associated(ofd_ctx)) return
    rc = ofd_ctx_destroy(ofd_ctx)
    ofd_ctx = c_null_ptr
  end subroutine ofd_state_final
end module ofd_state

!----------------------------------------------------------------------
! SparseProd.f90:4-60
!----------------------------------------------------------------------
subroutine SparseProdUP_3D(col, row, lhsK11, lhsG, lhsD1, lhsM, &
                           rhstmpu, rhstmpp, prodtmpu, prodtmpp, &
                           D_FLAG, P_FLAG, NNODZu, NSHLu, icntu, NSD)
  use ofd_state
  implicit none
  integer, intent(in) :: NNODZu, NSHLu, icntu, NSD
  integer, intent(in) :: col(NNODZu + 1), row(icntu), D_FLAG(NNODZu), P_FLAG(NNODZu)
  real(8), intent(in) :: lhsK11(NSD*NSD, icntu), lhsG(NSD, icntu), &
                         lhsD1(NSD, icntu), lhsM(icntu), &
                         rhstmpu(NNODZu, NSD), rhstmpp(NNODZu)
  real(8), intent(out) :: prodtmpu(NNODZu, NSD), prodtmpp(NNODZu)
  call ofd_check(ofd_check_pattern(ofd_ctx, int(NNODZu, c_int), int(icntu, c_int64_t), &
                                   col, row), 'SparseProdUP_3D')
  call ofd_check(ofd_set_lhs(ofd_ctx, lhsK11, lhsG, lhsD1, lhsM), 'SparseProdUP_3D')
  call ofd_check(ofd_spmv_up(ofd_ctx, OFD_FP64, rhstmpu, rhstmpp, prodtmpu, &
                             prodtmpp), 'SparseProdUP_3D')
end subroutine SparseProdUP_3D

!----------------------------------------------------------------------
! SparseGMRES_NS.f90:4-369: the increment is added to solu/solp, as the
! reference does (the caller zeroes them first)
!----------------------------------------------------------------------
subroutine SparseGMRES_up(col, row, IBC, IPER, D_FLAG, P_FLAG, &
                          rhsGu, rhsGp, solu, solp, &
                          lhsK11, lhsG, lhsD1, lhsM, &
                          icntu, Utol, Kspaceu, Kspaceu_mn, &
                          NNODZu, NSHLu, NSD)
  use ofd_state
  implicit none
  integer, intent(in) :: icntu, NNODZu, NSHLu, NSD, Kspaceu, Kspaceu_mn
  integer, intent(in) :: col(NNODZu + 1), row(icntu), IBC(NNODZu, 2*NSD + 1), &
                         IPER(NNODZu), D_FLAG(NNODZu), P_FLAG(NNODZu)
  real(8), intent(in) :: rhsGu(NNODZu, NSD), rhsGp(NNODZu), &
                         lhsK11(NSD*NSD, icntu), lhsG(NSD, icntu), &
                         lhsD1(NSD, icntu), lhsM(icntu), Utol
  real(8), intent(inout) :: solu(NNODZu, NSD), solp(NNODZu)
  real(8), allocatable :: du(:, :), dp(:)
  integer(c_int) :: iters
  real(c_double) :: resid
  call ofd_check(ofd_check_pattern(ofd_ctx, int(NNODZu, c_int), int(icntu, c_int64_t), &
                                   col, row), 'SparseGMRES_up')
  allocate (du(NNODZu, NSD), dp(NNODZu))
  call ofd_check(ofd_set_lhs(ofd_ctx, lhsK11, lhsG, lhsD1, lhsM), 'SparseGMRES_up')
  call ofd_check(ofd_set_rhs(ofd_ctx, rhsGu, rhsGp), 'SparseGMRES_up')
  call ofd_check(ofd_gmres_up(ofd_ctx, ofd_precision, ofd_ortho, Utol, &
                              int(Kspaceu, c_int), int(Kspaceu_mn, c_int), du, dp, &
                              iters, resid), 'SparseGMRES_up')
  solu = solu + du
  solp = solp + dp
  deallocate (du, dp)
end subroutine SparseGMRES_up

!----------------------------------------------------------------------
! SparseGMRES_tem.f90:1-256: the increment is added to sol (or sol = 0
! when |rhsG| = 0), as the reference does
!----------------------------------------------------------------------
subroutine SparseGMRES_tem(lhsM, Utol, col, row, rhsG, sol, Kspaceu, Kspaceu_mn, &
                           NNODZu, NSHLu, icntu, NSD)
  use ofd_state
  implicit none
  integer, intent(in) :: NNODZu, NSHLu, icntu, NSD, Kspaceu, Kspaceu_mn
  integer, intent(in) :: col(NNODZu + 1), row(icntu)
  real(8), intent(in) :: lhsM(icntu), rhsG(NNODZu), Utol
  real(8), intent(inout) :: sol(NNODZu)
  real(8), allocatable :: dT(:)
  integer(c_int) :: iters
  real(c_double) :: resid
  call ofd_check(ofd_check_pattern(ofd_ctx, int(NNODZu, c_int), int(icntu, c_int64_t), &
                                   col, row), 'SparseGMRES_tem')
  allocate (dT(NNODZu))
  call ofd_check(ofd_set_tem_system(ofd_ctx, lhsM, rhsG), 'SparseGMRES_tem')
  call ofd_check(ofd_gmres_tem(ofd_ctx, ofd_precision, ofd_ortho, Utol, &
                               int(Kspaceu, c_int), int(Kspaceu_mn, c_int), dT, iters, &
                               resid), 'SparseGMRES_tem')
  if (iters == 0) then
    sol = 0.0d0
  else
    sol = sol + dT
  end if
  deallocate (dT)
end subroutine SparseGMRES_tem

!----------------------------------------------------------------------
! SparseProd.f90:128-161
!----------------------------------------------------------------------
subroutine SparseProd_LS(lhsM, col, row, rhstmp, prodtmp, NNODZu, NSHLu, icntu, NSD)
  use ofd_state
  implicit none
  integer, intent(in) :: NNODZu, NSHLu, icntu, NSD
  integer, intent(in) :: col(NNODZu + 1), row(icntu)
  real(8), intent(in) :: lhsM(icntu), rhstmp(NNODZu)
  real(8), intent(out) :: prodtmp(NNODZu)
  call ofd_check(ofd_check_pattern(ofd_ctx, int(NNODZu, c_int), int(icntu, c_int64_t), &
                                   col, row), 'SparseProd_LS')
  call ofd_check(ofd_set_tem_system(ofd_ctx, lhs=lhsM), 'SparseProd_LS')
  call ofd_check(ofd_spmv_tem(ofd_ctx, OFD_FP64, rhstmp, prodtmp), 'SparseProd_LS')
end subroutine SparseProd_LS

!----------------------------------------------------------------------
! SparseProd.f90:167-242
!----------------------------------------------------------------------
subroutine SparseProd_NS_conv(col, row, lhsK11, lhsG, lhsD1, lhsM, lhsLS, lhsLSu, &
                              LHSPls, LHSUls, rhstmpu, rhstmpp, rhstmpls, &
                              prodtmpu, prodtmpp, prodtmpls, D_FLAG, P_FLAG, &
                              NNODZu, NSHLu, icntu, NSD)
  use ofd_state
  implicit none
  integer, intent(in) :: NNODZu, NSHLu, icntu, NSD
  integer, intent(in) :: col(NNODZu + 1), row(icntu), D_FLAG(NNODZu), P_FLAG(NNODZu)
  real(8), intent(in) :: lhsK11(NSD*NSD, icntu), lhsG(NSD, icntu), lhsD1(NSD, icntu), &
                         lhsM(icntu), lhsLS(icntu), lhsLSu(NSD, icntu), LHSPls(icntu), &
                         LHSUls(NSD, icntu), rhstmpu(NNODZu, NSD), rhstmpp(NNODZu), &
                         rhstmpls(NNODZu)
  real(8), intent(out) :: prodtmpu(NNODZu, NSD), prodtmpp(NNODZu), prodtmpls(NNODZu)
  call ofd_check(ofd_check_pattern(ofd_ctx, int(NNODZu, c_int), int(icntu, c_int64_t), &
                                   col, row), 'SparseProd_NS_conv')
  call ofd_check(ofd_set_nsvof_system(ofd_ctx, lhsK11, lhsG, lhsD1, lhsM, lhsLS, lhsLSu, &
                                      LHSPls, LHSUls), 'SparseProd_NS_conv')
  call ofd_check(ofd_spmv_nsvof(ofd_ctx, OFD_FP64, rhstmpu, rhstmpp, rhstmpls, prodtmpu, &
                                prodtmpp, prodtmpls), 'SparseProd_NS_conv')
end subroutine SparseProd_NS_conv

!----------------------------------------------------------------------
! SparseGMRES_NS_conv.f:1-406: the increment is added to solu/solp/solls,
! as the reference does
!----------------------------------------------------------------------
subroutine SparseGMRES_ns_conv(col, row, IBC, IPER, D_FLAG, P_FLAG, rhsGu, rhsGp, rhsGls, &
                               solu, solp, solls, lhsK11, lhsG, lhsD1, lhsM, lhsls, &
                               LHSlsu, LHSPls, LHSUls, icntu, Utol, Kspaceu, Kspaceu_mn, &
                               NNODZu, NSHLu, NSD)
  use ofd_state
  implicit none
  integer, intent(in) :: icntu, NNODZu, NSHLu, NSD, Kspaceu, Kspaceu_mn
  integer, intent(in) :: col(NNODZu + 1), row(icntu), IBC(NNODZu, 2*NSD + 1), &
                         IPER(NNODZu), D_FLAG(NNODZu), P_FLAG(NNODZu)
  real(8), intent(in) :: rhsGu(NNODZu, NSD), rhsGp(NNODZu), rhsGls(NNODZu), &
                         lhsK11(NSD*NSD, icntu), lhsG(NSD, icntu), lhsD1(NSD, icntu), &
                         lhsM(icntu), lhsls(icntu), LHSlsu(NSD, icntu), LHSPls(icntu), &
                         LHSUls(NSD, icntu), Utol
  real(8), intent(inout) :: solu(NNODZu, NSD), solp(NNODZu), solls(NNODZu)
  real(8), allocatable :: du(:, :), dp(:), dl(:)
  integer(c_int) :: iters
  real(c_double) :: resid
  call ofd_check(ofd_check_pattern(ofd_ctx, int(NNODZu, c_int), int(icntu, c_int64_t), &
                                   col, row), 'SparseGMRES_ns_conv')
  allocate (du(NNODZu, NSD), dp(NNODZu), dl(NNODZu))
  call ofd_check(ofd_set_nsvof_system(ofd_ctx, lhsK11, lhsG, lhsD1, lhsM, lhsls, LHSlsu, &
                                      LHSPls, LHSUls, rhsGu, rhsGp, rhsGls), &
                 'SparseGMRES_ns_conv')
  call ofd_check(ofd_gmres_nsvof(ofd_ctx, ofd_precision, ofd_ortho, Utol, &
                                 int(Kspaceu, c_int), int(Kspaceu_mn, c_int), du, dp, dl, &
                                 iters, resid), 'SparseGMRES_ns_conv')
  solu = solu + du
  solp = solp + dp
  solls = solls + dl
  deallocate (du, dp, dl)
end subroutine SparseGMRES_ns_conv

!----------------------------------------------------------------------
! assembleNavSto.f90:4-72 on the device-resident state: the system stays on
! the device for solveNS_resident (the module-state drop-in with the
! reference's name, which fills LHSK11 ... RHSGp, is in ofd_refmod.f90)
!----------------------------------------------------------------------
subroutine assembleNavSto_resident(assemble_tensor_flag, subdomain_flag)
  use ofd_state
  implicit none
  integer, intent(in) :: assemble_tensor_flag, subdomain_flag
  real(c_double) :: cfl
  call ofd_check(ofd_assemble_ns(ofd_ctx, int(assemble_tensor_flag, c_int), &
                                 int(subdomain_flag, c_int), cfl), 'assembleNavSto')
end subroutine assembleNavSto_resident

!----------------------------------------------------------------------
! device-resident linear solve of the assembled system to rtol (restarted
! GMRES with an FP64 outer residual); the increment stays on the device
! for ofd_newton_update
!----------------------------------------------------------------------
subroutine solveNS_resident(rtol, kspace, iters, relres)
  use ofd_state
  implicit none
  real(8), intent(in) :: rtol
  integer, intent(in) :: kspace
  integer, intent(out) :: iters
  real(8), intent(out) :: relres
  integer(c_int) :: it, cyc
  real(c_double) :: rel, itol
  ! a cycle stops once its preconditioned residual has fallen by itol;
  ! looser values restart early and cost cycles (both precisions)
  itol = 1.0d-4
  call ofd_check(ofd_solve(ofd_ctx, ofd_precision, OFD_ORTHO_CGS1, rtol, itol, &
                           int(kspace, c_int), 100_c_int, it, cyc, rel), &
                 'solveNS_resident')
  iters = it
  relres = rel
end subroutine solveNS_resident
