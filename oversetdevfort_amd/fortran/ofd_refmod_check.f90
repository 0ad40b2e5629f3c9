!======================================================================
! The reference's Newton sequence for one block over its own module state
! (solveOverset.f90:55-78 for one subdomain), every step a drop-in:
!   genSparStruc()                      -> col, row, icnt of the module
!   assembleNavSto(VEC+MAT, 1)          -> RHSGu, RHSGp, LHSK11 ... LHSM
!   SparseGMRES_up(col, row, ..., module arrays)
!   acg += du, ug += gami Delt du, pg += alfi gami Delt dp
!   assembleNavSto(VEC, 1)              -> the residual after the update
! Built against the reference's compiled modules (aAdjKeep, commonvars,
! commonpars) with ofd_refmod.f90; the state comes from a stream file and
! the results go to another, for a test that compares them with the oracle.
!   ofd_refmod_check in.bin out.bin
!======================================================================
program ofd_refmod_check
  use ofd_refmod
  implicit none
  character(len=512) :: fin, fout
  integer :: u, kspace, kmin, i
  integer(c_int32_t) :: hdr(5)
  real(8) :: tol, prm(11)
  real(8), allocatable :: solu(:, :), solp(:), ru0(:, :), rp0(:)

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open (newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
  read (u) hdr
  NNODE = hdr(1)
  NELEM = hdr(2)
  maxNSHL = hdr(3)
  kspace = hdr(4)
  kmin = hdr(5)
  NSD = 3
  read (u) tol, prm
  almi = prm(1); alfi = prm(2); gami = prm(3); beti = prm(4); Delt = prm(5)
  rhoa = prm(6); mua = prm(7); NS_kdc_a = prm(8); gravvec = prm(9:11)
  read (u) ELM_ID_set
  allocate (IEN(NELEM, maxNSHL), ELM_ID(NELEM), xg(NNODE, NSD), IBC(NNODE, 2*NSD + 2))
  allocate (IPER(NNODE), D_FLAG(NNODE), P_FLAG(NNODE))
  read (u) IEN, ELM_ID, xg, IBC
  IPER = [(i, i = 1, NNODE)]
  D_FLAG = 0
  P_FLAG = 1
  allocate (dg(NNODE, NSD), dgold(NNODE, NSD), ug(NNODE, NSD), ugold(NNODE, NSD), &
            ugm(NNODE, NSD), ugmold(NNODE, NSD), acg(NNODE, NSD), acgold(NNODE, NSD), &
            acgm(NNODE, NSD), acgmold(NNODE, NSD), pg(NNODE), pgold(NNODE))
  read (u) dg, dgold, ug, ugold, ugm, ugmold, acg, acgold, acgm, acgmold, pg, pgold
  close (u)

  call ofd_state_init(0)
  call genSparStruc()
  allocate (LHSK11(NSD*NSD, icnt), LHSG(NSD, icnt), LHSD1(NSD, icnt), LHSM(icnt))
  allocate (RHSGu(NNODE, NSD), RHSGp(NNODE), solu(NNODE, NSD), solp(NNODE))
  allocate (ru0(NNODE, NSD), rp0(NNODE))
  call assembleNavSto(ASSEMBLE_TENSOR_VEC + ASSEMBLE_TENSOR_MAT, 1)
  ru0 = RHSGu
  rp0 = RHSGp
  solu = 0.0d0
  solp = 0.0d0
  call SparseGMRES_up(col, row, IBC, IPER, D_FLAG, P_FLAG, RHSGu, RHSGp, solu, solp, &
                      LHSK11, LHSG, LHSD1, LHSM, icnt, tol, kspace, kmin, NNODE, maxNSHL, NSD)
  acg = acg + solu
  ug = ug + gami*Delt*solu
  pg = pg + alfi*gami*Delt*solp
  call assembleNavSto(ASSEMBLE_TENSOR_VEC, 1)

  open (newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  write (u) int(icnt, c_int32_t)
  write (u) col, row
  write (u) LHSK11, LHSG, LHSD1, LHSM, ru0, rp0, solu, solp, RHSGu, RHSGp
  close (u)
  call ofd_state_final()
end program ofd_refmod_check
