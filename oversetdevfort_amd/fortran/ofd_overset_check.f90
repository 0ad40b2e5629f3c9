!======================================================================
! The reference's overset boundary conditions and one staggered Newton
! iteration (solveOverset.f90:55-114) over its own module state, on a tet
! overset pair:
!   1. the reference's own setBCs_overset (setBCs.f90:124-167, compiled
!      here as setBCs_overset_cpu) with the donor search of the
!      interpolation_tet drop-in, subdomain flag 3;
!   2. the setBCs_overset drop-in (ofd_refmod.f90), subdomain flag 3;
!   3. genSparStruc(), then for subdomain 1 and then 2: IBC = 0,
!      setBCs_overset, assembleNavSto(VEC+MAT), SparseGMRES_up on the
!      module arrays, acg += du, ug += gami Delt du, pg += alfi gami Delt dp;
!      then IBC = 0, setBCs_overset(3), assembleNavSto(VEC, 3).
! The state comes from a stream file and the results go to another, for a
! test that compares them with the oracle.
!   ofd_overset_check in.bin out.bin
!======================================================================
program ofd_overset_check
  use ofd_refmod
  implicit none
  character(len=512) :: fin, fout
  integer :: u, kspace, kmin, i, b, n, isub
  integer(c_int32_t) :: hdr(8), fh(2)
  integer(c_int32_t), allocatable :: ibuf(:)
  real(8) :: tol, prm(11)
  real(8), allocatable :: solu(:, :), solp(:), ug0(:, :), acg0(:, :)

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open (newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
  read (u) hdr
  NNODE = hdr(1)
  NELEM = hdr(2)
  maxNSHL = hdr(3)
  kspace = hdr(4)
  kmin = hdr(5)
  NBOUND = hdr(6)
  NSD = 3
  read (u) tol, prm
  almi = prm(1); alfi = prm(2); gami = prm(3); beti = prm(4); Delt = prm(5)
  rhoa = prm(6); mua = prm(7); NS_kdc_a = prm(8); gravvec = prm(9:11)
  Dtgl = 1.0d0/Delt
  read (u) ELM_ID_set, BMESH_ID_set, SRC_ID_set
  allocate (IEN(NELEM, maxNSHL), ELM_ID(NELEM), ELMNSHL(NELEM), NodeID(NNODE), xg(NNODE, NSD))
  read (u) IEN, ELM_ID, NodeID, xg
  ELMNSHL = maxNSHL
  allocate (bound(NBOUND), BCugType(NBOUND, NSD), BCugValu(NBOUND, NSD))
  do b = 1, NBOUND
    read (u) fh
    bound(b)%FACE_ID = fh(1)
    n = fh(2)
    bound(b)%NNODE = n
    allocate (bound(b)%BNODES(n), bound(b)%L2SNODE(n), ibuf(n))
    read (u) ibuf
    bound(b)%BNODES = ibuf
    read (u) ibuf
    bound(b)%L2SNODE = ibuf
    deallocate (ibuf)
    allocate (ibuf(NSD))
    read (u) ibuf
    BCugType(b, :) = ibuf
    deallocate (ibuf)
    read (u) BCugValu(b, :)
  end do
  do isub = 1, NRES
    n = hdr(6 + isub)
    bmesh(isub)%NNODE = n
    bmesh(isub)%NSD = NSD
    allocate (bmesh(isub)%xg(n, NSD), bmesh(isub)%dg(n, NSD), bmesh(isub)%ug(n, NSD))
    read (u) bmesh(isub)%xg, bmesh(isub)%dg
  end do
  allocate (IBC(NNODE, 2*NSD + 2), IPER(NNODE), D_FLAG(NNODE), P_FLAG(NNODE))
  IPER = [(i, i = 1, NNODE)]
  D_FLAG = 0
  P_FLAG = 1
  allocate (dg(NNODE, NSD), dgold(NNODE, NSD), ug(NNODE, NSD), ugold(NNODE, NSD), &
            ugm(NNODE, NSD), ugmold(NNODE, NSD), acg(NNODE, NSD), acgold(NNODE, NSD), &
            acgm(NNODE, NSD), acgmold(NNODE, NSD), pg(NNODE), pgold(NNODE))
  read (u) dg, dgold, ug, ugold, ugm, ugmold, acg, acgold, acgm, acgmold, pg, pgold
  close (u)

  open (newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  call ofd_state_init(0)
  call genSparStruc()
  allocate (ug0(NNODE, NSD), acg0(NNODE, NSD))
  ug0 = ug
  acg0 = acg

  ! 1. the reference's own routine
  IBC = 0
  call setBCs_overset_cpu(3)
  write (u) int(IBC(:, 1:4), c_int32_t)
  write (u) ug, acg, bmesh(1)%ug, bmesh(2)%ug
  ug = ug0
  acg = acg0
  ! 2. the drop-in
  IBC = 0
  call setBCs_overset(3)
  write (u) int(IBC(:, 1:4), c_int32_t)
  write (u) ug, acg, bmesh(1)%ug, bmesh(2)%ug
  ug = ug0
  acg = acg0

  ! 3. one Newton iteration of solveOverset_stag, mesh 1 then mesh 2
  write (u) int(icnt, c_int32_t)
  allocate (LHSK11(NSD*NSD, icnt), LHSG(NSD, icnt), LHSD1(NSD, icnt), LHSM(icnt))
  allocate (RHSGu(NNODE, NSD), RHSGp(NNODE), solu(NNODE, NSD), solp(NNODE))
  do isub = 1, NRES
    IBC = 0
    call setBCs_overset(isub)
    call assembleNavSto(ASSEMBLE_TENSOR_VEC + ASSEMBLE_TENSOR_MAT, isub)
    write (u) int(IBC(:, 1:4), c_int32_t)
    write (u) RHSGu, RHSGp
    solu = 0.0d0
    solp = 0.0d0
    call SparseGMRES_up(col, row, IBC, IPER, D_FLAG, P_FLAG, RHSGu, RHSGp, solu, solp, &
                        LHSK11, LHSG, LHSD1, LHSM, icnt, tol, kspace, kmin, NNODE, maxNSHL, NSD)
    acg = acg + solu
    ug = ug + gami*Delt*solu
    pg = pg + alfi*gami*Delt*solp
    write (u) solu, solp, ug, acg, pg
  end do
  IBC = 0
  call setBCs_overset(3)
  call assembleNavSto(ASSEMBLE_TENSOR_VEC, 3)
  write (u) RHSGu, RHSGp
  close (u)
  call ofd_state_final()
end program ofd_overset_check
