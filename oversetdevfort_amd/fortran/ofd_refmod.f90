!======================================================================
! Drop-ins that read and write the reference's module state (aAdjKeep,
! commonvars, commonpars).  They compile only inside the reference's own
! build, after modules.f90 and next to ofd_iface.f90 / ofd_refapi.f90:
!
!   * assembleQuenching(assemble_tensor_flag, assemble_field_flag)
!     (solveQuenching.f90:4-120): copies the module's node fields, flags and
!     material to the device, assembles the NS + VOF and/or temperature
!     systems there (ofd_assemble_nsvof, ofd_assemble_tem) and copies the
!     assembled parts back into LHSK11 ... LHSTem, RHSGu ... RHSGtem.  The
!     reference's face terms (FaceAssembly_NS_weak_CF, FaceAssembly_NS_outflow,
!     FaceAssembly_ConvectiveH) are not part of this drop-in.
!   * interpolation_tet(mesh_id) (interpolation.f90:5-76): fills
!     bmesh(mesh_id)%ug from the device donor search over the elements with
!     ELM_ID == SRC_ID_SET(mesh_id).  The reference's own setBCs_overset
!     runs unchanged over it as well.
!   * setBCs_overset(subdomain_flag) (setBCs.f90:124-167): setBCs_CFD as
!     the reference, the other subdomain's nodes constrained, the donor
!     search and the receivers' ug, acg and IBC written on the device.
!   * genSparStruc() (genSparStruc.f90:4-47): uploads the module mesh (IEN,
!     ELM_ID, xg) and builds the pattern on the device, then allocates and
!     fills the module's col, row and icnt.
!   * assembleNavSto(assemble_tensor_flag, subdomain_flag)
!     (assembleNavSto.f90:4-72): copies the module node fields, IBC,
!     ELM_ID_SET and the parameters to the device, assembles there
!     (ofd_assemble_ns) and fills the module's RHSGu/RHSGp and
!     LHSK11/LHSG/LHSD1/LHSM the flags name, zeroed first as the reference
!     does.  The weak faces are the ones registered with ofd_set_weak_faces.
!
! With these, solveOverset_stag (solveOverset.f90:4-121) runs unchanged over
! the module state: genSparStruc once, then per Newton iteration
! setBCs_overset, assembleNavSto, SparseGMRES_up (ofd_refapi.f90, on the
! module arrays) and the update.  setBCs_overset replaces the reference's
! when setBCs.f90 is compiled with -DsetBCs_overset=setBCs_overset_cpu.
!======================================================================
module ofd_refmod
  use ofd_state
  use aAdjKeep
  use commonvars
  use commonpars
  implicit none
contains
  ! module node fields, IBC, the generalized-alpha constants and the quench
  ! material (param.f90:222-257) -> the device context
  subroutine ofd_push_quench_state()
    real(8) :: prm(OFD_NPARAMS), q(OFD_NQPARAMS)
    prm = (/almi, alfi, gami, beti, Delt, rhow, muw, NS_kdc_w, gravvec(1), gravvec(2), &
            gravvec(3)/)
    call ofd_check(ofd_set_params(ofd_ctx, prm), 'ofd_push_quench_state')
    q = (/rhoa, rhow, mua, muw, cpa, cpw, kappaa, kappaw, rhos, mus, cps, kappas, Ts, lh, &
          c_cond, c_evap, C_DMDOT, TEM_kdc, Tref, 102.0d0, LSC_kdc/)
    call ofd_check(ofd_set_quench_params(ofd_ctx, q), 'ofd_push_quench_state')
    call ofd_check(ofd_set_ibc(ofd_ctx, IBC, int(size(IBC, 2), c_int)), 'ofd_push_quench_state')
    call push(OFD_XG, xg)
    call push(OFD_DG, dg)
    call push(OFD_DGOLD, dgold)
    call push(OFD_UG, ug)
    call push(OFD_UGOLD, ugold)
    call push(OFD_UGM, ugm)
    call push(OFD_UGMOLD, ugmold)
    call push(OFD_ACG, acg)
    call push(OFD_ACGOLD, acgold)
    call push(OFD_ACGM, acgm)
    call push(OFD_ACGMOLD, acgmold)
    call push(OFD_PG, pg)
    call push(OFD_PGOLD, pgold)
    call push(OFD_PHIG, phig)
    call push(OFD_PHIGOLD, phigold)
    call push(OFD_RPHIG, rphig)
    call push(OFD_RPHIGOLD, rphigold)
    call push(OFD_TG, Tg)
    call push(OFD_TGOLD, Tgold)
    call push(OFD_RTG, rTg)
    call push(OFD_RTGOLD, rTgold)
  end subroutine ofd_push_quench_state

  ! module node fields, IBC, ELM_ID_SET and the NS parameters of
  ! IntElmAss_NS_overset (the air-phase rhoa, mua, NS_kdc_a of
  ! assembleNavSto.f90:252-254, gravvec) -> the device context
  subroutine ofd_push_ns_state()
    real(8) :: prm(OFD_NPARAMS)
    prm = (/almi, alfi, gami, beti, Delt, rhoa, mua, NS_kdc_a, gravvec(1), gravvec(2), &
            gravvec(3)/)
    call ofd_check(ofd_set_params(ofd_ctx, prm), 'ofd_push_ns_state')
    call ofd_check(ofd_set_subdomains(ofd_ctx, int(ELM_ID_SET(1), c_int32_t), &
                                      int(ELM_ID_SET(2), c_int32_t)), 'ofd_push_ns_state')
    call ofd_check(ofd_set_ibc(ofd_ctx, IBC, int(min(size(IBC, 2), 8), c_int)), &
                   'ofd_push_ns_state')
    call push(OFD_DG, dg)
    call push(OFD_DGOLD, dgold)
    call push(OFD_UG, ug)
    call push(OFD_UGOLD, ugold)
    call push(OFD_UGM, ugm)
    call push(OFD_UGMOLD, ugmold)
    call push(OFD_ACG, acg)
    call push(OFD_ACGOLD, acgold)
    call push(OFD_ACGM, acgm)
    call push(OFD_ACGMOLD, acgmold)
    call push(OFD_PG, pg)
  end subroutine ofd_push_ns_state

  subroutine push(field, a)
    integer(c_int), intent(in) :: field
    real(8), intent(in) :: a(*)
    call ofd_check(ofd_set_field(ofd_ctx, field, a), 'ofd_push_quench_state')
  end subroutine push
end module ofd_refmod

!----------------------------------------------------------------------
! solveQuenching.f90:4-120 (without the face terms)
!----------------------------------------------------------------------
subroutine assembleQuenching(assemble_tensor_flag, assemble_field_flag)
  use ofd_refmod
  implicit none
  integer, intent(in) :: assemble_tensor_flag, assemble_field_flag
  logical :: vec, mat
  real(c_double) :: cfl
  vec = iand(assemble_tensor_flag, ASSEMBLE_TENSOR_VEC) > 0
  mat = iand(assemble_tensor_flag, ASSEMBLE_TENSOR_MAT) > 0
  ! the reference zeroes every quench RHS / LHS array the flags name,
  ! whichever fields it then assembles
  if (vec) then
    RHSGu = 0.0d0
    RHSGp = 0.0d0
    RHSGls = 0.0d0
    RHSGtem = 0.0d0
  end if
  if (mat) then
    LHSK11 = 0.0d0
    LHSG = 0.0d0
    LHSD1 = 0.0d0
    LHSM = 0.0d0
    LHSLS = 0.0d0
    LHSULS = 0.0d0
    LHSLSU = 0.0d0
    LHSPLS = 0.0d0
    LHSTem = 0.0d0
  end if
  call ofd_push_quench_state()
  if (iand(assemble_field_flag, ASSEMBLE_FIELD_NS + ASSEMBLE_FIELD_VOF) > 0) then
    call ofd_check(ofd_assemble_nsvof(ofd_ctx, int(assemble_tensor_flag, c_int), cfl), &
                   'assembleQuenching')
    if (mat) call ofd_check(ofd_get_nsvof_system(ofd_ctx, k11=LHSK11, g=LHSG, d1=LHSD1, &
                                                 m=LHSM, ls=LHSLS, lsu=LHSLSU, pls=LHSPLS, &
                                                 uls=LHSULS), 'assembleQuenching')
    if (vec) call ofd_check(ofd_get_nsvof_system(ofd_ctx, rhsu=RHSGu, rhsp=RHSGp, &
                                                 rhsls=RHSGls), 'assembleQuenching')
  end if
  if (iand(assemble_field_flag, ASSEMBLE_FIELD_TEM) > 0) then
    call ofd_check(ofd_assemble_tem(ofd_ctx, int(assemble_tensor_flag, c_int)), &
                   'assembleQuenching')
    if (mat) call ofd_check(ofd_get_tem_system(ofd_ctx, lhs=LHSTem), 'assembleQuenching')
    if (vec) call ofd_check(ofd_get_tem_system(ofd_ctx, rhs=RHSGtem), 'assembleQuenching')
  end if
end subroutine assembleQuenching

!----------------------------------------------------------------------
! interpolation.f90:5-76: receivers are bmesh(mesh_id)%xg + %dg; donors
! the elements with ELM_ID == SRC_ID_SET(mesh_id), first in index order;
! receivers without a donor get 0, as the reference leaves them
!----------------------------------------------------------------------
subroutine interpolation_tet(mesh_id)
  use ofd_refmod
  implicit none
  integer, intent(in) :: mesh_id
  integer(c_int32_t), allocatable :: donor(:)
  integer :: n
  n = bmesh(mesh_id)%NNODE
  allocate (donor(n))
  call ofd_check(ofd_set_field(ofd_ctx, OFD_DG, dg), 'interpolation_tet')
  call ofd_check(ofd_set_field(ofd_ctx, OFD_UG, ug), 'interpolation_tet')
  call ofd_check(ofd_fringe_interp(ofd_ctx, int(SRC_ID_SET(mesh_id), c_int32_t), &
                                   int(n, c_int), bmesh(mesh_id)%xg, bmesh(mesh_id)%dg, &
                                   donor, bmesh(mesh_id)%ug), 'interpolation_tet')
  deallocate (donor)
end subroutine interpolation_tet

!----------------------------------------------------------------------
! genSparStruc.f90:4-47 (with Asadj :53-105) on the device: col(NNODE+1),
! row(icnt) and icnt of the module, 1-based, the reference's order
!----------------------------------------------------------------------
subroutine genSparStruc()
  use ofd_refmod
  implicit none
  integer(c_int64_t) :: n
  call ofd_check(ofd_upload_mesh(ofd_ctx, int(NNODE, c_int), int(NELEM, c_int), &
                                 int(maxNSHL, c_int), IEN, ELM_ID, xg), 'genSparStruc')
  call ofd_check(ofd_build_pattern(ofd_ctx, n), 'genSparStruc')
  icnt = int(n)
  if (allocated(col)) deallocate (col)
  if (allocated(row)) deallocate (row)
  allocate (col(NNODE + 1), row(icnt))
  call ofd_check(ofd_get_pattern(ofd_ctx, col, row), 'genSparStruc')
end subroutine genSparStruc

!----------------------------------------------------------------------
! assembleNavSto.f90:4-72 over the module state: alpha stage, zeroing,
! IntElmAss_NS_overset, the weak and outflow faces, on the device; the
! assembled RHSGu/RHSGp and LHSK11/LHSG/LHSD1/LHSM come back into the module
!----------------------------------------------------------------------
subroutine assembleNavSto(assemble_tensor_flag, subdomain_flag)
  use ofd_refmod
  implicit none
  integer, intent(in) :: assemble_tensor_flag, subdomain_flag
  real(c_double) :: cfl
  call ofd_push_ns_state()
  call ofd_check(ofd_assemble_ns(ofd_ctx, int(assemble_tensor_flag, c_int), &
                                 int(subdomain_flag, c_int), cfl), 'assembleNavSto')
  if (iand(assemble_tensor_flag, ASSEMBLE_TENSOR_VEC) > 0) &
    call ofd_check(ofd_get_rhs(ofd_ctx, RHSGu, RHSGp), 'assembleNavSto')
  if (iand(assemble_tensor_flag, ASSEMBLE_TENSOR_MAT) > 0) &
    call ofd_check(ofd_get_lhs(ofd_ctx, LHSK11, LHSG, LHSD1, LHSM), 'assembleNavSto')
end subroutine assembleNavSto

!----------------------------------------------------------------------
! setBCs_overset(subdomain_flag) (setBCs.f90:124-167) with the fringe on
! the device.  setBCs_CFD runs as in the reference (it stays in setBCs.f90,
! compiled with setBCs_overset renamed).  For every subdomain isub the flag
! selects: the other subdomain's nodes (NodeID == ELM_ID_SET(3-isub)) get
! IBC(:,1:4) = 1; the face whose FACE_ID is BMESH_ID_SET(isub) names the
! receivers; the donor search (interpolation_tet) fills bmesh(isub)%ug; and
! ofd_apply_fringe sets IBC(k,1:3) = 1, ug and the Dirichlet-consistent
! acg = ((ug - ugold) Dtgl + (gami - 1) acgold) / gami of each receiver k
! on the device, from which ug and acg return into the module.  A selected
! subdomain without its face stops with a message (the reference would
! index bound(0)).
!----------------------------------------------------------------------
subroutine setBCs_overset(subdomain_flag)
  use ofd_refmod
  implicit none
  integer, intent(in) :: subdomain_flag
  integer :: isub, b, ifac, i, n
  integer(c_int32_t), allocatable :: nodes(:)
  real(8), allocatable :: val(:, :)
  call setBCs_CFD()
  if (iand(subdomain_flag, 3) == 0) return
  call ofd_push_ns_state()
  do isub = 1, NRES
    if (iand(subdomain_flag, ishft(1, isub - 1)) == 0) cycle
    do i = 1, NNODE
      if (NodeID(i) == ELM_ID_SET(3 - isub)) IBC(i, 1:4) = 1
    end do
    ifac = 0
    do b = 1, NBOUND
      if (bound(b)%FACE_ID == BMESH_ID_SET(isub)) then
        ifac = b
        exit
      end if
    end do
    if (ifac == 0) then
      write (*, *) 'setBCs_overset: no boundary with FACE_ID = BMESH_ID_SET(', isub, ')'
      error stop 911
    end if
    call interpolation_tet(isub)
    n = bound(ifac)%NNODE
    allocate (nodes(n), val(n, NSD))
    do i = 1, n
      nodes(i) = int(bound(ifac)%BNODES(i), c_int32_t)
      val(i, :) = bmesh(isub)%ug(bound(ifac)%L2SNODE(i), :)
      IBC(bound(ifac)%BNODES(i), 1:3) = 1
    end do
    call ofd_check(ofd_apply_fringe(ofd_ctx, int(n, c_int), nodes, val), 'setBCs_overset')
    deallocate (nodes, val)
    ! the next subdomain's donors see these receivers' new values, as in the
    ! reference's loop
    call ofd_check(ofd_get_field(ofd_ctx, OFD_UG, ug), 'setBCs_overset')
    call ofd_check(ofd_get_field(ofd_ctx, OFD_ACG, acg), 'setBCs_overset')
  end do
end subroutine setBCs_overset
