!======================================================================
! Harness around the reference Fortran routines -- TEST INFRASTRUCTURE.
!
! Linked (by oracle/Makefile, target `ref`) against the reference's own
! object files compiled from their sources where they lie; nothing of the
! reference is copied.  It replaces driver.f90's main program: it fills the
! reference modules' state from a binary input file and calls the routines
! on the hot path directly, writing their outputs to a binary file.  It is
! used only to write golden vectors (tests/golden/make_golden.py) and, where
! built, as the CPU baseline.
!
! usage: ofd_refharness <in.bin> <out.bin>
!
! modes (first int32 of in.bin):
!   1  shape:  linhexshl / lintetshl / linprishl / genGPandGWtetpri /
!              get_inverse_3x3 / eval_shape_fem at given points
!   2  ns:     genSparStruc, IntElmAss_NS_overset + BCLhs_3D + FillSparseMat_3D,
!              SparseProdUP_3D on a given vector, SparseGMRES_up on the
!              assembled system
!   3  interp: interpolation_tet for receiver points
!   4  tem:    genSparStruc, IntElmAss_Tem_Quenching_STAB (with BCLhs_tem and
!              FillSparseMat_tem), SparseProd_LS on a given vector,
!              SparseGMRES_tem on the assembled system
!   7  io:     input_mesh (part.<id>.dat) and readSol (restart.<r>.1) in
!              the working directory, everything they read written out;
!              then writeSol (restart.<w>.1) from given fields
!   6  face:   genSparStruc, then FaceAssembly_NS_weak and
!              FaceAssembly_NS_outflow (as assembleNavSto calls them) over
!              one weak boundary (BCugType 2) of given tet faces
!   5  nsvof:  genSparStruc, IntElmAss_NSVOF_Quenching_STAB with the
!              5-field Dirichlet projection and scatter BCLhs_NS_conv /
!              FillSparseMat_NS_conv, SparseProd_NS_conv on a given vector,
!              SparseGMRES_ns_conv on the assembled system
!
! IntElmAss_NSVOF_Quenching_STAB calls BCLhs_3D and FillSparseMat_3D with
! the 5-field argument lists, which the reference's routines of those names
! do not take (solveQuenching.f90:526-534).  oracle/Makefile compiles
! solveQuenching.f90 with those two call names mapped to the adapters at
! the end of this file, which reorder the arguments and call the
! reference's own 5-field routines, BCLhs_NS_conv and FillSparseMat_NS_conv,
! with a zero second divergence block.
!======================================================================
program ofd_refharness
  use aAdjKeep
  use commonvars
  use commonpars
  use mpi
  implicit none

  character(len=512) :: fin, fout
  integer :: mode, ierr

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)

  call MPI_INIT(ierr)
  call MPI_COMM_SIZE(MPI_COMM_WORLD, numnodes, ierr)
  call MPI_COMM_RANK(MPI_COMM_WORLD, myid, ierr)
  ismaster = (myid == 0)
  iga = .false.
  NSD = 3
  NBOUND = 0

  open (10, file=trim(fin), access='stream', form='unformatted', status='old')
  ! unit 82: input_mesh opens unit 11 for the mesh file
  open (82, file=trim(fout), access='stream', form='unformatted', &
        status='replace')
  read (10) mode
  select case (mode)
  case (1)
    call run_shape()
  case (2)
    call run_ns()
  case (3)
    call run_interp()
  case (4)
    call run_tem()
  case (5)
    call run_nsvof()
  case (6)
    call run_face()
  case (7)
    call run_io()
  case default
    write (*, *) 'unknown mode', mode
  end select
  close (10)
  close (82)
  call MPI_FINALIZE(ierr)

contains

  subroutine run_shape()
    integer :: npt, ip, q
    real(8), allocatable :: pts(:, :)
    real(8) :: shl(8), sg(8, 3), gp8(8, 3), gw8(8), gp4(4, 3), gw4(4), &
               gp6(6, 3), gw6(6), gp1(1, 3), gw1(1)
    real(8) :: A(3, 3), Ai(3, 3), dj
    real(8) :: xl(6, 3), dl(6, 3), shlu(6), shgradgu(6, 3), dxidx(3, 3), &
               Gij(3, 3), Ginv(3, 3)
    integer :: ntet, npri
    real(8), allocatable :: xt(:, :, :), xp(:, :, :)
    real(8) :: xl4(4, 3), dl4(4, 3), shlu4(4), shgradgu4(4, 3)

    read (10) npt
    allocate (pts(npt, 3))
    read (10) pts
    ! reference-element shape functions at each point
    do ip = 1, npt
      call linhexshl(pts(ip, :), 8, 3, shl, sg)
      write (82) shl(1:8), sg(1:8, :)
    end do
    do ip = 1, npt
      call lintetshl(pts(ip, :), 4, 3, shlu4, shgradgu4)
      write (82) shlu4, shgradgu4
    end do
    do ip = 1, npt
      call linprishl(pts(ip, :), 6, 3, shlu, shgradgu)
      write (82) shlu, shgradgu
    end do
    ! quadrature rules
    call genGPandGWtetpri(gp1, gw1, 1)
    call genGPandGWtetpri(gp4, gw4, 4)
    call genGPandGWtetpri(gp6, gw6, 6)
    call genGPandGWtetpri(gp8, gw8, 8)
    write (82) gp1, gw1, gp4, gw4, gp6, gw6, gp8, gw8
    ! 3x3 inverse of each point-derived matrix
    do ip = 1, npt
      A(:, 1) = pts(ip, :) + (/2.0d0, 0.1d0, -0.3d0/)
      A(:, 2) = (/0.2d0, 1.5d0, 0.4d0/) - pts(ip, :)
      A(:, 3) = (/-0.1d0, 0.3d0, 1.7d0/) + 0.5d0*pts(ip, (/2, 3, 1/))
      call get_inverse_3x3(A, Ai, dj)
      write (82) Ai, dj
    end do
    ! eval_shape_fem on given tets, then given prisms, at each quadrature pt
    read (10) ntet
    allocate (xt(ntet, 4, 3))
    read (10) xt
    do ip = 1, ntet
      xl4 = xt(ip, :, :)
      dl4 = 0.01d0*xt(ip, :, (/3, 1, 2/))
      do q = 1, 4
        call eval_shape_fem(4, gp4(q, :), xl4, dl4, shlu4, shgradgu4, dxidx, &
                            Gij, Ginv)
        write (82) shlu4, shgradgu4, dxidx, Gij, Ginv, DetJ
      end do
    end do
    read (10) npri
    allocate (xp(npri, 6, 3))
    read (10) xp
    do ip = 1, npri
      xl = xp(ip, :, :)
      dl = 0.01d0*xp(ip, :, (/3, 1, 2/))
      do q = 1, 6
        call eval_shape_fem(6, gp6(q, :), xl, dl, shlu, shgradgu, dxidx, &
                            Gij, Ginv)
        write (82) shlu, shgradgu, dxidx, Gij, Ginv, DetJ
      end do
    end do
  end subroutine run_shape

  subroutine run_ns()
    integer :: nshl, ngauss, tflag, sflag, kspace, kmin, i
    real(8) :: prm(11), tol
    real(8), allocatable :: dA(:, :), uA(:, :), umA(:, :), aA(:, :), &
                            amA(:, :), pA(:), xu(:, :), xp(:), yu(:, :), &
                            yp(:), solu(:, :), solp(:)
    integer, allocatable :: iper(:), dflag(:), pflag(:)

    read (10) NNODE, NELEM, nshl, ngauss, tflag, sflag, ELM_ID_SET(1:2)
    maxNSHL = nshl
    allocate (IEN(NELEM, nshl), ELM_ID(NELEM), ELMNSHL(NELEM), &
              ELMNGAUSS(NELEM), IBC(NNODE, 8), xg(NNODE, 3), wg(NNODE))
    read (10) IEN, ELM_ID, IBC
    read (10) xg
    wg = 1.0d0
    ELMNSHL = nshl
    ELMNGAUSS = ngauss
    allocate (dA(NNODE, 3), uA(NNODE, 3), umA(NNODE, 3), aA(NNODE, 3), &
              amA(NNODE, 3), pA(NNODE))
    read (10) dA, uA, umA, aA, amA, pA, prm
    almi = prm(1); alfi = prm(2); gami = prm(3); beti = prm(4)
    Delt = prm(5); rhoa = prm(6); mua = prm(7); NS_kdc_a = prm(8)
    gravvec = prm(9:11)
    read (10) tol, kspace, kmin
    allocate (xu(NNODE, 3), xp(NNODE), yu(NNODE, 3), yp(NNODE))
    read (10) xu, xp

    call genSparStruc()
    allocate (LHSK11(9, icnt), LHSG(3, icnt), LHSD1(3, icnt), LHSM(icnt), &
              RHSGu(NNODE, 3), RHSGp(NNODE))
    LHSK11 = 0.0d0; LHSG = 0.0d0; LHSD1 = 0.0d0; LHSM = 0.0d0
    RHSGu = 0.0d0; RHSGp = 0.0d0
    call IntElmAss_NS_overset(dA, uA, umA, aA, amA, pA, tflag, sflag)

    write (82) icnt
    write (82) col, row
    write (82) LHSK11, LHSG, LHSD1, LHSM, RHSGu, RHSGp

    allocate (solu(NNODE, 3), solp(NNODE), iper(NNODE), dflag(NNODE), &
              pflag(NNODE))
    iper = (/(i, i=1, NNODE)/)
    dflag = 0; pflag = 0

    call SparseProdUP_3D(col, row, LHSK11, LHSG, LHSD1, LHSM, xu, xp, yu, yp, &
                         dflag, pflag, NNODE, nshl, icnt, NSD)
    write (82) yu, yp

    solu = 0.0d0; solp = 0.0d0
    if (kspace > 0) then
      call SparseGMRES_up(col, row, IBC, iper, dflag, pflag, RHSGu, RHSGp, &
                          solu, solp, LHSK11, LHSG, LHSD1, LHSM, icnt, tol, &
                          kspace, kmin, NNODE, nshl, NSD)
    end if
    write (82) solu, solp
  end subroutine run_ns

  subroutine run_interp()
    integer :: nshl, nrecv, srcid
    read (10) NNODE, NELEM, nshl, srcid, nrecv
    maxNSHL = nshl
    allocate (IEN(NELEM, nshl), ELM_ID(NELEM), ELMNSHL(NELEM), xg(NNODE, 3), &
              dg(NNODE, 3), ug(NNODE, 3))
    read (10) IEN, ELM_ID
    read (10) xg, dg, ug
    ELMNSHL = nshl
    SRC_ID_SET(1) = srcid
    bmesh(1)%NNODE = nrecv
    bmesh(1)%NSD = 3
    allocate (bmesh(1)%xg(nrecv, 3), bmesh(1)%dg(nrecv, 3), &
              bmesh(1)%ug(nrecv, 3))
    read (10) bmesh(1)%xg, bmesh(1)%dg
    call interpolation_tet(1)
    write (82) bmesh(1)%ug
  end subroutine run_interp

  subroutine run_tem()
    integer :: nshl, ngauss, tflag, kspace, kmin
    real(8) :: prm(4), qprm(19), tol
    real(8), allocatable :: dA(:, :), uA(:, :), umA(:, :), aA(:, :), &
                            amA(:, :), pA(:), phiA(:), rphiA(:), TA(:), rTA(:), &
                            x(:), y(:), sol(:)

    read (10) NNODE, NELEM, nshl, ngauss, tflag
    maxNSHL = nshl
    allocate (IEN(NELEM, nshl), ELM_ID(NELEM), ELMNSHL(NELEM), &
              ELMNGAUSS(NELEM), IBC(NNODE, 8), xg(NNODE, 3), wg(NNODE), &
              dgold(NNODE, 3), ugold(NNODE, 3))
    read (10) IEN, ELM_ID, IBC
    read (10) xg
    wg = 1.0d0
    dgold = 0.0d0
    ugold = 0.0d0
    ELMNSHL = nshl
    ELMNGAUSS = ngauss
    allocate (dA(NNODE, 3), uA(NNODE, 3), umA(NNODE, 3), aA(NNODE, 3), &
              amA(NNODE, 3), pA(NNODE), phiA(NNODE), rphiA(NNODE), TA(NNODE), &
              rTA(NNODE))
    read (10) dA, uA, umA, phiA, TA, rTA, prm, qprm
    aA = 0.0d0; amA = 0.0d0; pA = 0.0d0; rphiA = 0.0d0
    almi = prm(1); alfi = prm(2); gami = prm(3); Delt = prm(4)
    rhoa = qprm(1); rhow = qprm(2); mua = qprm(3); muw = qprm(4)
    cpa = qprm(5); cpw = qprm(6); kappaa = qprm(7); kappaw = qprm(8)
    rhos = qprm(9); mus = qprm(10); cps = qprm(11); kappas = qprm(12)
    Ts = qprm(13); lh = qprm(14); c_cond = qprm(15); c_evap = qprm(16)
    C_DMDOT = qprm(17); TEM_kdc = qprm(18); Tref = qprm(19)
    read (10) tol, kspace, kmin
    allocate (x(NNODE), y(NNODE), sol(NNODE))
    read (10) x

    call genSparStruc()
    allocate (LHSTem(icnt), RHSGTem(NNODE))
    LHSTem = 0.0d0
    RHSGTem = 0.0d0
    call IntElmAss_Tem_Quenching_STAB(dA, uA, umA, aA, amA, pA, phiA, rphiA, &
                                      TA, rTA, tflag)
    write (82) icnt
    write (82) col, row
    write (82) LHSTem, RHSGTem

    call SparseProd_LS(LHSTem, col, row, x, y, NNODE, nshl, icnt, NSD)
    write (82) y

    sol = 0.0d0
    if (kspace > 0) then
      call SparseGMRES_tem(LHSTem, tol, col, row, RHSGTem, sol, kspace, kmin, &
                           NNODE, nshl, icnt, NSD)
    end if
    write (82) sol
  end subroutine run_tem

  subroutine run_nsvof()
    integer :: nshl, ngauss, tflag, kspace, kmin, i
    real(8) :: prm(4), qprm(19), extra(6), tol
    real(8), allocatable :: dA(:, :), uA(:, :), umA(:, :), aA(:, :), &
                            amA(:, :), pA(:), phiA(:), rphiA(:), TA(:), rTA(:), &
                            xu(:, :), xp(:), xl(:), yu(:, :), yp(:), yl(:), &
                            solu(:, :), solp(:), soll(:)
    integer, allocatable :: iper(:), dflag(:), pflag(:)

    read (10) NNODE, NELEM, nshl, ngauss, tflag
    maxNSHL = nshl
    allocate (IEN(NELEM, nshl), ELM_ID(NELEM), ELMNSHL(NELEM), &
              ELMNGAUSS(NELEM), IBC(NNODE, 8), xg(NNODE, 3), wg(NNODE))
    read (10) IEN, ELM_ID, IBC
    read (10) xg
    wg = 1.0d0
    ELMNSHL = nshl
    ELMNGAUSS = ngauss
    allocate (dA(NNODE, 3), uA(NNODE, 3), umA(NNODE, 3), aA(NNODE, 3), &
              amA(NNODE, 3), pA(NNODE), phiA(NNODE), rphiA(NNODE), TA(NNODE), &
              rTA(NNODE))
    read (10) dA, uA, umA, aA, amA, pA, phiA, rphiA, TA, rTA, prm, qprm, extra
    almi = prm(1); alfi = prm(2); gami = prm(3); Delt = prm(4)
    rhoa = qprm(1); rhow = qprm(2); mua = qprm(3); muw = qprm(4)
    cpa = qprm(5); cpw = qprm(6); kappaa = qprm(7); kappaw = qprm(8)
    rhos = qprm(9); mus = qprm(10); cps = qprm(11); kappas = qprm(12)
    Ts = qprm(13); lh = qprm(14); c_cond = qprm(15); c_evap = qprm(16)
    C_DMDOT = qprm(17); TEM_kdc = qprm(18); Tref = qprm(19)
    gravvec = extra(1:3); NS_kdc_w = extra(4); NS_kdc_a = extra(5); LSC_kdc = extra(6)
    read (10) tol, kspace, kmin
    allocate (xu(NNODE, 3), xp(NNODE), xl(NNODE), yu(NNODE, 3), yp(NNODE), yl(NNODE))
    read (10) xu, xp, xl

    call genSparStruc()
    allocate (LHSK11(9, icnt), LHSG(3, icnt), LHSD1(3, icnt), LHSD2(3, icnt), &
              LHSM(icnt), LHSls(icnt), LHSlsu(3, icnt), LHSPls(icnt), &
              LHSUls(3, icnt), RHSGu(NNODE, 3), RHSGp(NNODE), RHSGls(NNODE))
    LHSK11 = 0.0d0; LHSG = 0.0d0; LHSD1 = 0.0d0; LHSD2 = 0.0d0; LHSM = 0.0d0
    LHSls = 0.0d0; LHSlsu = 0.0d0; LHSPls = 0.0d0; LHSUls = 0.0d0
    RHSGu = 0.0d0; RHSGp = 0.0d0; RHSGls = 0.0d0
    call IntElmAss_NSVOF_Quenching_STAB(dA, uA, umA, aA, amA, pA, phiA, rphiA, &
                                        TA, rTA, tflag)
    write (82) icnt
    write (82) col, row
    write (82) LHSK11, LHSG, LHSD1, LHSM, LHSls, LHSlsu, LHSPls, LHSUls, &
               RHSGu, RHSGp, RHSGls

    allocate (solu(NNODE, 3), solp(NNODE), soll(NNODE), iper(NNODE), dflag(NNODE), &
              pflag(NNODE))
    iper = (/(i, i=1, NNODE)/)
    dflag = 0; pflag = 0
    call SparseProd_NS_conv(col, row, LHSK11, LHSG, LHSD1, LHSM, LHSls, LHSlsu, &
                            LHSPls, LHSUls, xu, xp, xl, yu, yp, yl, dflag, pflag, &
                            NNODE, nshl, icnt, NSD)
    write (82) yu, yp, yl

    solu = 0.0d0; solp = 0.0d0; soll = 0.0d0
    if (kspace > 0) then
      call SparseGMRES_ns_conv(col, row, IBC, iper, dflag, pflag, RHSGu, RHSGp, &
                               RHSGls, solu, solp, soll, LHSK11, LHSG, LHSD1, LHSM, &
                               LHSls, LHSlsu, LHSPls, LHSUls, icnt, tol, kspace, kmin, &
                               NNODE, nshl, NSD)
    end if
    write (82) solu, solp, soll
  end subroutine run_nsvof

  subroutine run_face()
    integer :: nshl, ngauss, tflag, sflag, nface, ngb
    real(8) :: prm(11)
    real(8), allocatable :: dA(:, :), uA(:, :), umA(:, :), aA(:, :), &
                            amA(:, :), pA(:)

    read (10) NNODE, NELEM, nshl, ngauss, tflag, sflag, ELM_ID_SET(1:2), nface, ngb
    maxNSHL = nshl
    NSHLBmax = 3
    allocate (IEN(NELEM, nshl), ELM_ID(NELEM), ELMNSHL(NELEM), &
              ELMNGAUSS(NELEM), IBC(NNODE, 8), xg(NNODE, 3), wg(NNODE))
    read (10) IEN, ELM_ID, IBC
    read (10) xg
    wg = 1.0d0
    ELMNSHL = nshl
    ELMNGAUSS = ngauss
    allocate (dA(NNODE, 3), uA(NNODE, 3), umA(NNODE, 3), aA(NNODE, 3), &
              amA(NNODE, 3), pA(NNODE))
    read (10) dA, uA, umA, aA, amA, pA, prm
    almi = prm(1); alfi = prm(2); gami = prm(3); beti = prm(4)
    Delt = prm(5); rhoa = prm(6); mua = prm(7); NS_kdc_a = prm(8)
    gravvec = prm(9:11)
    NBOUND = 1
    allocate (bound(1), BCugType(1, 3))
    BCugType = 2
    bound(1)%FACE_ID = 1
    bound(1)%NFACE = nface
    allocate (bound(1)%F2E(nface), bound(1)%FACE_OR(nface), bound(1)%NGAUSSB(nface))
    read (10) bound(1)%F2E, bound(1)%FACE_OR
    bound(1)%NGAUSSB = ngb

    call genSparStruc()
    allocate (LHSK11(9, icnt), LHSG(3, icnt), LHSD1(3, icnt), LHSM(icnt), &
              RHSGu(NNODE, 3), RHSGp(NNODE))
    LHSK11 = 0.0d0; LHSG = 0.0d0; LHSD1 = 0.0d0; LHSM = 0.0d0
    RHSGu = 0.0d0; RHSGp = 0.0d0
    call FaceAssembly_NS_weak(dA, uA, umA, aA, amA, pA, tflag, sflag)
    call FaceAssembly_NS_outflow(dA, uA, umA, aA, amA, pA, tflag, sflag)
    write (82) icnt
    write (82) col, row
    write (82) LHSK11, LHSG, LHSD1, LHSM, RHSGu, RHSGp
  end subroutine run_face

  subroutine run_io()
    integer :: id, rstep, wstep, b, nsh

    read (10) id, fem_flag, rstep, wstep, bmesh_ID_set(1:2)
    call input_mesh(id)
    nsh = size(IEN, 2)
    write (82) NSD, nsh, NSHLBmax, NNODE, NELEM, NBOUND, NPATCH
    write (82) xg, NodeID, ELMNSHL, IEN, ELM_ID
    do b = 1, NBOUND
      write (82) bound(b)%FACE_ID, bound(b)%NFACE, bound(b)%NNODE
      write (82) bound(b)%FACE_IEN, bound(b)%NSHLB, bound(b)%F2E, bound(b)%FACE_OR, &
                 bound(b)%BNODES, bound(b)%L2SNODE, bound(b)%L2SELEM
    end do
    call input_bmesh()
    do b = 1, 2
      write (82) bmesh(b)%NNODE
      write (82) bmesh(b)%XG, bmesh(b)%owned
    end do
    allocate (dgold(NNODE, 3), ugold(NNODE, 3), ugmold(NNODE, 3), acgold(NNODE, 3), &
              acgmold(NNODE, 3), phigold(NNODE), rphigold(NNODE), pgold(NNODE), &
              Tgold(NNODE), rTgold(NNODE))
    call readSol(rstep)
    write (82) rstep, time
    write (82) dgold, ugold, ugmold, acgold, acgmold, phigold, rphigold, pgold, Tgold, rTgold
    allocate (dg(NNODE, 3), ug(NNODE, 3), ugm(NNODE, 3), acg(NNODE, 3), acgm(NNODE, 3), &
              phig(NNODE), rphig(NNODE), pg(NNODE))
    read (10) time
    read (10) dg, ug, ugm, acg, acgm, phig, rphig, pg, Tgold, rTgold
    call writeSol(wstep)
  end subroutine run_io

end program ofd_refharness

!----------------------------------------------------------------------
! Adapters for the two 5-field calls of IntElmAss_NSVOF_Quenching_STAB
! (see the header): same element arrays, the reference's 5-field routines.
!----------------------------------------------------------------------
subroutine ofdh_bclhs_nsvof(nshl, iel, xKebe11, xGebe, xDebe1, xMebe, Rhsu, Rhsp, &
                            xLSebe, xLSUebe, xULSebe, xPLSebe, Rhsphi)
  implicit none
  integer, intent(in) :: nshl, iel
  real(8) :: xKebe11(9, nshl, nshl), xGebe(3, nshl, nshl), xDebe1(3, nshl, nshl), &
             xMebe(nshl, nshl), Rhsu(3, nshl), Rhsp(nshl), xLSebe(nshl, nshl), &
             xLSUebe(3, nshl, nshl), xULSebe(3, nshl, nshl), xPLSebe(nshl, nshl), &
             Rhsphi(nshl)
  real(8) :: xDebe2(3, nshl, nshl)
  xDebe2 = 0.0d0
  call BCLhs_NS_conv(nshl, iel, xKebe11, xGebe, xDebe1, xDebe2, xMebe, xLSebe, &
                     xLSUebe, xPLSebe, xULSebe, Rhsu, Rhsp, Rhsphi)
end subroutine ofdh_bclhs_nsvof

subroutine ofdh_fill_nsvof(nshl, iel, xKebe11, xGebe, xDebe1, xMebe, xLSebe, xLSUebe, &
                           xULSebe, xPLSebe)
  implicit none
  integer, intent(in) :: nshl, iel
  real(8) :: xKebe11(9, nshl, nshl), xGebe(3, nshl, nshl), xDebe1(3, nshl, nshl), &
             xMebe(nshl, nshl), xLSebe(nshl, nshl), xLSUebe(3, nshl, nshl), &
             xULSebe(3, nshl, nshl), xPLSebe(nshl, nshl)
  real(8) :: xDebe2(3, nshl, nshl)
  xDebe2 = 0.0d0
  call FillSparseMat_NS_conv(nshl, iel, xKebe11, xGebe, xDebe1, xDebe2, xMebe, &
                             xLSebe, xLSUebe, xPLSebe, xULSebe)
end subroutine ofdh_fill_nsvof
