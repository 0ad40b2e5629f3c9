! Adapter for the fine-scale residual form of the NS assembly (test
! infrastructure only; nothing of the reference is copied).
!
! The recipe compiles assembleNavSto.f90 a second time with its call of
! e3RHS_3D_fluid (assembleNavSto.f90:284-288) mapped to this routine.  It
! sets the fine-scale velocity uprime = -tauM rLi, the line the reference
! keeps commented out at e3RHS.f90:31, and calls the reference's own
! e3Rhs_3D_fluid with it.  tauBar stays what the caller passes (0,
! assembleNavSto.f90:268).  The live call passes the shape Hessian in the
! uprime position (zero on linear elements); this replaces it.
subroutine ofdh_e3rhs_vms(nshl, rhoi, mui, ui, aci, umi, acmi, uadvi, pri, rLi, fi, &
                          duidxi, ddidxi, tauM, tauP, tauC, tauBar, kap_dc, gwt, shgu, &
                          shgradgu, shhessgu, Rhsu, Rhsp)
  use aAdjKeep
  use commonvars
  implicit none
  integer, intent(in) :: nshl
  real(8), intent(in) :: rhoi, mui, ui(NSD), aci(NSD), umi(NSD), acmi(NSD), uadvi(NSD), &
                         pri, rLi(NSD), fi(NSD), duidxi(NSD, NSD), ddidxi(NSD, NSD), &
                         tauM, tauP, tauC, tauBar, kap_dc, gwt, shgu(nshl), &
                         shgradgu(nshl, NSD), shhessgu(*)
  real(8), intent(inout) :: Rhsu(NSD, nshl), Rhsp(nshl)
  real(8) :: uprime(NSD)

  uprime(:) = -tauM * rLi(:)
  call e3Rhs_3D_fluid(nshl, rhoi, mui, ui, aci, umi, acmi, uadvi, pri, rLi, fi, duidxi, &
                      ddidxi, tauM, tauP, tauC, tauBar, kap_dc, gwt, shgu, shgradgu, &
                      uprime, Rhsu, Rhsp)
end subroutine ofdh_e3rhs_vms
