! nftdrop: test program of the Fortran drop-in layer (TEST INFRASTRUCTURE).
!
! Plays the reference's mpsolver: fills the COMMON blocks as the reference's `input` does
! (lippsLFroude_implgwopks3D_unstr3.f :2172-2679; mesh3d.d, inputinc.d, input_mg_v3.d from
! the working directory), hands them to the GPU with nft_attach, then calls the drop-in
! replacements with the reference's own argument lists and leading dimensions:
!   derivative(unk,dpdxa,idx)          :3943     laplc(press,alap,impfl)        :1288
!   prfr(press,u*,v*,w*,u,v,w,impfl)   :1536     bcv(uout,vout,wout,flag)       :1583
!   div(ua,va,wa,rhs,iflg)             :2096     precon(rhs,p,r,iprc,elas,impfl) :482
!   mpdatam(uuu,unk(1,k),k)            as at :309-312 (columns of unk(npoinx,4) by
!                                      sequence association); mpdatamper as at :315-318
!                                      when inputinc.d sets iper = 2 (the periodic box,
!                                      whose coincident sets follow the mesh, :2569-2583)
!   gcrk(p,unk1(1,1),unk1(1,2),unk1(1,3),uuu(1,1),uuu(1,2),uuu(1,3),1)  as at :362-365
!   courant(uuu,cour,iflag)            :1974     stats(unk)                     :2847
! Inputs: vec_in.bin, 8 float64 node arrays (p, u, v, w, r, U, V, W). Outputs: every result
! array in call order, then /itero/ (niter, nitsm, icount, eer, eem), to dropin_out.bin.
!
! Usage: nftdrop <nlev> <precon> <precision> <lord> <strat>
program nftdrop
  use iso_c_binding
  use nft_hip
  implicit none
  ! COMMON blocks as the reference declares them (read by nft_attach and the drop-ins)
  integer :: npoin, nedge, nbface, nbpoin, nboun
  common/n/npoin,nedge,nbface,nbpoin,nboun
  real(8) :: coord, vol, sn, sb, sbb, anp
  integer :: iedge, kbface, ibpoin
  common/xyz/coord(npoinx,3)
  common/volume/vol(npoinx)
  common/edg/iedge(nedgex,2)
  common/wge/sn(nedgex,3)
  common/bound/kbface(nfacesx,1:2),sb(nfacesx,3)
  common/bpoints/ibpoin(nfacesx,2),anp(nfacesx,3)
  common/terms/sbb(nfacesx,3)
  real(8) :: delt, adapt, afree(5), p0inf, rh0inf, u0inf, v0inf, w0inf
  common/sss/delt,adapt
  common/frflo/afree,p0inf,rh0inf,u0inf,v0inf,w0inf
  integer :: itr, niter, nitsm, icount
  real(8) :: eps0, eer, eem
  common/itero/ itr,eps0,niter,nitsm,icount,eer,eem
  real(8) :: dxabl,dxabr,dyabd,dyabu,dzabd,dzabu,towxl,towxr,towyd,towyu,towzd,towzu
  integer :: irl
  common/absspec/dxabl,dxabr,dyabd,dyabu,dzabd,dzabu,towxl,towxr,towyd,towyu,towzd,towzu,irl
  integer :: mg_flag, itrp1, itrp2, itrp_c1, itrp_c2, itrp_cc
  real(8) :: omega, sm_flag1, sm_flag2
  common/vcycle/mg_flag,omega,itrp1,itrp2,itrp_c1,itrp_c2,itrp_cc
  common/smooth/sm_flag1,sm_flag2
  integer :: iper, ncoin, ncoin2, ncoin4, icoin, icoin2, icoin4, irecog1
  common/per/iper
  common/coin/ncoin,icoin(nfacesx*2,2),irecog1(nfacesx)
  common/per2/icoin2(nfacesx,2),ncoin2
  common/per4/icoin4(nfacesx,4),ncoin4

  ! mpsolver's work arrays, dimensioned with npoinx as the reference's
  real(8), allocatable :: unk(:,:), uuu(:,:), unk1(:,:), vin(:,:)
  real(8), allocatable :: a1(:), a2(:), a3(:), pp(:), rr(:), cour(:)
  integer(c_int), allocatable, target :: cols(:)
  integer :: nlev, pk, prec, lord, i, k, ii, ntime, nrkstage, istst, mrestart, isfera
  integer :: ielast, lordin, itrp, match_flag, ncol
  real(8) :: strat, cfl, relax, re, betap, elas
  character(len=64) :: arg

  call get_command_argument(1, arg); read(arg,*) nlev
  call get_command_argument(2, arg); read(arg,*) pk
  call get_command_argument(3, arg); read(arg,*) prec
  call get_command_argument(4, arg); read(arg,*) lord
  call get_command_argument(5, arg); read(arg,*) strat

  ! control files (input :2265-2321)
  open(8, file='input_mg_v3.d', status='old')
  read(8,*); read(8,*) mg_flag; read(8,*); read(8,*); read(8,*) sm_flag1, sm_flag2
  read(8,*); read(8,*) omega; read(8,*); read(8,*)
  read(8,*) itrp1, itrp2, itrp_c1, itrp_c2, itrp_cc; read(8,*); read(8,*) match_flag
  close(8)
  open(1, file='inputinc.d', status='old')
  read(1,*); read(1,*) ntime, nrkstage, istst, mrestart, isfera, iper
  read(1,*); read(1,*) cfl, delt, adapt, relax
  read(1,*); read(1,*) p0inf, u0inf, v0inf, w0inf, re
  read(1,*); read(1,*); read(1,*) ielast, eps0, itr, lordin
  read(1,*); read(1,*) itrp, betap
  read(1,*); read(1,*) irl
  if (irl == 1) then
    read(1,*); read(1,*) dxabl, dxabr, dyabd, dyabu, dzabd, dzabu
    read(1,*); read(1,*) towxl, towxr, towyd, towyu, towzd, towzu
  end if
  close(1)
  niter = 0; nitsm = 0; icount = 0; eer = 0.0d0; eem = 0.0d0

  ! the mesh (input :2343-2501), into the COMMON arrays
  open(5, file='mesh3d.d', status='old')
  read(5,*) npoin, nedge, nbface, nbpoin, nboun
  do i = 1, npoin
    read(5,*) ii, (coord(i,k), k=1,3), vol(i)
  end do
  do i = 1, nedge
    read(5,*) ii, iedge(i,1), iedge(i,2)
  end do
  do i = 1, nedge
    read(5,*) ii, (sn(i,k), k=1,3)
  end do
  do i = 1, nbface
    read(5,*) ii, kbface(i,1), kbface(i,2)
  end do
  do i = 1, nbface
    read(5,*) ii, (sb(i,k), k=1,3)
  end do
  do i = 1, nboun
    read(5,*) ii, ibpoin(i,1), ibpoin(i,2)
  end do
  do i = 1, nboun
    read(5,*) ii, (sbb(i,k), k=1,3)
  end do
  do i = 1, nboun
    read(5,*) ii, (anp(i,k), k=1,3)
  end do
  ncoin = 0; ncoin2 = 0; ncoin4 = 0
  if (iper == 2) then
    read(5,*) ncoin
    do i = 1, ncoin
      read(5,*) icoin(i,1), icoin(i,2)
    end do
    read(5,*) ncoin2
    do i = 1, ncoin2
      read(5,*) icoin2(i,1), icoin2(i,2)
    end do
    read(5,*) ncoin4
    do i = 1, ncoin4
      read(5,*) icoin4(i,1), icoin4(i,2), icoin4(i,3), icoin4(i,4)
    end do
  end if
  close(5)

  ncol = 0
  if (nlev > 0) ncol = npoin/nlev
  allocate(cols(ncol + 1))
  do i = 0, ncol
    cols(i+1) = i*nlev
  end do
  call nft_attach(ncol, cols, pk, prec, lord, strat, 300.0d0)

  allocate(vin(npoin,8))
  open(20, file='vec_in.bin', access='stream', form='unformatted', status='old')
  read(20) vin
  close(20)
  allocate(unk(npoinx,4), uuu(npoinx,3), unk1(npoinx,4))
  allocate(a1(npoinx), a2(npoinx), a3(npoinx), pp(npoinx), rr(npoinx), cour(npoinx))
  unk = 0.0d0; uuu = 0.0d0; unk1 = 0.0d0
  open(30, file='dropin_out.bin', access='stream', form='unformatted', status='replace')

  ! derivative, idx = 1, 2, 3
  do k = 1, 3
    call derivative(vin(1,1), a1, k)
    write(30) a1(1:npoin)
  end do
  ! laplc, impfl = 0, 1
  do k = 0, 1
    call laplc(vin(1,1), a1, k)
    write(30) a1(1:npoin)
  end do
  ! prfr(p, u*, v*, w*, u, v, w, 1)
  call prfr(vin(1,1), vin(1,2), vin(1,3), vin(1,4), a1, a2, a3, 1)
  write(30) a1(1:npoin), a2(1:npoin), a3(1:npoin)
  ! bcv in place, flag 0 then 1
  do k = 0, 1
    a1(1:npoin) = vin(:,2); a2(1:npoin) = vin(:,3); a3(1:npoin) = vin(:,4)
    call bcv(a1, a2, a3, dble(k))
    write(30) a1(1:npoin), a2(1:npoin), a3(1:npoin)
  end do
  ! div, iflg = -1 and 1 (the latter also prints the reference's diagnostics line)
  call div(vin(1,2), vin(1,3), vin(1,4), a1, -1)
  write(30) a1(1:npoin)
  call div(vin(1,2), vin(1,3), vin(1,4), a1, 1)
  write(30) a1(1:npoin)
  ! precon(rhs, p, r, iprc, elas, impfl)
  elas = 0.0d0
  call precon(vin(1,5), pp, rr, 1, elas, 1)
  write(30) pp(1:npoin)
  ! mpdatam(uuu, unk(1,k), k) for u, v, w, theta' as mpsolver calls it
  uuu(1:npoin,1) = vin(:,6); uuu(1:npoin,2) = vin(:,7); uuu(1:npoin,3) = vin(:,8)
  unk(1:npoin,1) = vin(:,2); unk(1:npoin,2) = vin(:,3); unk(1:npoin,3) = vin(:,4)
  unk(1:npoin,4) = vin(:,1)
  do k = 1, 4
    if (iper == 2) then
      call mpdatamper(uuu, unk(1,k), k)
    else
      call mpdatam(uuu, unk(1,k), k)
    end if
  end do
  do k = 1, 4
    write(30) unk(1:npoin,k)
  end do
  ! gcrk(p, u*, v*, w*, u, v, w, 1) with mpsolver's argument arrays (:362-365): the
  ! corrected velocities land in uuu
  unk1(1:npoin,1) = vin(:,2); unk1(1:npoin,2) = vin(:,3); unk1(1:npoin,3) = 0.1d0*vin(:,4)
  pp = 0.0d0
  call gcrk(pp, unk1(1,1), unk1(1,2), unk1(1,3), uuu(1,1), uuu(1,2), uuu(1,3), 1)
  write(30) pp(1:npoin), uuu(1:npoin,1), uuu(1:npoin,2), uuu(1:npoin,3)
  write(30) dble(niter), dble(nitsm), dble(icount), eer, eem
  ! courant and stats of arrays given here (the reference's prints)
  uuu(1:npoin,1) = vin(:,6); uuu(1:npoin,2) = vin(:,7); uuu(1:npoin,3) = vin(:,8)
  call courant(uuu, cour, 1)
  write(30) cour(1:npoin)
  unk(1:npoin,1) = vin(:,2); unk(1:npoin,2) = vin(:,3); unk(1:npoin,3) = vin(:,4)
  unk(1:npoin,4) = vin(:,1)
  call stats(unk)
  close(30)
  call nft_check(nft_finalize(nft_context), 'nft_finalize')
end program nftdrop
