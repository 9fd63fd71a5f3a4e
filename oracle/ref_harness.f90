! Harness around the flang-compiled reference (TEST INFRASTRUCTURE ONLY).
!
! Linked with oracle/_ref/ref_l<k>.o, the reference file compiled with its program unit
! renamed to the subroutine `nftmain` (see oracle/build_ref.sh). Two modes:
!   nft_ref run : call nftmain, i.e. the reference program itself (input + mpsolver);
!   nft_ref ops : open the reference's input units, call `input`, repeat the set-up part
!                 of mpsolver (lippsLFroude_implgwopks3D_unstr3.f:108-170, 250-252), then
!                 call the reference's own derivative/laplc/bcv/prfr/div/precon/mpdatam/gcrk
!                 on the vectors of vec_in.bin and write every result to vec_out.bin.
!                 With iper=2 (the periodic box) the set-up closes vol over the coincident
!                 nodes first, as mpsolver does (:99-103), and mpdatamper takes the place of
!                 mpdatam, as in mpsolver's step (:314-319).
!   nft_ref mgops : the same set-up including the coarse meshes (beta_timestep_unstr and
!                 compute_d_unstr for levels 2 and 3), then the multigrid pieces:
!                 restriction_unstr, prolongation_unstr, laplc_c, laplc_cc, and precon and
!                 gcrk with the V-cycle as input_mg_v3.d sets it (mg_flag=1).
! The COMMON declarations below restate the reference's block layouts at its default
! PARAMETER sizes (they must match for the blocks to alias).
program nft_ref_harness
  implicit real (a-h,o-z)
  parameter (nedgex=4000000,npoinx=800000,nfacesx=110000)
  common/n/npoin,nedge,nbface,nbpoin,nboun
  common/metric/cosa(npoinx),gmm(npoinx),radious
  common/xyz/coord(npoinx,3)
  common/frflo/afree(5),p0inf,rh0inf,u0inf,v0inf,w0inf
  common/tsteps/timestp(npoinx),istst,mrestart,time
  common/sss/delt,adapt
  common/coefficients/coefw(npoinx)
  common/profil/rh0(npoinx),th0(npoinx),the(npoinx),ue(npoinx),ve(npoinx),lipps
  common/ctherm/ rgas,cp,cap,strat,g,th00,tt00,pr00,rh00,u00,v00,u0z,v0z
  common/doperator/d1(npoinx),d2(npoinx),d3(npoinx)
  common/beta/betaunstr1,betaunstr2,betaunstr3
  common/itero/ itr,eps0,niter,nitsm,icount,eer,eem
  common/sf/isfera
  common/elast/ cfels,iflels
  common/n_c/npoin_c,nedge_c,nbface_c,nbpoin_c,nboun_c
  common/n_cc/npoin_cc,nedge_cc,nbface_cc,nbpoin_cc,nboun_cc
  common/per/iper
  common/volume/vol(npoinx)
  real, save :: alpha(npoinx), v(npoinx,8), o1(npoinx), o2(npoinx), o3(npoinx), o4(npoinx)
  real, save :: uuu(npoinx,3), wk(npoinx)
  character(len=16) :: mode
  integer :: ip, k

  call get_command_argument(1, mode)
  if (trim(mode) == 'run') then
    call nftmain
    stop
  end if

  open(5,file='mesh3d.d',access='sequential',status='old')
  open(35,file='mesh3d_c.d',access='sequential',status='old')
  open(45,file='mesh3d_cc.d',access='sequential',status='old')
  open(1,file='inputinc.d',access='sequential',status='old')
  open(8,file='input_mg_v3.d',access='sequential',status='old')
  open(12,file='inputre.d',access='sequential',status='old')
  open(55,file='match1.d',access='sequential',status='old')
  open(56,file='match2.d',access='sequential',status='old')
  open(21,file='harness.d',access='sequential',status='old')
  call input
  read(21,*) strat_in
  if (iper == 2) then
    call cyclicr(vol)
    call cyclicr2(vol)
    call cyclicr4(vol)
  end if

  ! mpsolver set-up (Cartesian, lipps=0, implgw=1)
  iflels=0
  afree(1)=u0inf
  afree(2)=v0inf
  afree(3)=w0inf
  afree(4)=0.
  afree(5)=1.
  rh00=afree(5)
  lipps=0
  strat=strat_in
  tetbar=300.
  tet00=tetbar
  tt00=tetbar
  th00=tetbar
  call rhprof
  call thprof
  call absorbers(alpha)
  r=radious*float(isfera)+(1.-float(isfera))
  do ip=1,npoin
    y=coord(ip,2)
    cosa(ip)=cos(y)*float(isfera)+(1.-float(isfera))
    gmm(ip)=1.+coord(ip,3)/r*float(isfera)
    rh0(ip)=rh0(ip)*gmm(ip)**2*cosa(ip)
  end do
  call beta_timestep_unstr(1)
  call compute_d_unstr(1)
  do ip=1,npoin
    ue(ip)=afree(1)
    ve(ip)=afree(2)
    timestp(ip)=delt
  end do
  do ip=1,npoin
    dthedz=tet00*strat*1
    gth0=g/th0(ip)
    coefw(ip)=1.+0.25*timestp(1)**2*gth0*dthedz
    coefw(ip)=1./coefw(ip)
  end do

  open(22,file='vec_in.bin',access='stream',form='unformatted',status='old')
  do k=1,8
    read(22) (v(ip,k),ip=1,npoin)
  end do
  close(22)
  open(23,file='vec_out.bin',access='stream',form='unformatted',status='replace')
  if (trim(mode) == 'mgops') then
    call beta_timestep_unstr(2)
    call beta_timestep_unstr(3)
    call compute_d_unstr(2)
    call compute_d_unstr(3)
    write(23) (d2(ip),ip=1,npoin_c)
    write(23) (d3(ip),ip=1,npoin_cc)
    write(23) betaunstr2, betaunstr3
    call restriction_unstr(v(1,1),o1,1)
    write(23) (o1(ip),ip=1,npoin_c)
    call restriction_unstr(v(1,2),o1,2)
    write(23) (o1(ip),ip=1,npoin_cc)
    call prolongation_unstr(v(1,3),o1,1)
    write(23) (o1(ip),ip=1,npoin)
    call prolongation_unstr(v(1,4),o1,2)
    write(23) (o1(ip),ip=1,npoin_c)
    call laplc_c(v(1,1),o1,1)
    write(23) (o1(ip),ip=1,npoin_c)
    call laplc_c(v(1,1),o1,0)
    write(23) (o1(ip),ip=1,npoin_c)
    call laplc_cc(v(1,1),o1,1)
    write(23) (o1(ip),ip=1,npoin_cc)
    call precon(v(1,5),o1,wk,1,elas,1)
    write(23) (o1(ip),ip=1,npoin)
    o4(1:npoin)=0.
    nitsm=0
    icount=0
    call gcrk(o4,v(1,2),v(1,3),v(1,4),o1,o2,o3,1)
    write(23) (o4(ip),ip=1,npoin)
    write(23) (o1(ip),ip=1,npoin), (o2(ip),ip=1,npoin), (o3(ip),ip=1,npoin)
    write(23) real(niter), eer, eem
    close(23)
    stop
  end if
  write(23) (alpha(ip),ip=1,npoin)
  write(23) (d1(ip),ip=1,npoin)
  write(23) (coefw(ip),ip=1,npoin)
  write(23) betaunstr1

  do k=1,3
    call derivative(v(1,1),o1,k)
    write(23) (o1(ip),ip=1,npoin)
  end do
  call laplc(v(1,1),o1,0)
  write(23) (o1(ip),ip=1,npoin)
  call laplc(v(1,1),o1,1)
  write(23) (o1(ip),ip=1,npoin)
  do k=0,1
    o1(1:npoin)=v(1:npoin,2)
    o2(1:npoin)=v(1:npoin,3)
    o3(1:npoin)=v(1:npoin,4)
    call bcv(o1,o2,o3,real(k))
    write(23) (o1(ip),ip=1,npoin), (o2(ip),ip=1,npoin), (o3(ip),ip=1,npoin)
  end do
  do k=0,1
    call prfr(v(1,1),v(1,2),v(1,3),v(1,4),o1,o2,o3,k)
    write(23) (o1(ip),ip=1,npoin), (o2(ip),ip=1,npoin), (o3(ip),ip=1,npoin)
  end do
  call div(v(1,2),v(1,3),v(1,4),o1,-1)
  write(23) (o1(ip),ip=1,npoin)
  call precon(v(1,5),o1,wk,1,elas,1)
  write(23) (o1(ip),ip=1,npoin)
  uuu(1:npoin,1)=v(1:npoin,6)
  uuu(1:npoin,2)=v(1:npoin,7)
  uuu(1:npoin,3)=v(1:npoin,8)
  do k=1,4,3
    o1(1:npoin)=v(1:npoin,1)
    if (iper == 2) then
      call mpdatamper(uuu,o1,k)
    else
      call mpdatam(uuu,o1,k)
    end if
    write(23) (o1(ip),ip=1,npoin)
  end do
  o4(1:npoin)=0.
  nitsm=0
  icount=0
  call gcrk(o4,v(1,2),v(1,3),v(1,4),o1,o2,o3,1)
  write(23) (o4(ip),ip=1,npoin)
  write(23) (o1(ip),ip=1,npoin), (o2(ip),ip=1,npoin), (o3(ip),ip=1,npoin)
  write(23) real(niter), eer, eem
  close(23)
end program nft_ref_harness
