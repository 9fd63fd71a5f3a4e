! iso_c_binding layer over libnft_hip.so (C ABI in include/nft.h).
!
! Module nft_hip: bind(C) types and interfaces of every nft_* entry point, plus
! nft_attach, which hands the reference's COMMON-block mesh and control data to nft_init.
!
! External subroutines derivative, laplc, prfr, bcv, div, precon, gcrk, mpdatam and
! mpdatamper keep the names and argument lists of the reference's subroutines
! (lippsLFroude_implgwopks3D_unstr3.f :3943, :1288, :1536, :1583, :2096, :482, :1776,
! :3029, :3617) so callers with implicit interfaces (mpsolver) link to them unchanged; they move
! their arguments to the GPU, run the HIP kernels and copy the results back. gcrk fills
! COMMON /itero/ as the reference does (:1953-1963); div with iflg=1 prints the same
! diagnostics line (:2148-2168).
!
! Compile with the same default-real kind as the reference build (FP64:
! -fdefault-real-8); npoinx must equal the reference's PARAMETER.
module nft_hip
  use iso_c_binding
  implicit none
  integer, parameter :: npoinx = 800000, nedgex = 4000000, nfacesx = 110000

  type, bind(C) :: nft_mesh_t
    integer(c_int) :: npoin, nedge, nbface, nboun, ld_poin, ld_edge, ld_face
    type(c_ptr) :: iedge, sn, kbface, sb, ibpoin, sbb, anp, coord, vol
    integer(c_int) :: ncol
    type(c_ptr) :: col_start
  end type

  type, bind(C) :: nft_params_t
    real(c_double) :: dt, u0, v0, w0, strat, theta0
    integer(c_int) :: lord, itr, exit_mode
    real(c_double) :: eps0
    integer(c_int) :: precon, sm_flag, itrp1
    real(c_double) :: omega
    real(c_double) :: dxabl, dxabr, dyabd, dyabu, dzabd, dzabu
    real(c_double) :: towxl, towxr, towyd, towyu, towzd, towzu
    integer(c_int) :: irl
    real(c_double) :: dom(6)
    integer(c_int) :: precision
    integer(c_int) :: itraj
  end type

  type, bind(C) :: nft_gcr_stats_t
    integer(c_int) :: niter, nitsm, icount, converged
    real(c_double) :: eer, eem, res_max, tol
  end type

  type, bind(C) :: nft_run_stats_t
    integer(c_int) :: steps, gcr_solves, gcr_iters, gcr_unconverged
    real(c_double) :: ms_total, ms_gcr, ms_mpdata
  end type

  ! per-step diagnostics (nft_diag in nft.h)
  type, bind(C) :: nft_diag_t
    real(c_double) :: div(4)
    real(c_double) :: eer, eem
    integer(c_int) :: niter, nitav, crnic
    real(c_double) :: cour(4)
    real(c_double) :: u(4), v(4), w(4), t(4)
    integer(c_int64_t) :: step
  end type

  ! V-cycle settings of precon's multigrid branch (nft_multigrid in nft.h)
  type, bind(C) :: nft_multigrid_t
    integer(c_int) :: mg_flag, itrp2, itrp_c1, itrp_c2, itrp_cc, transfer
    real(c_double) :: omega_c
  end type

  ! doubly periodic box (nft_periodic in nft.h)
  type, bind(C) :: nft_periodic_t
    integer(c_int) :: ncoin, ncoin2, ncoin4, ld_coin, ld_coin2, ld_coin4
    type(c_ptr) :: icoin, icoin2, icoin4
    integer(c_int) :: vol_closed
  end type

  type(c_ptr), save :: nft_context = c_null_ptr
  real(c_double), save :: nft_dt = 0.0d0

  interface
    integer(c_int) function nft_init(mesh, params, ctx) bind(C, name="nft_init")
      import :: nft_mesh_t, nft_params_t, c_ptr, c_int
      type(nft_mesh_t), intent(in) :: mesh
      type(nft_params_t), intent(in) :: params
      type(c_ptr), intent(out) :: ctx
    end function
    integer(c_int) function nft_set_multigrid(ctx, m1, match1, m2, match2, mg) bind(C, name="nft_set_multigrid")
      import :: nft_mesh_t, nft_multigrid_t, c_ptr, c_int
      type(c_ptr), value :: ctx
      type(nft_mesh_t), intent(in) :: m1, m2
      type(c_ptr), value :: match1, match2
      type(nft_multigrid_t), intent(in) :: mg
    end function
    integer(c_int) function nft_init_periodic(mesh, params, per, ctx) bind(C, name="nft_init_periodic")
      import :: nft_mesh_t, nft_params_t, nft_periodic_t, c_ptr, c_int
      type(nft_mesh_t), intent(in) :: mesh
      type(nft_params_t), intent(in) :: params
      type(nft_periodic_t), intent(in) :: per
      type(c_ptr), intent(out) :: ctx
    end function
    integer(c_int) function nft_finalize(ctx) bind(C, name="nft_finalize")
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
    end function
    type(c_ptr) function nft_last_error() bind(C, name="nft_last_error")
      import :: c_ptr
    end function
    integer(c_int) function nft_derivative(ctx, unk, dpdxa, idx) bind(C, name="nft_derivative")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: unk(*)
      real(c_double), intent(out) :: dpdxa(*)
      integer(c_int), value :: idx
    end function
    integer(c_int) function nft_laplc(ctx, press, alap, impfl) bind(C, name="nft_laplc")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: press(*)
      real(c_double), intent(out) :: alap(*)
      integer(c_int), value :: impfl
    end function
    integer(c_int) function nft_prfr(ctx, press, ustr, vstr, wstr, uout, vout, wout, impfl) &
        bind(C, name="nft_prfr")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: press(*), ustr(*), vstr(*), wstr(*)
      real(c_double), intent(out) :: uout(*), vout(*), wout(*)
      integer(c_int), value :: impfl
    end function
    integer(c_int) function nft_bcv(ctx, u, v, w, flag) bind(C, name="nft_bcv")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(inout) :: u(*), v(*), w(*)
      real(c_double), value :: flag
    end function
    integer(c_int) function nft_div(ctx, ua, va, wa, rhs, iflg) bind(C, name="nft_div")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: ua(*), va(*), wa(*)
      real(c_double), intent(out) :: rhs(*)
      integer(c_int), value :: iflg
    end function
    integer(c_int) function nft_precon(ctx, rhs, p, impfl) bind(C, name="nft_precon")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: rhs(*)
      real(c_double), intent(out) :: p(*)
      integer(c_int), value :: impfl
    end function
    integer(c_int) function nft_gcrk(ctx, p, u1, v1, w1, u2, v2, w2, impfl, stats) &
        bind(C, name="nft_gcrk")
      import :: c_ptr, c_int, c_double, nft_gcr_stats_t
      type(c_ptr), value :: ctx
      real(c_double), intent(inout) :: p(*)
      real(c_double), intent(in) :: u1(*), v1(*), w1(*)
      real(c_double), intent(out) :: u2(*), v2(*), w2(*)
      integer(c_int), value :: impfl
      type(nft_gcr_stats_t), intent(out) :: stats
    end function
    integer(c_int) function nft_mpdatam(ctx, uuu, ld, unk, iflag) bind(C, name="nft_mpdatam")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: uuu(*)
      integer(c_int), value :: ld
      real(c_double), intent(inout) :: unk(*)
      integer(c_int), value :: iflag
    end function
    integer(c_int) function nft_state_init(ctx) bind(C, name="nft_state_init")
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function nft_run_steps(ctx, nsteps, stats) bind(C, name="nft_run_steps")
      import :: c_ptr, c_int, nft_run_stats_t
      type(c_ptr), value :: ctx
      integer(c_int), value :: nsteps
      type(nft_run_stats_t), intent(out) :: stats
    end function
    ! full-precision state file (replaces Restart_Out_old / Restart_In_old :4779-4890)
    integer(c_int) function nft_save_state(ctx, path) bind(C, name="nft_save_state")
      import :: c_ptr, c_int, c_char
      type(c_ptr), value :: ctx
      character(kind=c_char), intent(in) :: path(*)
    end function
    integer(c_int) function nft_load_state(ctx, path) bind(C, name="nft_load_state")
      import :: c_ptr, c_int, c_char
      type(c_ptr), value :: ctx
      character(kind=c_char), intent(in) :: path(*)
    end function
    integer(c_int) function nft_get_state(ctx, name, host, ld) bind(C, name="nft_get_state")
      import :: c_ptr, c_int, c_double, c_char
      type(c_ptr), value :: ctx
      character(kind=c_char), intent(in) :: name(*)
      real(c_double), intent(out) :: host(*)
      integer(c_int), value :: ld
    end function
    integer(c_int) function nft_set_state(ctx, name, host, ld) bind(C, name="nft_set_state")
      import :: c_ptr, c_int, c_double, c_char
      type(c_ptr), value :: ctx
      character(kind=c_char), intent(in) :: name(*)
      real(c_double), intent(in) :: host(*)
      integer(c_int), value :: ld
    end function
    integer(c_int) function nft_diagnostics(ctx, courant_flag, d) bind(C, name="nft_diagnostics")
      import :: c_ptr, c_int, nft_diag_t
      type(c_ptr), value :: ctx
      integer(c_int), value :: courant_flag
      type(nft_diag_t), intent(out) :: d
    end function
    integer(c_int) function nft_courant(ctx, uuu, ld, cour, iflag, d) bind(C, name="nft_courant")
      import :: c_ptr, c_int, c_double, nft_diag_t
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: uuu(*)
      integer(c_int), value :: ld
      real(c_double), intent(out) :: cour(*)
      integer(c_int), value :: iflag
      type(nft_diag_t), intent(out) :: d
    end function
    integer(c_int) function nft_stats(ctx, unk, ld, d) bind(C, name="nft_stats")
      import :: c_ptr, c_int, c_double, nft_diag_t
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: unk(*)
      integer(c_int), value :: ld
      type(nft_diag_t), intent(out) :: d
    end function
    integer(c_int) function nft_get_field(ctx, name, host) bind(C, name="nft_get_field")
      import :: c_ptr, c_int, c_double, c_char
      type(c_ptr), value :: ctx
      character(kind=c_char), intent(in) :: name(*)
      real(c_double), intent(out) :: host(*)
    end function
    integer(c_int) function nft_get_gcr_stats(ctx, stats) bind(C, name="nft_get_gcr_stats")
      import :: c_ptr, c_int, nft_gcr_stats_t
      type(c_ptr), value :: ctx
      type(nft_gcr_stats_t), intent(out) :: stats
    end function
  end interface

contains

  ! the reference's diagnostic lines (div :2165-2167, courant :2090-2092, stats :2906-2910);
  ! parts: 1 div, 2 courant, 4 stats (sums of these)
  subroutine nft_print_diag(d, parts)
    type(nft_diag_t), intent(in) :: d
    integer, intent(in) :: parts
    if (iand(parts, 1) /= 0) print 201, d%div, d%eer, d%eem, d%niter, d%nitav
    if (iand(parts, 2) /= 0) then
      write(6,*) d%crnic
      print 202, d%cour
    end if
    if (iand(parts, 4) /= 0) print 203, d%u, d%v, d%w, d%t
201 format(1x,'dvmx,dvmn,dvav,dvsd:',4e11.4/1x,' eer,eem:',2e11.4,1x,'niter,nitav:',4i4)
202 format(1x,'crnmx,crnmn,crnav,crnsd:',4e11.4)
203 format(1x,'umx,umn,uav,usd:',4e11.4/1x,'vmx,vmn,vav,vsd:',4e11.4/ &
           1x,'wmx,wmn,wav,wsd:',4e11.4/1x,'tmx,tmn,tav,tsd:',4e11.4)
  end subroutine

  subroutine nft_check(rc, what)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: what
    character(kind=c_char), pointer :: msg(:)
    character(len=512) :: text
    integer :: k
    if (rc >= 0) return
    call c_f_pointer(nft_last_error(), msg, [512])
    text = ' '
    do k = 1, 512
      if (msg(k) == c_null_char) exit
      text(k:k) = msg(k)
    end do
    write(*,'(a,a,a,i0,a,a)') 'nft: ', what, ' failed (', rc, '): ', trim(text)
    error stop 1
  end subroutine

  ! Hands the reference's COMMON-block mesh and control data to nft_init. Call after the
  ! reference's `input`. ncol/col_start: vertical columns for the line preconditioner
  ! (ncol = 0: none); precon_kind: 1 point relaxation, 2 vertical line (two-point-flux
  ! tridiagonal), 3 vertical line (in-column band of laplc); precision: 0 FP64, 1 mixed
  ! (FP32 MPDATA and line solve), 2 mixed with the FP32 operator inside GCR; strat/theta0: the values mpsolver hard-codes (:134-135).
  ! With iper = 2 in COMMON /per/ (the periodic box) the coincident sets of /coin/, /per2/
  ! and /per4/ go to nft_init_periodic; vol_closed = 1 says that COMMON vol already holds
  ! mpsolver's sums over the sets (:99-103: attach after them), 0 or absent that it does not
  ! (attach before them, e.g. right after `input`).
  subroutine nft_attach(ncol, col_start, precon_kind, precision, lord, strat, theta0, vol_closed)
    integer, intent(in) :: ncol, precon_kind, precision, lord
    integer(c_int), intent(in), target :: col_start(*)
    real(8), intent(in) :: strat, theta0
    integer, intent(in), optional :: vol_closed
    integer :: iper, ncoin, ncoin2, ncoin4, icoin, icoin2, icoin4, irecog1
    common/per/iper
    common/coin/ncoin,icoin(nfacesx*2,2),irecog1(nfacesx)
    common/per2/icoin2(nfacesx,2),ncoin2
    common/per4/icoin4(nfacesx,4),ncoin4
    target :: icoin, icoin2, icoin4
    type(nft_periodic_t) :: per
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
    target :: coord, vol, sn, sb, sbb, anp, iedge, kbface, ibpoin
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
    type(nft_mesh_t) :: m
    type(nft_params_t) :: p
    m%npoin = npoin; m%nedge = nedge; m%nbface = nbface; m%nboun = nboun
    m%ld_poin = npoinx; m%ld_edge = nedgex; m%ld_face = nfacesx
    m%iedge = c_loc(iedge); m%sn = c_loc(sn); m%kbface = c_loc(kbface); m%sb = c_loc(sb)
    m%ibpoin = c_loc(ibpoin); m%sbb = c_loc(sbb); m%anp = c_loc(anp)
    m%coord = c_loc(coord); m%vol = c_loc(vol)
    m%ncol = ncol
    m%col_start = c_loc(col_start)
    p%dt = delt; p%u0 = u0inf; p%v0 = v0inf; p%w0 = w0inf
    p%strat = strat; p%theta0 = theta0
    p%lord = lord; p%itr = itr; p%exit_mode = 0; p%eps0 = eps0
    p%precon = precon_kind; p%sm_flag = nint(sm_flag1); p%itrp1 = itrp1; p%omega = omega
    p%dxabl = dxabl; p%dxabr = dxabr; p%dyabd = dyabd; p%dyabu = dyabu; p%dzabd = dzabd; p%dzabu = dzabu
    p%towxl = towxl; p%towxr = towxr; p%towyd = towyd; p%towyu = towyu; p%towzd = towzd; p%towzu = towzu
    p%irl = irl
    p%dom = [0.0d0, 15000.0d0, -6000.0d0, 6000.0d0, 0.0d0, 6000.0d0]
    p%precision = precision
    p%itraj = 0   ! mpsolver hard-codes itraj=0 (:111); nft_init accepts 1 for velprd
    nft_dt = delt
    if (iper == 2) then
      per%ncoin = ncoin; per%ncoin2 = ncoin2; per%ncoin4 = ncoin4
      per%ld_coin = 2 * nfacesx; per%ld_coin2 = nfacesx; per%ld_coin4 = nfacesx
      per%icoin = c_loc(icoin); per%icoin2 = c_loc(icoin2); per%icoin4 = c_loc(icoin4)
      per%vol_closed = 0
      if (present(vol_closed)) per%vol_closed = vol_closed
      call nft_check(nft_init_periodic(m, p, per, nft_context), 'nft_init_periodic')
    else
      call nft_check(nft_init(m, p, nft_context), 'nft_init')
    end if
    if (mg_flag == 1) call nft_attach_multigrid(itrp2, itrp_c1, itrp_c2, itrp_cc)
  end subroutine

  ! The coarse meshes and node maps of precon's multigrid branch as `input` leaves them in
  ! COMMON (:2343-2672; /edg_c/ ... /match/), handed to nft_set_multigrid. The arrays are
  ! passed as read, so the reference's level-2 metrics (sn_cc y/z from sn_c, :2467-2468)
  ! and its transfers (transfer 0) are reproduced.
  subroutine nft_attach_multigrid(itrp2, itrp_c1, itrp_c2, itrp_cc)
    integer, intent(in) :: itrp2, itrp_c1, itrp_c2, itrp_cc
    integer :: npoin_c, nedge_c, nbface_c, nbpoin_c, nboun_c
    integer :: npoin_cc, nedge_cc, nbface_cc, nbpoin_cc, nboun_cc
    common/n_c/npoin_c,nedge_c,nbface_c,nbpoin_c,nboun_c
    common/n_cc/npoin_cc,nedge_cc,nbface_cc,nbpoin_cc,nboun_cc
    real(8) :: coord_c, vol_c, sn_c, sb_c, sbb_c, anp_c
    real(8) :: coord_cc, vol_cc, sn_cc, sb_cc, sbb_cc, anp_cc
    integer :: iedge_c, kbface_c, ibpoin_c, iedge_cc, kbface_cc, ibpoin_cc
    common/xyz_c/coord_c(npoinx,3)
    common/volume_c/vol_c(npoinx)
    common/edg_c/iedge_c(nedgex,2)
    common/wge_c/sn_c(nedgex,3)
    common/bound_c/kbface_c(nfacesx,1:2),sb_c(nfacesx,3)
    common/bpoints_c/ibpoin_c(nfacesx,2),anp_c(nfacesx,3)
    common/terms_c/sbb_c(nfacesx,3)
    common/xyz_cc/coord_cc(npoinx,3)
    common/volume_cc/vol_cc(npoinx)
    common/edg_cc/iedge_cc(nedgex,2)
    common/wge_cc/sn_cc(nedgex,3)
    common/bound_cc/kbface_cc(nfacesx,1:2),sb_cc(nfacesx,3)
    common/bpoints_cc/ibpoin_cc(nfacesx,2),anp_cc(nfacesx,3)
    common/terms_cc/sbb_cc(nfacesx,3)
    integer :: match1, match2
    common/match/match1(npoinx),match2(npoinx)
    target :: coord_c, vol_c, sn_c, sb_c, sbb_c, anp_c, iedge_c, kbface_c, ibpoin_c
    target :: coord_cc, vol_cc, sn_cc, sb_cc, sbb_cc, anp_cc, iedge_cc, kbface_cc, ibpoin_cc
    target :: match1, match2
    type(nft_mesh_t) :: m1, m2
    type(nft_multigrid_t) :: mg
    m1%npoin = npoin_c; m1%nedge = nedge_c; m1%nbface = nbface_c; m1%nboun = nboun_c
    m1%ld_poin = npoinx; m1%ld_edge = nedgex; m1%ld_face = nfacesx
    m1%iedge = c_loc(iedge_c); m1%sn = c_loc(sn_c); m1%kbface = c_loc(kbface_c); m1%sb = c_loc(sb_c)
    m1%ibpoin = c_loc(ibpoin_c); m1%sbb = c_loc(sbb_c); m1%anp = c_loc(anp_c)
    m1%coord = c_loc(coord_c); m1%vol = c_loc(vol_c); m1%ncol = 0; m1%col_start = c_null_ptr
    m2%npoin = npoin_cc; m2%nedge = nedge_cc; m2%nbface = nbface_cc; m2%nboun = nboun_cc
    m2%ld_poin = npoinx; m2%ld_edge = nedgex; m2%ld_face = nfacesx
    m2%iedge = c_loc(iedge_cc); m2%sn = c_loc(sn_cc); m2%kbface = c_loc(kbface_cc); m2%sb = c_loc(sb_cc)
    m2%ibpoin = c_loc(ibpoin_cc); m2%sbb = c_loc(sbb_cc); m2%anp = c_loc(anp_cc)
    m2%coord = c_loc(coord_cc); m2%vol = c_loc(vol_cc); m2%ncol = 0; m2%col_start = c_null_ptr
    mg%mg_flag = 1; mg%itrp2 = itrp2; mg%itrp_c1 = itrp_c1; mg%itrp_c2 = itrp_c2; mg%itrp_cc = itrp_cc
    mg%transfer = 0; mg%omega_c = 0.0d0
    call nft_check(nft_set_multigrid(nft_context, m1, c_loc(match1), m2, c_loc(match2), mg), &
                   'nft_set_multigrid')
  end subroutine

end module nft_hip

! ---- drop-in replacements of the reference subroutines ---------------------------------

subroutine derivative(unk, dpdxa, idx)
  use nft_hip
  implicit none
  real(8) :: unk(*), dpdxa(*)
  integer :: idx
  call nft_check(nft_derivative(nft_context, unk, dpdxa, idx), 'derivative')
end subroutine

subroutine laplc(press, alap, impfl)
  use nft_hip
  implicit none
  real(8) :: press(*), alap(*)
  integer :: impfl
  call nft_check(nft_laplc(nft_context, press, alap, impfl), 'laplc')
end subroutine

subroutine prfr(press, ustr, vstr, wstr, uout, vout, wout, impfl)
  use nft_hip
  implicit none
  real(8) :: press(*), ustr(*), vstr(*), wstr(*), uout(*), vout(*), wout(*)
  integer :: impfl
  call nft_check(nft_prfr(nft_context, press, ustr, vstr, wstr, uout, vout, wout, impfl), 'prfr')
end subroutine

subroutine bcv(uout, vout, wout, flag)
  use nft_hip
  implicit none
  real(8) :: uout(*), vout(*), wout(*), flag
  call nft_check(nft_bcv(nft_context, uout, vout, wout, flag), 'bcv')
end subroutine

subroutine div(ua, va, wa, rhs, iflg)
  use nft_hip
  implicit none
  real(8) :: ua(*), va(*), wa(*), rhs(*)
  integer :: iflg
  integer :: npoin, nedge, nbface, nbpoin, nboun
  common/n/npoin,nedge,nbface,nbpoin,nboun
  integer :: itr, niter, nitsm, icount, nitav, ip
  real(8) :: eps0, eer, eem, divmx, divmn, divav, divsd
  common/itero/ itr,eps0,niter,nitsm,icount,eer,eem
  call nft_check(nft_div(nft_context, ua, va, wa, rhs, iflg), 'div')
  if (iflg == 1) then
    divmx = -1.d15
    divmn = 1.d15
    divav = 0.d0
    do ip = 1, npoin
      divmx = max(divmx, rhs(ip)*nft_dt)
      divmn = min(divmn, rhs(ip)*nft_dt)
      divav = divav + rhs(ip)*nft_dt
    end do
    divav = divav/dble(npoin)
    divsd = 0.d0
    do ip = 1, npoin
      divsd = divsd + (rhs(ip)*nft_dt - divav)**2
    end do
    divsd = sqrt(divsd/npoin)
    nitav = nitsm/max(icount, 1)
    print 205, divmx, divmn, divav, divsd, eer, eem, niter, nitav
205 format(1x,'dvmx,dvmn,dvav,dvsd:',4e11.4/1x,' eer,eem:',2e11.4,1x,'niter,nitav:',4i4)
  end if
end subroutine

subroutine precon(rhs, p, r, iprc, elas, impfl)
  use nft_hip
  implicit none
  real(8) :: rhs(*), p(*), r(*), elas
  integer :: iprc, impfl
  integer :: npoin, nedge, nbface, nbpoin, nboun
  common/n/npoin,nedge,nbface,nbpoin,nboun
  if (iprc == 0) then
    p(1:npoin) = rhs(1:npoin)
    return
  end if
  call nft_check(nft_precon(nft_context, rhs, p, impfl), 'precon')
end subroutine

subroutine gcrk(p, u1, v1, w1, u2, v2, w2, impfl)
  use nft_hip
  implicit none
  real(8) :: p(*), u1(*), v1(*), w1(*), u2(*), v2(*), w2(*)
  integer :: impfl
  integer :: itr, niter, nitsm, icount, rc
  real(8) :: eps0, eer, eem
  common/itero/ itr,eps0,niter,nitsm,icount,eer,eem
  type(nft_gcr_stats_t) :: st
  rc = nft_gcrk(nft_context, p, u1, v1, w1, u2, v2, w2, impfl, st)
  call nft_check(rc, 'gcrk')
  niter = st%niter
  nitsm = nitsm + st%niter
  icount = icount + 1
  eer = st%eer
  eem = st%eem
end subroutine

subroutine mpdatam(uuu, unk, iflag)
  use nft_hip
  implicit none
  real(8) :: uuu(npoinx,3), unk(*)
  integer :: iflag
  call nft_check(nft_mpdatam(nft_context, uuu, npoinx, unk, iflag), 'mpdatam')
end subroutine

! mpdatamper (reference :3617-3941), which mpsolver calls in the periodic box (:314-319):
! nft_mpdatam on the periodic context that nft_attach made from iper = 2
subroutine mpdatamper(uuu, unk, iflag)
  use nft_hip
  implicit none
  real(8) :: uuu(npoinx,3), unk(*)
  integer :: iflag
  call nft_check(nft_mpdatam(nft_context, uuu, npoinx, unk, iflag), 'mpdatamper')
end subroutine

! courant(uuu,cour,iflag) (reference :1974-2094) on the GPU: the Courant numbers of the
! uuu given and their statistics; returns cour and prints the reference's lines. The
! device-resident time-step state is left as it is.
subroutine courant(uuu, cour, iflag)
  use nft_hip
  implicit none
  real(8) :: uuu(npoinx,3), cour(*)
  integer :: iflag
  type(nft_diag_t) :: d
  call nft_check(nft_courant(nft_context, uuu, npoinx, cour, iflag, d), 'courant')
  call nft_print_diag(d, 2)
end subroutine

! stats(unk) (reference :2847-2912) on the GPU: prints the reference's lines for the unk
! given; the device-resident time-step state is left as it is.
subroutine stats(unk)
  use nft_hip
  implicit none
  real(8) :: unk(npoinx,4)
  type(nft_diag_t) :: d
  call nft_check(nft_stats(nft_context, unk, npoinx, d), 'stats')
  call nft_print_diag(d, 4)
end subroutine
