! nftdrv: Fortran driver of the device-resident NFT time loop.
!
! Reads the reference's own input files from the working directory (mesh3d.d,
! inputinc.d, input_mg_v3.d; layouts of lippsLFroude_implgwopks3D_unstr3.f `input`,
! :2265-2321 and :2343-2501), hands the mesh to nft_init through iso_c_binding, runs the
! mpsolver initialisation and `ntime` steps of its DO ITIME loop on the GPU, prints the
! reference's per-step iteration line and writes the final state (unk, uuu, prpot) as
! raw float64 to nftdrv_state.bin. With mg_flag=1 in input_mg_v3.d it also reads the coarse
! meshes mesh3d_c.d / mesh3d_cc.d and the node maps match1.d / match2.d and runs precon's
! multigrid branch. With iper=2 in inputinc.d (the periodic box) it reads the coincident
! sets that follow the mesh (:2569-2583) and runs the periodic scheme (nft_init_periodic).
!
! Usage: nftdrv <nlev> <precon> <precision> <lord> <strat> [exit_mode] [itraj]
!   nlev      levels per column when nodes are numbered column by column (0: no columns)
!   precon    1 point relaxation (reference), 2 vertical line (two-point flux), 3 vertical line
!             (in-column band of laplc)
!   precision 0 FP64, 1 mixed, 2 mixed with the FP32 operator inside GCR
!   lord      GCR order; strat: N^2/g (the reference hard-codes these in mpsolver)
!   itraj     0 extrapolated advector (reference), 1 velprd
program nftdrv
  use iso_c_binding
  use nft_hip
  implicit none
  integer :: npoin, nedge, nbface, nbpoin, nboun, i, k, ii, nlev, pk, prec, lord, em, itj
  integer :: ntime, nrkstage, istst, mrestart, isfera, iper, ielast, itr, lordin, itrp, irl
  integer :: mg_flag, itrp1, itrp2, itrp_c1, itrp_c2, itrp_cc, match_flag
  real(8) :: cfl, delt, adapt, relax, p0inf, u0inf, v0inf, w0inf, re, eps0, betap, strat
  real(8) :: sm1, sm2, omega
  real(8) :: ab(6), tw(6)
  type mesh_arrays
    integer :: npoin = 0, nedge = 0, nbface = 0, nboun = 0
    real(8), allocatable :: coord(:,:), vol(:), sn(:,:), sb(:,:), sbb(:,:), anp(:,:)
    integer(c_int), allocatable :: iedge(:,:), kbface(:,:), ibpoin(:,:)
  end type
  type(mesh_arrays), target :: fine, lv(2)
  integer(c_int), allocatable, target :: cols(:), match1(:), match2(:)
  type(nft_mesh_t) :: m1, m2
  type(nft_multigrid_t) :: mg
  type(nft_periodic_t) :: per
  integer(c_int), allocatable, target :: coin(:,:), coin2(:,:), coin4(:,:)
  integer :: ncs(3) = 0
  real(8), allocatable :: state(:,:)
  type(nft_mesh_t) :: m
  type(nft_params_t) :: p
  type(nft_run_stats_t) :: rs
  type(nft_diag_t) :: dg
  real(8) :: ms_total
  type(nft_gcr_stats_t) :: gs
  character(len=64) :: arg
  integer :: nitsm_total, step

  call get_command_argument(1, arg); read(arg,*) nlev
  call get_command_argument(2, arg); read(arg,*) pk
  call get_command_argument(3, arg); read(arg,*) prec
  call get_command_argument(4, arg); read(arg,*) lord
  call get_command_argument(5, arg); read(arg,*) strat
  em = 0
  if (command_argument_count() >= 6) then
    call get_command_argument(6, arg); read(arg,*) em
  end if
  itj = 0
  if (command_argument_count() >= 7) then
    call get_command_argument(7, arg); read(arg,*) itj
  end if

  open(8, file='input_mg_v3.d', status='old')
  read(8,*); read(8,*) mg_flag; read(8,*); read(8,*); read(8,*) sm1, sm2
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
  ab = 0; tw = 0
  if (irl == 1) then
    read(1,*); read(1,*) ab
    read(1,*); read(1,*) tw
  end if
  close(1)
  if (isfera /= 0 .or. (iper /= 0 .and. iper /= 2)) then
    write(*,*) 'nftdrv: only the Cartesian box, closed (iper=0) or periodic (iper=2), is supported'
    error stop 2
  end if

  call read_mesh('mesh3d.d', fine)
  npoin = fine%npoin; nedge = fine%nedge; nbface = fine%nbface; nboun = fine%nboun

  if (nlev > 0) then
    allocate(cols(npoin/nlev + 1))
    do i = 0, npoin/nlev
      cols(i+1) = i*nlev
    end do
  else
    allocate(cols(1)); cols(1) = 0
  end if

  call c_view(fine, m)
  m%ncol = 0
  if (nlev > 0) m%ncol = npoin/nlev
  m%col_start = c_loc(cols)

  p%dt = delt; p%u0 = u0inf; p%v0 = v0inf; p%w0 = w0inf
  p%strat = strat; p%theta0 = 300.0d0
  p%lord = lord; p%itr = itr; p%exit_mode = em; p%eps0 = eps0
  p%precon = pk; p%sm_flag = nint(sm1); p%itrp1 = itrp1; p%omega = omega
  p%dxabl = ab(1); p%dxabr = ab(2); p%dyabd = ab(3); p%dyabu = ab(4); p%dzabd = ab(5); p%dzabu = ab(6)
  p%towxl = tw(1); p%towxr = tw(2); p%towyd = tw(3); p%towyu = tw(4); p%towzd = tw(5); p%towzu = tw(6)
  p%irl = irl
  p%dom = [0.0d0, 15000.0d0, -6000.0d0, 6000.0d0, 0.0d0, 6000.0d0]
  p%precision = prec
  p%itraj = itj

  if (iper == 2) then
    per%ncoin = ncs(1); per%ncoin2 = ncs(2); per%ncoin4 = ncs(3)
    per%ld_coin = size(coin, 1); per%ld_coin2 = size(coin2, 1); per%ld_coin4 = size(coin4, 1)
    per%icoin = c_loc(coin); per%icoin2 = c_loc(coin2); per%icoin4 = c_loc(coin4)
    per%vol_closed = 0
    call nft_check(nft_init_periodic(m, p, per, nft_context), 'nft_init_periodic')
  else
    call nft_check(nft_init(m, p, nft_context), 'nft_init')
  end if
  if (mg_flag == 1) then
    ! precon's multigrid branch: coarse meshes and node maps as `input` reads them
    ! (:2415-2672), including its level-2 metrics quirk (sn_cc y/z from sn_c, :2467-2468)
    call read_mesh('mesh3d_c.d', lv(1))
    call read_mesh('mesh3d_cc.d', lv(2))
    lv(2)%sn(1:lv(2)%nedge, 2:3) = lv(1)%sn(1:lv(2)%nedge, 2:3)
    allocate(match1(lv(1)%npoin), match2(lv(2)%npoin))
    open(55, file='match1.d', status='old')
    read(55,*) match1
    close(55)
    open(56, file='match2.d', status='old')
    read(56,*) match2
    close(56)
    call c_view(lv(1), m1)
    call c_view(lv(2), m2)
    mg%mg_flag = 1; mg%itrp2 = itrp2; mg%itrp_c1 = itrp_c1; mg%itrp_c2 = itrp_c2; mg%itrp_cc = itrp_cc
    mg%transfer = 0; mg%omega_c = 0.0d0
    call nft_check(nft_set_multigrid(nft_context, m1, c_loc(match1), m2, c_loc(match2), mg), &
                   'nft_set_multigrid')
  end if
  call nft_check(nft_state_init(nft_context), 'nft_state_init')
  ! the reference's diagnostics after the initialisation (:270-272) and every step
  ! (:456-459), computed on the GPU
  call nft_check(nft_diagnostics(nft_context, 1, dg), 'diagnostics')
  call nft_print_diag(dg, 7)
  write(*,*) 'end of initialisation, the time loop starts'
  nitsm_total = 0
  ms_total = 0.0d0
  do step = 1, ntime
    call nft_check(nft_run_steps(nft_context, 1, rs), 'nft_run_steps')
    ms_total = ms_total + rs%ms_total
    write(*,*) step, p%dt, p%dt
    call nft_check(nft_diagnostics(nft_context, 1, dg), 'diagnostics')
    call nft_print_diag(dg, 7)
  end do
  write(*,'(a,f12.3,a)') ' GPU time in the time loop ', ms_total, ' ms'
  allocate(state(npoin, 8))
  call nft_check(nft_get_state(nft_context, 'unk'//c_null_char, state(1,1), npoin), 'state')
  call nft_check(nft_get_state(nft_context, 'uuu'//c_null_char, state(1,5), npoin), 'state')
  call nft_check(nft_get_state(nft_context, 'prpot'//c_null_char, state(1,8), npoin), 'state')
  open(30, file='nftdrv_state.bin', access='stream', form='unformatted', status='replace')
  write(30) state
  close(30)
  call nft_check(nft_finalize(nft_context), 'nft_finalize')

contains

  ! a mesh file in the layout `input` reads (:2343-2501), Cartesian and non-periodic
  subroutine read_mesh(fname, mt)
    character(len=*), intent(in) :: fname
    type(mesh_arrays), intent(inout) :: mt
    integer :: nbp, i, k, ii, ldf
    open(5, file=fname, status='old')
    read(5,*) mt%npoin, mt%nedge, mt%nbface, nbp, mt%nboun
    ldf = max(mt%nbface, mt%nboun, 1)
    allocate(mt%coord(mt%npoin,3), mt%vol(mt%npoin), mt%iedge(max(mt%nedge,1),2), mt%sn(max(mt%nedge,1),3))
    allocate(mt%kbface(ldf,2), mt%sb(ldf,3), mt%ibpoin(ldf,2), mt%sbb(ldf,3), mt%anp(ldf,3))
    do i = 1, mt%npoin
      read(5,*) ii, (mt%coord(i,k), k=1,3), mt%vol(i)
    end do
    do i = 1, mt%nedge
      read(5,*) ii, mt%iedge(i,1), mt%iedge(i,2)
    end do
    do i = 1, mt%nedge
      read(5,*) ii, (mt%sn(i,k), k=1,3)
    end do
    do i = 1, mt%nbface
      read(5,*) ii, mt%kbface(i,1), mt%kbface(i,2)
    end do
    do i = 1, mt%nbface
      read(5,*) ii, (mt%sb(i,k), k=1,3)
    end do
    do i = 1, mt%nboun
      read(5,*) ii, mt%ibpoin(i,1), mt%ibpoin(i,2)
    end do
    do i = 1, mt%nboun
      read(5,*) ii, (mt%sbb(i,k), k=1,3)
    end do
    do i = 1, mt%nboun
      read(5,*) ii, (mt%anp(i,k), k=1,3)
    end do
    if (iper == 2 .and. fname == 'mesh3d.d') then
      ! the coincident sets of the periodic box (:2569-2583)
      read(5,*) k
      ncs(1) = k
      allocate(coin(max(k,1),2))
      do i = 1, k
        read(5,*) coin(i,1), coin(i,2)
      end do
      read(5,*) k
      ncs(2) = k
      allocate(coin2(max(k,1),2))
      do i = 1, k
        read(5,*) coin2(i,1), coin2(i,2)
      end do
      read(5,*) k
      ncs(3) = k
      allocate(coin4(max(k,1),4))
      do i = 1, k
        read(5,*) coin4(i,1), coin4(i,2), coin4(i,3), coin4(i,4)
      end do
    end if
    close(5)
  end subroutine

  subroutine c_view(mt, mc)
    type(mesh_arrays), target, intent(in) :: mt
    type(nft_mesh_t), intent(out) :: mc
    mc%npoin = mt%npoin; mc%nedge = mt%nedge; mc%nbface = mt%nbface; mc%nboun = mt%nboun
    mc%ld_poin = mt%npoin; mc%ld_edge = max(mt%nedge,1); mc%ld_face = max(mt%nbface,mt%nboun,1)
    mc%iedge = c_loc(mt%iedge); mc%sn = c_loc(mt%sn); mc%kbface = c_loc(mt%kbface); mc%sb = c_loc(mt%sb)
    mc%ibpoin = c_loc(mt%ibpoin); mc%sbb = c_loc(mt%sbb); mc%anp = c_loc(mt%anp)
    mc%coord = c_loc(mt%coord); mc%vol = c_loc(mt%vol)
    mc%ncol = 0; mc%col_start = c_null_ptr
  end subroutine
end program nftdrv
