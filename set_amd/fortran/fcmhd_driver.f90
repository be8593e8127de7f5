! fcmhd_driver.f90 -- the solver's Fortran entry point on MI355X.
!
! Same control flow as SET's program MIK (FCMHD.f90:121-250) with
! CUDA_START_MGD (cuf_kernel.cuf:402-949): read the FCMHD input file
! (Set_Storage, Storage_module.f90:43-160), optionally the restart file
! (Fill_data, :191-209), move mesh and state to the device (Set_CUDA),
! run the step loop with the lagged dt starting from 2e-9, print progress
! every print_step steps, bring the state back (Send_data_to_Host) and write
! the output file (Save_Storage, :163-189).  Every kernel launch goes
! through the C ABI (module setmhd).
!
! Observability, in the text formats of FCMHD.f90:6-119: with a probe file
! (one 1-based cell number per line) every print_step the probed densities
! are appended to 1D_rhoT.txt (write_T_rho) through setmhd_probe, one
! gather instead of a read per cell; with 'dump' every print_step writes
! output_data_<k>.txt (x, y, rho: Print_mult) and the end output_data.txt
! (x, y, rho, u, v, p, |V|: Print).  A dump's state is taken with
! setmhd_download_async and written at the next print, so its copy overlaps
! the steps in between.
!
! With 'reorder' the device holds the mesh renumbered for locality
! (SETMHD_F_REORDER); files, probes and dumps keep the input's numbering and
! the results are the same bits.
!
! Restart: the reference resumes from Fill_data's time and state with the
! initial dt 2e-9 (cuf_kernel.cuf:758); so does this driver.  It also writes
! output.bin.dt (the time and the lagged dt the next step would use, two
! REAL(8)), and with 'continue' a restart reads restart.bin.dt and takes that
! dt, so N steps, a restart and M steps give the bits of N+M steps.
!
! usage: fcmhd_driver input.bin output.bin all_step [print_step] [restart.bin|-]
!                     [cartesian|reference] [probes.txt|-] [dump|-] [reorder|-]
!                     [continue]
program fcmhd_driver
  use iso_c_binding
  use setmhd
  implicit none

  real(8) :: host_time_all, vv, t, dt_next, tdump
  integer(4) :: host_N_cell, host_N_gran
  real(8), allocatable :: host_Cell_par(:, :), host_Cell_center(:, :), host_Cell_Volume(:)
  integer(4), allocatable :: host_Cell_gran(:, :)
  real(8), allocatable :: host_Gran_normal(:, :), host_Gran_square(:), host_Gran_center(:, :)
  integer(4), allocatable :: host_Gran_neighbour(:, :), host_Gran_neighbour_TVD(:, :)
  integer(4), allocatable :: host_Gran_type(:)
  real(8), allocatable :: host_Gran_POTOK(:, :)
  character(len=512) :: fin, fout, arg, frestart, fprobes
  integer :: all_step, print_step, step, chunk, nargs, u, ios, nprobe, i, k_dump, itmp
  integer(4), allocatable :: probes(:)
  real(8), allocatable :: pv(:, :)
  logical :: dump
  logical :: cartesian, reorder, cont
  real(8) :: t_dt, dt_start
  type(setmhd_config) :: cfg
  type(c_ptr) :: ctx
  integer(c_int) :: st

  nargs = command_argument_count()
  if (nargs < 3) then
    print *, 'usage: fcmhd_driver input.bin output.bin all_step [print_step] [restart.bin|-] ', &
             '[cartesian|reference] [probes.txt|-] [dump|-] [reorder|-] [continue]'
    stop 2
  end if
  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  call get_command_argument(3, arg)
  read(arg, *) all_step
  print_step = 50000
  if (nargs >= 4) then
    call get_command_argument(4, arg)
    read(arg, *) print_step
  end if
  frestart = ''
  if (nargs >= 5) call get_command_argument(5, frestart)
  if (trim(frestart) == '-') frestart = ''
  cartesian = .false.
  if (nargs >= 6) then
    call get_command_argument(6, arg)
    cartesian = trim(arg) == 'cartesian'
  end if
  fprobes = ''
  if (nargs >= 7) call get_command_argument(7, fprobes)
  if (trim(fprobes) == '-') fprobes = ''
  dump = .false.
  if (nargs >= 8) then
    call get_command_argument(8, arg)
    dump = trim(arg) == 'dump'
  end if
  reorder = .false.
  if (nargs >= 9) then
    call get_command_argument(9, arg)
    reorder = trim(arg) == 'reorder'
  end if
  cont = .false.
  if (nargs >= 10) then
    call get_command_argument(10, arg)
    cont = trim(arg) == 'continue'
  end if
  nprobe = 0
  if (len_trim(fprobes) > 0) then
    open(newunit=u, file=trim(fprobes), action='read', status='old')
    do
      read(u, *, iostat=ios) itmp
      if (ios /= 0) exit
      nprobe = nprobe + 1
    end do
    rewind(u)
    allocate(probes(nprobe), pv(8, nprobe))
    do i = 1, nprobe
      read(u, *) probes(i)
    end do
    close(u)
  end if
  k_dump = 0

  ! ---- Set_Storage --------------------------------------------------------
  open(newunit=u, file=trim(fin), form='unformatted', access='stream', action='read', status='old')
  read(u) host_time_all
  read(u) host_N_cell
  read(u) host_N_gran
  allocate(host_Cell_par(16, host_N_cell), host_Cell_center(3, host_N_cell))
  allocate(host_Cell_Volume(host_N_cell), host_Cell_gran(6, host_N_cell))
  allocate(host_Gran_normal(3, host_N_gran), host_Gran_square(host_N_gran))
  allocate(host_Gran_center(3, host_N_gran), host_Gran_neighbour(2, host_N_gran))
  allocate(host_Gran_neighbour_TVD(2, host_N_gran), host_Gran_type(host_N_gran))
  allocate(host_Gran_POTOK(9, host_N_gran))
  read(u) host_Cell_par
  read(u) host_Cell_center
  read(u) host_Cell_Volume
  read(u) host_Cell_gran
  read(u) host_Gran_normal
  read(u) host_Gran_square
  read(u) host_Gran_center
  read(u) host_Gran_neighbour
  read(u) host_Gran_neighbour_TVD
  read(u) vv
  if (vv /= 121.0_8) stop 'FCMHD input: sentinel 121 missing'
  read(u) host_Gran_type
  read(u) vv
  if (vv /= 122.0_8) stop 'FCMHD input: sentinel 122 missing'
  read(u) host_Gran_POTOK
  read(u) vv
  if (vv /= 123.0_8) stop 'FCMHD input: sentinel 123 missing'
  close(u)
  print *, 'N_cell = ', host_N_cell, ' N_gran = ', host_N_gran

  ! ---- Fill_data (restart) ------------------------------------------------
  if (len_trim(frestart) > 0) then
    open(newunit=u, file=trim(frestart), form='unformatted', access='stream', action='read', status='old')
    read(u) host_time_all
    read(u) host_Cell_par
    read(u) vv
    close(u)
    if (vv /= 321.0_8) stop 'restart file: sentinel 321 missing'
  end if
  dt_start = 0.000000002_8   ! time_step2, cuf_kernel.cuf:758
  if (cont) then
    if (len_trim(frestart) == 0) stop 'continue needs a restart file'
    open(newunit=u, file=trim(frestart)//'.dt', form='unformatted', access='stream', action='read', &
         status='old')
    read(u) t_dt
    read(u) dt_start
    close(u)
    if (t_dt /= host_time_all) stop 'restart file and its .dt disagree on the time'
  end if

  ! ---- Set_CUDA -------------------------------------------------------------
  if (cartesian) then
    call setmhd_config_cartesian(cfg, 0.4_8, 0.1_8)
  else
    call setmhd_config_reference(cfg)
  end if
  if (reorder) cfg%flags = ior(cfg%flags, SETMHD_F_REORDER)
  st = setmhd_create(ctx, cfg)
  if (st /= SETMHD_OK) stop 'setmhd_create failed'
  st = setmhd_upload_mesh(ctx, host_N_cell, host_N_gran, host_Cell_par, host_Cell_center, &
       host_Cell_Volume, host_Cell_gran, host_Gran_normal, host_Gran_square, host_Gran_center, &
       host_Gran_neighbour, host_Gran_neighbour_TVD, host_Gran_type, host_Gran_POTOK, host_time_all)
  call setmhd_check(ctx, st, 'upload_mesh')
  st = setmhd_set_dt(ctx, dt_start)
  call setmhd_check(ctx, st, 'set_dt')

  ! ---- CUDA_START_MGD step loop ----------------------------------------------
  step = 0
  do while (step < all_step)
    chunk = min(print_step, all_step - step)
    st = setmhd_step(ctx, int(chunk, c_int32_t))
    call setmhd_check(ctx, st, 'step')
    step = step + chunk
    st = setmhd_sync(ctx, t, dt_next)
    call setmhd_check(ctx, st, 'sync')
    print *, 'Step = ', step, '  step_time = ', dt_next, ' all time = ', t
    if (nprobe > 0) then
      st = setmhd_probe(ctx, int(nprobe, c_int32_t), probes, pv, t)
      call setmhd_check(ctx, st, 'probe')
      call append_probes(t, pv(1, :))
    end if
    if (dump) then
      if (k_dump > 0) call collect_dump()
      st = setmhd_download_async(ctx)
      call setmhd_check(ctx, st, 'download_async')
      k_dump = k_dump + 1
    end if
  end do
  if (dump .and. k_dump > 0) call collect_dump()

  ! ---- Send_data_to_Host + Save_Storage --------------------------------------
  st = setmhd_download(ctx, host_Cell_par, host_time_all)
  call setmhd_check(ctx, st, 'download')
  st = setmhd_sync(ctx, t, dt_next)
  call setmhd_check(ctx, st, 'sync')
  call setmhd_destroy(ctx)
  open(newunit=u, file=trim(fout)//'.dt', form='unformatted', access='stream', action='write', &
       status='replace')
  write(u) t
  write(u) dt_next
  close(u)
  open(newunit=u, file=trim(fout), form='unformatted', access='stream', action='write', status='replace')
  write(u) host_time_all
  write(u) host_Cell_par
  write(u) 321.0_8
  close(u)
  print *, 'saved ', trim(fout), ' time = ', host_time_all
  if (dump) call dump_state()

contains

  ! writes dump k_dump from the snapshot queued at the previous print
  subroutine collect_dump()
    st = setmhd_download_wait(ctx, host_Cell_par, tdump)
    call setmhd_check(ctx, st, 'download_wait')
    call dump_density(k_dump, tdump)
  end subroutine

  ! SET's dimensional units: the reference scales with default-real
  ! (single-precision) literals, so these are their single-precision values.
  real(8) function sx(v)
    real(8), intent(in) :: v
    sx = v * real(319.319, 8)
  end function

  ! probe series line, the layout of write_T_rho (FCMHD.f90:98-119)
  subroutine append_probes(tt, rho)
    real(8), intent(in) :: tt, rho(:)
    integer :: w
    open(newunit=w, file='1D_rhoT.txt', position='append', action='write', status='unknown')
    write(w, '(18es23.15)') tt * real(146.262, 8), rho * real(0.073, 8)
    close(w)
  end subroutine

  ! density dump, the layout of Print_mult (FCMHD.f90:52-96)
  subroutine dump_density(k, tt)
    integer, intent(in) :: k
    real(8), intent(in) :: tt
    character(len=64) :: name
    integer :: w, c
    write(name, '(A, I0, A)') 'output_data_', k, '.txt'
    open(newunit=w, file=trim(name), status='replace', action='write')
    write(w, '(A)') 'TITLE = "HP"'
    write(w, '(A)') 'VARIABLES = x, y, rho'
    write(w, '(A, ES12.5)') 'ZONE T="HP", STRANDID=1, SOLUTIONTIME=', tt * real(146.262, 8)
    do c = 1, host_N_cell
      write(w, '(3(ES15.7, 2X))') sx(host_Cell_center(1, c)), sx(host_Cell_center(2, c)), &
          host_Cell_par(1, c) * real(0.073, 8)
    end do
    close(w)
  end subroutine

  ! final state, the layout of Print (FCMHD.f90:6-50)
  subroutine dump_state()
    integer :: w, c
    real(8) :: uu, vv2
    open(newunit=w, file='output_data.txt', status='replace', action='write')
    write(w, '(A)') 'TITLE = "HP"  VARIABLES = x, y, rho, u, v, p, |V|'
    do c = 1, host_N_cell
      uu = host_Cell_par(2, c) * real(10.3565, 8)
      vv2 = host_Cell_par(3, c) * real(10.3565, 8)
      write(w, '(6(ES15.7, 2X))') sx(host_Cell_center(1, c)), sx(host_Cell_center(2, c)), &
          host_Cell_par(1, c) * real(0.073, 8), uu, vv2, host_Cell_par(5, c) * real(1.30962, 8), &
          sqrt(uu**2 + vv2**2)
    end do
    close(w)
  end subroutine
end program fcmhd_driver
