! setmhd.f90 -- iso_c_binding interface to the C ABI of include/setmhd.h.
!
! This is the thin layer through which a Fortran driver shaped like SET's
! FCMHD.f90 / cuf_kernel.cuf reaches the gfx950 kernels: Set_CUDA becomes
! setmhd_upload_mesh, the step loop of CUDA_START_MGD becomes setmhd_step,
! Send_data_to_Host becomes setmhd_download.  Arrays are passed as the
! reference declares them (Storage_module.f90:16-37): column-major, 1-based
! numbers.
module setmhd
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: SETMHD_OK = 0
  integer(c_int), parameter :: SETMHD_F_ZERO_BW = 1, SETMHD_F_AXISYM = 2
  integer(c_int), parameter :: SETMHD_F_SOURCE_MK = 4, SETMHD_F_HELIO_HACKS = 8
  integer(c_int), parameter :: SETMHD_F_FAST = 16, SETMHD_F_REORDER = 32, SETMHD_F_PER_CELL = 64

  type, bind(C) :: setmhd_config
    integer(c_int32_t) :: riemann, recon, time_scheme, flags
    real(c_double) :: cfl, dt_cap, r_skip_face, r_hold_cell
    integer(c_int32_t) :: precision, device, tile_x, tile_y, kchunk
  end type setmhd_config

  interface
    subroutine setmhd_config_reference(cfg) bind(C, name="setmhd_config_reference")
      import :: setmhd_config
      type(setmhd_config), intent(out) :: cfg
    end subroutine

    subroutine setmhd_config_cartesian(cfg, cfl, r_hold) bind(C, name="setmhd_config_cartesian")
      import :: setmhd_config, c_double
      type(setmhd_config), intent(out) :: cfg
      real(c_double), value :: cfl, r_hold
    end subroutine

    integer(c_int) function setmhd_create(ctx, cfg) bind(C, name="setmhd_create")
      import :: c_int, c_ptr, setmhd_config
      type(c_ptr), intent(out) :: ctx
      type(setmhd_config), intent(in) :: cfg
    end function

    subroutine setmhd_destroy(ctx) bind(C, name="setmhd_destroy")
      import :: c_ptr
      type(c_ptr), value :: ctx
    end subroutine

    integer(c_int) function setmhd_upload_mesh(ctx, n_cell, n_gran, cell_par, cell_center, &
        cell_volume, cell_gran, gran_normal, gran_square, gran_center, gran_neighbour, &
        gran_neighbour_tvd, gran_type, gran_potok, t0) bind(C, name="setmhd_upload_mesh")
      import :: c_int, c_int32_t, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: n_cell, n_gran
      real(c_double), intent(in) :: cell_par(*), cell_center(*), cell_volume(*)
      integer(c_int32_t), intent(in) :: cell_gran(*)
      real(c_double), intent(in) :: gran_normal(*), gran_square(*), gran_center(*)
      integer(c_int32_t), intent(in) :: gran_neighbour(*), gran_neighbour_tvd(*), gran_type(*)
      real(c_double), intent(in) :: gran_potok(*)
      real(c_double), value :: t0
    end function

    integer(c_int) function setmhd_mesh_order(n_cell, n_gran, cell_center, gran_normal, &
        gran_neighbour, cell_perm, gran_perm) bind(C, name="setmhd_mesh_order")
      import :: c_int, c_int32_t, c_double
      integer(c_int32_t), value :: n_cell, n_gran
      real(c_double), intent(in) :: cell_center(*), gran_normal(*)
      integer(c_int32_t), intent(in) :: gran_neighbour(*)
      integer(c_int32_t), intent(out) :: cell_perm(*), gran_perm(*)
    end function

    integer(c_int) function setmhd_set_dt(ctx, dt) bind(C, name="setmhd_set_dt")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double), value :: dt
    end function

    integer(c_int) function setmhd_step(ctx, nsteps) bind(C, name="setmhd_step")
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: nsteps
    end function

    integer(c_int) function setmhd_sync(ctx, t, dt_next) bind(C, name="setmhd_sync")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: t, dt_next
    end function

    integer(c_int) function setmhd_download(ctx, cell_par, t) bind(C, name="setmhd_download")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double), intent(inout) :: cell_par(*)
      real(c_double), intent(out) :: t
    end function

    integer(c_int) function setmhd_download_async(ctx) bind(C, name="setmhd_download_async")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function

    integer(c_int) function setmhd_download_wait(ctx, cell_par, t) bind(C, name="setmhd_download_wait")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double), intent(inout) :: cell_par(*)
      real(c_double), intent(out) :: t
    end function

    integer(c_int) function setmhd_probe(ctx, n, cells, out, t) bind(C, name="setmhd_probe")
      import :: c_int, c_int32_t, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: n
      integer(c_int32_t), intent(in) :: cells(*)
      real(c_double), intent(inout) :: out(*)
      real(c_double), intent(out) :: t
    end function

    integer(c_int) function setmhd_last_error(ctx, buf, n) bind(C, name="setmhd_last_error")
      import :: c_int, c_ptr, c_char, c_size_t
      type(c_ptr), value :: ctx
      character(kind=c_char), intent(out) :: buf(*)
      integer(c_size_t), value :: n
    end function
  end interface

contains

  ! prints the context's last message and stops, as the reference STOPs
  subroutine setmhd_check(ctx, st, where)
    type(c_ptr), intent(in) :: ctx
    integer(c_int), intent(in) :: st
    character(*), intent(in) :: where
    character(kind=c_char) :: buf(256)
    integer :: i, ierr
    character(len=256) :: msg
    if (st == SETMHD_OK) return
    ierr = setmhd_last_error(ctx, buf, int(256, c_size_t))
    msg = ''
    do i = 1, 256
      if (buf(i) == c_null_char) exit
      msg(i:i) = buf(i)
    end do
    print *, 'setmhd error in ', where, ': status ', st, ' ', trim(msg)
    stop 1
  end subroutine setmhd_check

end module setmhd
