! cgl_mod.f90 -- Fortran binding of libcgl's C ABI (include/cgl.h).
!
! The reference calls its lattice-update subroutines with implicit
! interfaces and whole arrays by address (Source1.f90 S:216-251).  This
! module gives the same call site explicit iso_c_binding interfaces to the
! GPU library: host arrays keep the reference's layout, ri(0:18,zl,yl,xl)
! and rho(zl,yl,xl), and are passed by address.
module cgl
  use iso_c_binding
  implicit none

  integer(c_int32_t), parameter :: CGL_BC_PERIODIC = 0, CGL_BC_WALL = 1, CGL_BC_MIRROR = 2
  integer(c_int32_t), parameter :: CGL_PREC_F64 = 0, CGL_PREC_F32 = 1, CGL_PREC_F16 = 2
  integer(c_int32_t), parameter :: CGL_LAYOUT_RB = 0, CGL_LAYOUT_FS = 1
  integer(c_int32_t), parameter :: CGL_OK = 0, CGL_ERR_ARG = -1, CGL_ERR_HIP = -2, CGL_ERR_UNSUPPORTED = -3, &
                                   CGL_ERR_RANGE = -4

  type, bind(C) :: cgl_params
    integer(c_int32_t) :: nx, ny, nz
    integer(c_int32_t) :: bc(6)
    integer(c_int32_t) :: precision, npar, xchunk, device, layout, tile_rows
    real(c_double) :: sigma, rho0, nu_r, nu_b, beta
  end type cgl_params

  ! one box of a batch (cgl_batch_create)
  type, bind(C) :: cgl_box
    integer(c_int32_t) :: nx, ny, nz
    integer(c_int32_t) :: reserved = 0
    real(c_double) :: nu_r, nu_b
  end type cgl_box

  interface
    integer(c_int) function cgl_create(h, p) bind(C, name="cgl_create")
      import :: c_int, c_ptr, cgl_params
      type(c_ptr), intent(out) :: h
      type(cgl_params), intent(in) :: p
    end function
    integer(c_int) function cgl_destroy(h) bind(C, name="cgl_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    integer(c_int) function cgl_upload_state(h, ri, bi) bind(C, name="cgl_upload_state")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(in) :: ri(*), bi(*)
    end function
    integer(c_int) function cgl_download_state(h, ri, bi) bind(C, name="cgl_download_state")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: ri(*), bi(*)
    end function
    integer(c_int) function cgl_init_cylinder(h, xc, yc, r0, kprof, nmodes, amp) bind(C, name="cgl_init_cylinder")
      import :: c_int, c_ptr, c_double, c_int32_t
      type(c_ptr), value :: h
      real(c_double), value :: xc, yc, r0, kprof
      integer(c_int32_t), value :: nmodes
      real(c_double), intent(in) :: amp(*)
    end function
    integer(c_int) function cgl_step(h, nsteps) bind(C, name="cgl_step")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: h
      integer(c_int32_t), value :: nsteps
    end function
    integer(c_int) function cgl_calmacros(h, rho, rho_r, rho_b, ux, uy, uz, phi) bind(C, name="cgl_calmacros")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: rho(*), rho_r(*), rho_b(*), ux(*), uy(*), uz(*), phi(*)
    end function
    integer(c_int) function cgl_sync(h) bind(C, name="cgl_sync")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    integer(c_int) function cgl_time_steps(h, nsteps, ms_total, ms_moments, ms_collide) bind(C, name="cgl_time_steps")
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: h
      integer(c_int32_t), value :: nsteps
      real(c_double), intent(out) :: ms_total, ms_moments, ms_collide
    end function
    integer(c_int) function cgl_set_particles(h, center, up, rp, rad, mass, inertia, cita, fg) &
        bind(C, name="cgl_set_particles")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(in) :: center(*), up(*), rp(*), rad(*), mass(*), inertia(*), cita(*)
      real(c_double), value :: fg
    end function
    integer(c_int) function cgl_set_particle_geometry(h, rep_plane, rep_on, dom_lo, dom_hi, dom_on, sym_rules) &
        bind(C, name="cgl_set_particle_geometry")
      import :: c_int, c_ptr, c_double, c_int32_t
      type(c_ptr), value :: h
      real(c_double), intent(in) :: rep_plane(6), dom_lo(3), dom_hi(3)
      integer(c_int32_t), intent(in) :: rep_on(6), dom_on(9)
      integer(c_int32_t), value :: sym_rules
    end function
    integer(c_int) function cgl_checkpoint_save(h, path) bind(C, name="cgl_checkpoint_save")
      import :: c_int, c_ptr, c_char
      type(c_ptr), value :: h
      character(kind=c_char), intent(in) :: path(*)   ! NUL-terminated: trim(name)//c_null_char
    end function
    integer(c_int) function cgl_checkpoint_load(h, path) bind(C, name="cgl_checkpoint_load")
      import :: c_int, c_ptr, c_char
      type(c_ptr), value :: h
      character(kind=c_char), intent(in) :: path(*)
    end function
    integer(c_int) function cgl_get_particles(h, center, up, rp, force, torque, fcap, tcap, frep) &
        bind(C, name="cgl_get_particles")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: center(*), up(*), rp(*), force(*), torque(*), fcap(*), tcap(*), frep(*)
    end function
    integer(c_int) function cgl_get_particle_shb(h, shb) bind(C, name="cgl_get_particle_shb")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: shb(*)
    end function
    ! each mask by c_loc of an int32 (zl, yl, xl) array, or c_null_ptr to skip it
    integer(c_int) function cgl_download_masks(h, walls, walls_all, pin, pout) bind(C, name="cgl_download_masks")
      import :: c_int, c_ptr
      type(c_ptr), value :: h, walls, walls_all, pin, pout
    end function
    integer(c_int) function cgl_init_interface(h, z0, kprof) bind(C, name="cgl_init_interface")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), value :: z0, kprof
    end function
    ! cgl_calmacros with each output by c_loc, or c_null_ptr to skip it: the
    ! opening stage of a staged step needs no output at all
    integer(c_int) function cgl_calmacros_ptr(h, rho, rho_r, rho_b, ux, uy, uz, phi) bind(C, name="cgl_calmacros")
      import :: c_int, c_ptr
      type(c_ptr), value :: h, rho, rho_r, rho_b, ux, uy, uz, phi
    end function
    integer(c_int) function cgl_derivatives(h, gradx, grady, gradz, n_x, n_y, n_z, flag) &
        bind(C, name="cgl_derivatives")
      import :: c_int, c_ptr, c_double, c_int32_t
      type(c_ptr), value :: h
      real(c_double), intent(out) :: gradx(*), grady(*), gradz(*), n_x(*), n_y(*), n_z(*)
      integer(c_int32_t), intent(out) :: flag(*)
    end function
    integer(c_int) function cgl_tuned_xchunk(h, xchunk, tuning) bind(C, name="cgl_tuned_xchunk")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: h
      integer(c_int32_t), intent(out) :: xchunk, tuning
    end function
    integer(c_int) function cgl_tuned_params(h, xchunk, xcd_order, tile_rows, tuning) bind(C, name="cgl_tuned_params")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: h
      integer(c_int32_t), intent(out) :: xchunk, xcd_order, tile_rows, tuning
    end function
    integer(c_int) function cgl_phase_field_form(h, scatter) bind(C, name="cgl_phase_field_form")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: h
      integer(c_int32_t), intent(out) :: scatter
    end function
    ! the stages of a staged step, in order after cgl_calmacros (include/cgl.h)
    integer(c_int) function cgl_capillary(h) bind(C, name="cgl_capillary")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    integer(c_int) function cgl_momentum(h) bind(C, name="cgl_momentum")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    integer(c_int) function cgl_repulsive(h) bind(C, name="cgl_repulsive")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    integer(c_int) function cgl_move(h) bind(C, name="cgl_move")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    integer(c_int) function cgl_set_walls(h) bind(C, name="cgl_set_walls")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    integer(c_int) function cgl_refill(h) bind(C, name="cgl_refill")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    integer(c_int) function cgl_collide_stream(h) bind(C, name="cgl_collide_stream")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    ! each field by c_loc of a double (zl, yl, xl) array, or c_null_ptr to skip it
    integer(c_int) function cgl_particle_band(h, rho, rho_r, rho_b, ux, uy, uz) bind(C, name="cgl_particle_band")
      import :: c_int, c_ptr
      type(c_ptr), value :: h, rho, rho_r, rho_b, ux, uy, uz
    end function
    ! sums(7); slice_red by c_loc of a double (zl) array, or c_null_ptr
    integer(c_int) function cgl_diagnostics(h, sums, slice_red) bind(C, name="cgl_diagnostics")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: sums(7)
      type(c_ptr), value :: slice_red
    end function
    integer(c_int) function cgl_diagnostics_noalloc(h, sums, slice_red) bind(C, name="cgl_diagnostics_noalloc")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: sums(7)
      type(c_ptr), value :: slice_red
    end function
    integer(c_int) function cgl_batch_create(h, p, nbox, boxes) bind(C, name="cgl_batch_create")
      import :: c_int, c_ptr, c_int32_t, cgl_params, cgl_box
      type(c_ptr), intent(out) :: h
      type(cgl_params), intent(in) :: p
      integer(c_int32_t), value :: nbox
      type(cgl_box), intent(in) :: boxes(*)
    end function
    ! box indices are 0-based, as in the C ABI
    integer(c_int) function cgl_box_upload_state(h, box, ri, bi) bind(C, name="cgl_box_upload_state")
      import :: c_int, c_ptr, c_double, c_int32_t
      type(c_ptr), value :: h
      integer(c_int32_t), value :: box
      real(c_double), intent(in) :: ri(*), bi(*)
    end function
    integer(c_int) function cgl_box_init_cylinder(h, box, xc, yc, r0, kprof, nmodes, amp) &
        bind(C, name="cgl_box_init_cylinder")
      import :: c_int, c_ptr, c_double, c_int32_t
      type(c_ptr), value :: h
      integer(c_int32_t), value :: box
      real(c_double), value :: xc, yc, r0, kprof
      integer(c_int32_t), value :: nmodes
      real(c_double), intent(in) :: amp(*)
    end function
    integer(c_int) function cgl_box_download_state(h, box, ri, bi) bind(C, name="cgl_box_download_state")
      import :: c_int, c_ptr, c_double, c_int32_t
      type(c_ptr), value :: h
      integer(c_int32_t), value :: box
      real(c_double), intent(out) :: ri(*), bi(*)
    end function
    integer(c_int) function cgl_box_calmacros(h, box, rho, rho_r, rho_b, ux, uy, uz, phi) &
        bind(C, name="cgl_box_calmacros")
      import :: c_int, c_ptr, c_double, c_int32_t
      type(c_ptr), value :: h
      integer(c_int32_t), value :: box
      real(c_double), intent(out) :: rho(*), rho_r(*), rho_b(*), ux(*), uy(*), uz(*), phi(*)
    end function
    integer(c_int) function cgl_probe_copy_bandwidth(device, bytes, reps, gbps) bind(C, name="cgl_probe_copy_bandwidth")
      import :: c_int, c_int32_t, c_int64_t, c_double
      integer(c_int32_t), value :: device, reps
      integer(c_int64_t), value :: bytes
      real(c_double), intent(out) :: gbps
    end function
    integer(c_int) function cgl_device_name(h, buf, len) bind(C, name="cgl_device_name")
      import :: c_int, c_ptr, c_char, c_int32_t
      type(c_ptr), value :: h
      character(kind=c_char), intent(out) :: buf(*)
      integer(c_int32_t), value :: len
    end function
    type(c_ptr) function cgl_last_error_c() bind(C, name="cgl_last_error")
      import :: c_ptr
    end function
  end interface

contains

  ! error stop with the library's message when rc /= 0 (the reference has no
  ! error handling; the drop-in stops instead of continuing on bad state)
  subroutine cgl_check(rc, what)
    integer(c_int), intent(in) :: rc
    character(*), intent(in) :: what
    character(kind=c_char), pointer :: msg(:)
    integer :: n
    if (rc == 0) return
    call c_f_pointer(cgl_last_error_c(), msg, [512])
    n = 0
    do while (n < 512)
      if (msg(n + 1) == c_null_char) exit
      n = n + 1
    end do
    write(*, '(a, a, i0, a, 512a)') what, ': libcgl error ', rc, ': ', msg(1:n)
    error stop 1
  end subroutine cgl_check
end module cgl
