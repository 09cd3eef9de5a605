! cgl_driver.f90 -- host driver of the GPU lattice update.
!
! Rebuilds the reference's `program main` (Source1.f90 S:28-253) around
! libcgl: the physical parameters keep the reference's names (sigma, rho0,
! tau_R, tau_B, beta, Rad_c; S:57-84), host fields keep its layout
! ri(0:18,zl,yl,xl), rho_N(zl,yl,xl), and the field dump keeps the format of
! output_results (OUT_<t>.dat, S:898-924).  The time loop (S:215-252) becomes
! cgl_step calls between output steps; fields cross to the host only at the
! output cadence.
!
! Input: an optional namelist file (first command-line argument) with group
! &cgl (see the declarations below).  The initial state is the reference's
! initialization (S:850-896) for a liquid cylinder along z through the box
! centre with nmodes cosine modes of wavelength zl (ic = 'cylinder'), or a
! flat interface at z = z0 with red fluid below (ic = 'interface').  With
! zwalls the z faces are half-way bounce-back walls with ghost layers (the
! reference's z=1 kind, S:1557-1590); ic_on_device builds either initial
! state on the device.  npar > 0 adds spheres, their arrays allocated to
! npar: centres px, py, pz, radius p_rad, density p_rho and contact angle
! p_cita (degrees) per sphere, from a second namelist group &particles in
! the same file (p_rad, p_rho, p_cita default to the scalars rad_p, rho_p,
! cita_deg of &cgl), or from particle_file, one sphere per line:
! x y z radius density angle.  Mass and moment of inertia as S:81-82, Fg as
! S:84 with the first sphere's radius; their state is appended to
! PARTICLES.dat at every output step.  nbox > 0 runs a batch of independent
! periodic boxes instead, one per line of box_file:
! nx ny nz tau_R tau_B R nmodes amp(1..nmodes), each a z-axis cylinder of
! radius R built on the device; BOXES.dat gets the red and blue mass and the
! largest speed of every box at every output step, and dump_binary the
! fields of the first dump_boxes boxes (all when negative).  ref_faces selects the
! reference's own box: walls at x = 1 and z = 1, mirrors at x = xl, y = 1,
! y = yl and z = zl (S:1443-1592), and for particles its repulsion planes
! (S:1607-1612), Lagrange in-domain box (S:1832) and symmetry-plane rules of
! movement (S:1272-1335).  ckpt_every > 0 writes a checkpoint CKPT_<t>.bin
! every ckpt_every steps; restart = 'file' with restart_step = t resumes from
! one (after the same set-up, which the checkpoint then overwrites).
program cgl_driver
  use iso_c_binding
  use cgl
  implicit none
  integer :: xl = 32, yl = 32, zl = 128, nsteps = 2000, out_every = 100, prec = 0, xchunk = 0, nmodes = 1, layout = 0
  integer :: tile_rows = 0
  double precision :: sigma = 0.01d0, rho0 = 1.0d0, tau_R = 1.0d0, tau_B = 1.0d0, beta = 0.7d0, Rad_c = 8.0d0
  double precision :: amp(16) = 0.0d0
  logical :: ic_on_device = .false., write_tecplot = .true., dump_binary = .false., zwalls = .false.
  logical :: ref_faces = .false.
  integer :: ckpt_every = 0, restart_step = 0
  character(len=256) :: restart = ' '
  character(len=300) :: ckpt_name
  character(len=256) :: outdir = '.'
  character(len=16) :: ic = 'cylinder'
  integer :: npar = 0, nbox = 0, dump_boxes = -1
  double precision :: z0 = 0.0d0, rad_p = 10.0d0, rho_p = 1.5d0, cita_deg = 90.0d0
  character(len=256) :: particle_file = ' ', box_file = ' '
  ! per-sphere set-up, allocated to npar: centre, radius, density, contact angle (degrees)
  double precision, allocatable :: px(:), py(:), pz(:), p_rad(:), p_rho(:), p_cita(:)
  namelist /cgl/ xl, yl, zl, nsteps, out_every, prec, xchunk, sigma, rho0, tau_R, tau_B, beta, Rad_c, &
                 nmodes, amp, ic_on_device, write_tecplot, dump_binary, outdir, layout, ic, z0, zwalls, &
                 npar, rad_p, rho_p, cita_deg, particle_file, ref_faces, ckpt_every, restart, restart_step, &
                 nbox, box_file, dump_boxes, tile_rows
  namelist /particles/ px, py, pz, p_rad, p_rho, p_cita

  double precision, allocatable :: ri(:,:,:,:), bi(:,:,:,:)
  double precision, allocatable :: rho(:,:,:), rho_R(:,:,:), rho_B(:,:,:), rho_N(:,:,:), ux(:,:,:), uy(:,:,:), uz(:,:,:)
  integer(c_int32_t), allocatable, target :: walls_all(:,:,:), pout(:,:,:)
  ! particle state in the reference's layout: center(par,3) etc. (S:45-50)
  double precision, allocatable :: center(:,:), up(:,:), rp(:,:), fp(:,:), tp(:,:), fcap(:,:), tcap(:,:), frep(:,:)
  double precision, allocatable :: prad(:), pmass(:), pinert(:), pcita(:)
  double precision :: fg
  double precision :: wi(0:18), nu_R, nu_B, ms(3)
  double precision, parameter :: PI = 4.0d0*datan(1.0d0)
  type(cgl_params) :: p
  type(c_ptr) :: h
  character(len=512) :: arg
  integer :: t, nchunk, u, ios
  double precision :: wcap_old = 0.0d0   ! output4's accumulated capillary work (S:202)
  integer :: t_breakup = -1              ! first output step with the jet's axis blue (ic = 'cylinder')

  amp(1) = 0.8d0
  if (command_argument_count() >= 1) then
    call get_command_argument(1, arg)
    open(newunit=u, file=trim(arg), status='old', action='read', iostat=ios)
    if (ios /= 0) then
      write(*,*) 'cannot open ', trim(arg)
      error stop 2
    end if
    read(u, nml=cgl)
    close(u)
  end if
  if (npar < 0) then
    write(*,*) 'npar must not be negative'
    error stop 2
  end if
  if (npar > 0) call read_particles()

  ! relaxation times to viscosities (inverse of S:71-72)
  nu_R = (tau_R - 0.5d0)/3.0d0
  nu_B = (tau_B - 0.5d0)/3.0d0
  wi(0) = 1.0d0/3.0d0
  wi(1:6) = 1.0d0/18.0d0
  wi(7:18) = 1.0d0/36.0d0
  write(*,*) 'box', xl, yl, zl, ' nu_R=', nu_R, ' nu_B=', nu_B, ' precision', prec

  p%nx = xl; p%ny = yl; p%nz = zl
  p%bc = CGL_BC_PERIODIC
  if (zwalls) p%bc(5:6) = CGL_BC_WALL
  if (ref_faces) p%bc = [CGL_BC_WALL, CGL_BC_MIRROR, CGL_BC_MIRROR, CGL_BC_MIRROR, CGL_BC_WALL, CGL_BC_MIRROR]
  p%precision = prec
  p%npar = npar
  p%xchunk = xchunk
  p%device = 0
  p%layout = layout
  p%tile_rows = tile_rows
  p%sigma = sigma; p%rho0 = rho0; p%nu_r = nu_R; p%nu_b = nu_B; p%beta = beta
  if (nbox > 0) then
    call run_batch()
    stop
  end if
  call cgl_check(cgl_create(h, p), 'cgl_create')

  allocate(rho(zl,yl,xl), rho_R(zl,yl,xl), rho_B(zl,yl,xl), rho_N(zl,yl,xl), ux(zl,yl,xl), uy(zl,yl,xl), uz(zl,yl,xl))
  allocate(walls_all(zl,yl,xl), pout(zl,yl,xl))
  walls_all = 0
  pout = 0
  if (ic_on_device .and. trim(ic) == 'cylinder') then
    call cgl_check(cgl_init_cylinder(h, dble(xl)/2.0d0 + 0.5d0, dble(yl)/2.0d0 + 0.5d0, Rad_c, &
                   6.0d0*beta*0.134d0, int(nmodes, c_int32_t), amp), 'cgl_init_cylinder')
  else if (ic_on_device .and. trim(ic) == 'interface') then
    call cgl_check(cgl_init_interface(h, z0, 6.0d0*beta*0.134d0), 'cgl_init_interface')
  else
    allocate(ri(0:18,zl,yl,xl), bi(0:18,zl,yl,xl))
    if (trim(ic) == 'interface') then
      call interface_state()
    else
      call initial_state()
    end if
    call cgl_check(cgl_upload_state(h, ri, bi), 'cgl_upload_state')
    deallocate(ri, bi)
  end if
  if (npar > 0) then
    allocate(center(npar,3), up(npar,3), rp(npar,3), fp(npar,3), tp(npar,3), fcap(npar,3), tcap(npar,3), &
             frep(npar,3), prad(npar), pmass(npar), pinert(npar), pcita(npar))
    center(:,1) = px; center(:,2) = py; center(:,3) = pz
    up = 0.0d0; rp = 0.0d0
    prad = p_rad
    pmass = 4.0d0/3.0d0*p_rho*PI*p_rad**3            ! S:81
    pinert = 2.0d0/5.0d0*pmass*p_rad**2              ! S:82
    pcita = p_cita/180.0d0*PI
    fg = -PI*2.0d0*p_rad(1)*sigma                    ! S:84
    call cgl_check(cgl_set_particles(h, center, up, rp, prad, pmass, pinert, pcita, fg), 'cgl_set_particles')
    if (ref_faces) call cgl_check(cgl_set_particle_geometry(h, &
        [1.5d0, dble(xl) - 1.5d0, 1.5d0, dble(yl) - 1.5d0, 1.5d0, dble(zl) - 0.5d0], [1, 1, 1, 1, 1, 1], &
        [0.0d0, 1.5d0, 0.0d0], [dble(xl) - 0.5d0, dble(yl) - 0.5d0, dble(zl) - 0.5d0], &
        [0, 0, 1, 1, 0, 1, 0, 0, 1], 1_c_int32_t), 'cgl_set_particle_geometry')
    open(newunit=u, file=trim(outdir)//'/PARTICLES.dat', status='replace')
    write(u, '(a)') 'VARIABLES=T,P,X,Y,Z,U,V,W,FX,FY,FZ,FCX,FCY,FCZ'
    close(u)
  end if

  t = 0
  if (len_trim(restart) > 0) then
    call cgl_check(cgl_checkpoint_load(h, trim(restart)//c_null_char), 'cgl_checkpoint_load')
    t = restart_step
  end if
  do
    if (ckpt_every > 0 .and. t > restart_step .and. mod(t, max(ckpt_every, 1)) == 0) then
      write(ckpt_name, '(a, "/CKPT_", i0, ".bin")') trim(outdir), t
      call cgl_check(cgl_checkpoint_save(h, trim(ckpt_name)//c_null_char), 'cgl_checkpoint_save')
    end if
    if (mod(t, out_every) == 0 .or. t == nsteps) then
      call cgl_check(cgl_calmacros(h, rho, rho_R, rho_B, ux, uy, uz, rho_N), 'cgl_calmacros')
      write(*,'(i8, 3(1x, es22.15))') t, sum(rho_R), sum(rho_B), maxval(sqrt(ux**2 + uy**2 + uz**2))
      if (npar > 0) call output_particles(t)
      if (npar > 0) call output_result2(t)
      if (ref_faces) call output3(t)
      if (trim(ic) == 'cylinder') call output_neck(t)
      call output4(t)
      if (write_tecplot) call output_results(t)
      if (dump_binary) call dump_fields(t)
    end if
    if (t >= nsteps) exit
    nchunk = min(out_every - mod(t, out_every), nsteps - t)
    if (ckpt_every > 0) nchunk = min(nchunk, ckpt_every - mod(t, ckpt_every))
    call cgl_check(cgl_step(h, int(nchunk, c_int32_t)), 'cgl_step')
    t = t + nchunk
  end do
  call cgl_check(cgl_sync(h), 'cgl_sync')
  call cgl_check(cgl_destroy(h), 'cgl_destroy')

contains

  ! the spheres' set-up: particle_file, or the &particles group of the
  ! namelist file
  subroutine read_particles()
    integer :: k, up_, ios_
    allocate(px(npar), py(npar), pz(npar), p_rad(npar), p_rho(npar), p_cita(npar))
    px = 0.0d0; py = 0.0d0; pz = 0.0d0
    p_rad = rad_p; p_rho = rho_p; p_cita = cita_deg
    if (len_trim(particle_file) > 0) then
      open(newunit=up_, file=trim(particle_file), status='old', action='read', iostat=ios_)
      if (ios_ /= 0) then
        write(*,*) 'cannot open ', trim(particle_file)
        error stop 2
      end if
      do k = 1, npar
        read(up_, *, iostat=ios_) px(k), py(k), pz(k), p_rad(k), p_rho(k), p_cita(k)
        if (ios_ /= 0) then
          write(*,*) trim(particle_file), ': fewer than npar =', npar, 'lines of x y z radius density angle'
          error stop 2
        end if
      end do
      close(up_)
    else
      open(newunit=up_, file=trim(arg), status='old', action='read')
      read(up_, nml=particles, iostat=ios_)
      close(up_)
      if (ios_ /= 0) then
        write(*,*) 'npar > 0 needs particle_file or a &particles group (px, py, pz, ...)'
        error stop 2
      end if
    end if
  end subroutine read_particles

  ! a batch of independent periodic boxes in one handle (cgl_batch_create):
  ! z-axis cylinders built on the device, per-box output lines
  subroutine run_batch()
    type(cgl_box), allocatable :: boxes(:)
    integer, allocatable :: nm(:)
    double precision, allocatable :: r0(:), am(:,:), tr_(:), tb_(:)
    double precision, allocatable :: f(:,:,:,:)
    integer :: b, k, ub, uo, tt, nc
    character(len=64) :: name
    allocate(boxes(nbox), nm(nbox), r0(nbox), am(16, nbox), tr_(nbox), tb_(nbox))
    am = 0.0d0
    open(newunit=ub, file=trim(box_file), status='old', action='read')
    do b = 1, nbox
      read(ub, *) boxes(b)%nx, boxes(b)%ny, boxes(b)%nz, tr_(b), tb_(b), r0(b), nm(b), (am(k, b), k = 1, nm(b))
      boxes(b)%reserved = 0
      boxes(b)%nu_r = (tr_(b) - 0.5d0)/3.0d0
      boxes(b)%nu_b = (tb_(b) - 0.5d0)/3.0d0
    end do
    close(ub)
    p%bc = CGL_BC_PERIODIC
    p%npar = 0
    call cgl_check(cgl_batch_create(h, p, int(nbox, c_int32_t), boxes), 'cgl_batch_create')
    do b = 1, nbox
      call cgl_check(cgl_box_init_cylinder(h, int(b - 1, c_int32_t), dble(boxes(b)%nx)/2.0d0 + 0.5d0, &
                     dble(boxes(b)%ny)/2.0d0 + 0.5d0, r0(b), 6.0d0*beta*0.134d0, int(nm(b), c_int32_t), &
                     am(:, b)), 'cgl_box_init_cylinder')
    end do
    open(newunit=uo, file=trim(outdir)//'/BOXES.dat', status='replace')
    write(uo, '(a)') 'VARIABLES=T,BOX,MASS_R,MASS_B,UMAX'
    close(uo)
    tt = 0
    do
      if (mod(tt, out_every) == 0 .or. tt == nsteps) then
        open(newunit=uo, file=trim(outdir)//'/BOXES.dat', status='old', position='append')
        do b = 1, nbox
          allocate(f(boxes(b)%nz, boxes(b)%ny, boxes(b)%nx, 7))
          call cgl_check(cgl_box_calmacros(h, int(b - 1, c_int32_t), f(:,:,:,1), f(:,:,:,2), f(:,:,:,3), &
                         f(:,:,:,4), f(:,:,:,5), f(:,:,:,6), f(:,:,:,7)), 'cgl_box_calmacros')
          write(uo, '(i8, 1x, i4, 3(1x, es24.16e3))') tt, b, sum(f(:,:,:,2)), sum(f(:,:,:,3)), &
                maxval(sqrt(f(:,:,:,4)**2 + f(:,:,:,5)**2 + f(:,:,:,6)**2))
          if (dump_binary .and. (dump_boxes < 0 .or. b <= dump_boxes)) then
            write(name, '("/BOX_", i0, "_FIELDS_", i0, ".bin")') b, tt
            open(newunit=ub, file=trim(outdir)//trim(name), access='stream', form='unformatted', status='replace')
            write(ub) f
            close(ub)
          end if
          deallocate(f)
        end do
        close(uo)
      end if
      if (tt >= nsteps) exit
      nc = min(out_every - mod(tt, out_every), nsteps - tt)
      call cgl_check(cgl_step(h, int(nc, c_int32_t)), 'cgl_step')
      tt = tt + nc
    end do
    call cgl_check(cgl_sync(h), 'cgl_sync')
    call cgl_check(cgl_destroy(h), 'cgl_destroy')
  end subroutine run_batch

  ! initialization (S:850-896) for a z-axis cylinder through the box centre
  subroutine initial_state()
    integer :: x, y, z, m
    double precision :: x0, y0, dis, dis_x, star, plus, ph, rh, rr
    x0 = dble(xl)/2.0d0 + 0.5d0
    y0 = dble(yl)/2.0d0 + 0.5d0
    do x = 1, xl
      do y = 1, yl
        do z = 1, zl
          dis = (dble(x) - x0)**2 + (dble(y) - y0)**2
          dis_x = 0.0d0
          do m = 1, nmodes
            dis_x = dis_x + amp(m)*cos(2.0d0*PI*dble(m)*(dble(z) - 1.0d0)/dble(zl))
          end do
          star = Rad_c + dis_x - sqrt(dis)
          ph = tanh(star*6.0d0*beta*0.134d0)
          plus = sigma/(Rad_c + dis_x)*3.0d0
          if (ph > 0.0d0) then
            rh = rho0 + plus
          else
            rh = rho0
          end if
          rr = rh*(ph + 1.0d0)/2.0d0
          ri(:,z,y,x) = rr*wi*1.0d0
          bi(:,z,y,x) = (rh - rr)*wi*1.0d0
        end do
      end do
    end do
  end subroutine initial_state

  ! flat interface at z = z0, red below, at rest; same profile and
  ! equilibria as initialization (S:874, S:883-889)
  subroutine interface_state()
    integer :: x, y, z
    double precision :: ph, rr
    do x = 1, xl
      do y = 1, yl
        do z = 1, zl
          ph = tanh((z0 - dble(z))*6.0d0*beta*0.134d0)
          rr = rho0*(ph + 1.0d0)/2.0d0
          ri(:,z,y,x) = rr*wi*1.0d0
          bi(:,z,y,x) = (rho0 - rr)*wi*1.0d0
        end do
      end do
    end do
  end subroutine interface_state

  ! particle state and masks at output steps
  subroutine output_particles(tt)
    integer, intent(in) :: tt
    integer :: k, uo
    call cgl_check(cgl_get_particles(h, center, up, rp, fp, tp, fcap, tcap, frep), 'cgl_get_particles')
    call cgl_check(cgl_download_masks(h, c_null_ptr, c_loc(walls_all), c_null_ptr, c_loc(pout)), 'cgl_download_masks')
    open(newunit=uo, file=trim(outdir)//'/PARTICLES.dat', status='old', position='append')
    do k = 1, npar
      write(uo, '(i8, 1x, i4, 12(1x, es24.16e3))') tt, k, center(k,:), up(k,:), fp(k,:), fcap(k,:)
    end do
    close(uo)
  end subroutine output_particles

  ! The neck of a z-axis jet (ic = 'cylinder'), the quantity output3 follows
  ! for the reference's y-axis thread (S:1673-1715), restated for a jet
  ! along z: per z slice R(z) = sqrt(red area / pi), the red area being the
  ! device sum of (1 + phi)/2 over the slice (cgl_diagnostics).  NECK.dat
  ! gets t, t*, the neck radius min R, max R, the neck's z (1-based), the
  ! amplitude (max R - min R)/2 and, at the z where it is lowest, phi on the
  ! axis (the mean over the 2 x 2 nodes nearest it).  The first output step
  ! at which that phi is below 0 (the axis has turned blue: the thread has
  ! broken) is reported once.
  subroutine output_neck(tt)
    integer, intent(in) :: tt
    real(c_double), target :: area(zl)
    real(c_double) :: sums(7)
    double precision :: r(zl), ax(zl), tr
    integer :: z, zn, uo, xa, ya
    call cgl_check(cgl_diagnostics(h, sums, c_loc(area)), 'cgl_diagnostics')
    r = sqrt(area/PI)
    xa = xl/2
    ya = yl/2
    do z = 1, zl
      ax(z) = (rho_N(z,ya,xa) + rho_N(z,ya+1,xa) + rho_N(z,ya,xa+1) + rho_N(z,ya+1,xa+1))/4.0d0
    end do
    zn = minloc(r, 1)
    tr = sqrt(Rad_c**3*rho0/sigma)
    if (tt == 0) then
      open(newunit=uo, file=trim(outdir)//'/NECK.dat', status='replace')
      write(uo, '(a)') 'VARIABLES=T,T*,RMIN,RMAX,ZMIN,AMP,PHI_AXIS_MIN'
    else
      open(newunit=uo, file=trim(outdir)//'/NECK.dat', status='unknown', position='append')
    end if
    write(uo, '(i8, 1x, es24.16e3, 2(1x, es24.16e3), 1x, i6, 2(1x, es24.16e3))') tt, tt/tr, r(zn), maxval(r), zn, &
          (maxval(r) - r(zn))/2.0d0, minval(ax)
    close(uo)
    if (t_breakup < 0 .and. minval(ax) < 0.0d0) then
      t_breakup = tt
      write(*, '(a, i0)') ' breakup at step ', tt
    end if
  end subroutine output_neck

  ! output_result2 (S:1420-1441): one history file P_<p>.dat per particle,
  ! the reference's columns and format; appended at every output step (the
  ! reference rewrites the file from its start)
  subroutine output_result2(tt)
    integer, intent(in) :: tt
    integer :: k, uo
    character(len=50) :: filename
    double precision :: tr
    tr = sqrt(Rad_c**3*rho0/sigma)                   ! S:92
    do k = 1, npar
      write(filename, *) k
      if (tt == 0) then
        open(newunit=uo, file=trim(outdir)//'/P_'//trim(adjustl(filename))//'.dat', status='replace')
        write(uo, '(a)') ' VARIABLES=t,t*,cx,cy,cz,upx,upy,upz,rpx,rpy,rpz,Fpx,Fpy,Fpz,Tpx,Tpy,Tpz,Fcx,Fcy,Fcz,'// &
                         'Tcx,Tcy,Tcz,Frex,Frey,Frez'
      else
        open(newunit=uo, file=trim(outdir)//'/P_'//trim(adjustl(filename))//'.dat', status='unknown', &
             position='append')
      end if
      write(uo, "(1(1xI6),25(1xF12.6))") tt, tt/tr, center(k,:), up(k,:), rp(k,:), fp(k,:), tp(k,:), fcap(k,:), &
            tcap(k,:), frep(k,:)
      close(uo)
    end do
  end subroutine output_result2

  ! output3 (S:1673-1715): thinnest fluid thread on the plane x = xl - 1 of
  ! the reference's box -- per y, the last sign change of rho_N along z
  ! between live fluid nodes, h = (zl - 1/2) - that height -- with its y;
  ! appended to hmin.dat
  subroutine output3(tt)
    integer, intent(in) :: tt
    integer :: x, y, z, zn, y0, uo
    double precision :: bottom(yl), h, hmin, tr, lamda_d
    tr = sqrt(Rad_c**3*rho0/sigma)
    lamda_d = (dble(yl) - 2.0d0)*2.0d0                ! S:62
    x = xl - 1
    hmin = dble(zl)
    y0 = 0
    bottom = 1.0d0
    do y = 2, yl - 1
      do z = 2, zl - 1
        zn = z + 1
        if (pout(z,y,x) == 0 .and. walls_all(z,y,x) == 0 .and. pout(zn,y,x) == 0 .and. walls_all(zn,y,x) == 0) then
          if (rho_N(z,y,x)*rho_N(zn,y,x) < 0.0d0) bottom(y) = dble(zn) - rho_N(zn,y,x)/(rho_N(zn,y,x) - rho_N(z,y,x))
        end if
      end do
    end do
    do y = 2, yl - 1
      h = (dble(zl) - 0.5d0) - bottom(y)
      if (h < hmin) then
        hmin = h
        y0 = y
      end if
    end do
    if (tt == 0) then
      open(newunit=uo, file=trim(outdir)//'/hmin.dat', status='replace')
      write(uo, '(a)') ' VARIABLES=t,t*,h,y0'
    else
      open(newunit=uo, file=trim(outdir)//'/hmin.dat', status='unknown', position='append')
    end if
    write(uo, "(1(1xI8),3(1xF12.6))") tt, tt/tr, hmin/Rad_c, y0/lamda_d
    close(uo)
  end subroutine output3

  ! output4 (S:1719-1797): energy budget appended to Energy.dat -- surface
  ! energy (sigma / 2 per live node with a live neighbour of the other
  ! colour), pressure term and kinetic energy of the red fluid, particle
  ! kinetic energy, capillary work accumulated over the output steps, and
  ! the wetted area Shb, each scaled by 8 / (pi Rad_c^2 sigma) (Shb by
  ! 8 / (pi Rad_c^2)), particles set up on a symmetry plane counted half.
  ! Sums over x, y, z = 2 .. n-1.  The reference reads walls_all, rho and uz
  ! at x index z (walls_all(z,y,z), rho(z,y,z), uz(z,y,z), S:1745-1750); so
  ! does this where z <= xl, which covers the reference's cube, and at the
  ! node's own x where z > xl.  A restart starts the work sum from zero.
  subroutine output4(tt)
    integer, intent(in) :: tt
    integer, parameter :: cx(18) = [1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0]
    integer, parameter :: cy(18) = [0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1]
    integer, parameter :: cz(18) = [0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1]
    integer :: i, x, y, z, xq, k, uo
    double precision :: tr, er, epv, emov, esur, area, emov_p, wcap, emov_psum, wcap_sum, s_emov, s_wcap, s_shb
    double precision, allocatable :: shb(:)
    tr = sqrt(Rad_c**3*rho0/sigma)
    er = PI*Rad_c**2*sigma
    epv = 0.0d0
    emov = 0.0d0
    area = 0.0d0
    do x = 2, xl - 1
      do y = 2, yl - 1
        do z = 2, zl - 1
          xq = z
          if (xq > xl) xq = x
          if (walls_all(z,y,xq) <= 0) then
            if (rho_N(z,y,x) > 0.0d0) then
              epv = epv + (rho(z,y,x) - rho0)/3.0d0
              emov = emov + 0.5d0*rho(z,y,xq)*(ux(z,y,x)**2 + uy(z,y,x)**2 + uz(z,y,xq)**2)
            end if
            do i = 1, 18
              if (rho_N(z,y,x)*rho_N(z+cz(i),y+cy(i),x+cx(i)) < 0.0d0 .and. &
                  walls_all(z+cz(i),y+cy(i),x+cx(i)) <= 0) then
                area = area + 1.0d0
                exit
              end if
            end do
          end if
        end do
      end do
    end do
    epv = epv/er*8.0d0
    emov = emov/er*8.0d0
    esur = area/2.0d0*sigma/er*8.0d0
    s_emov = 0.0d0
    s_wcap = 0.0d0
    s_shb = 0.0d0
    if (npar > 0) then
      allocate(shb(npar))
      call cgl_check(cgl_get_particle_shb(h, shb), 'cgl_get_particle_shb')
      do k = 1, npar
        emov_p = 0.5d0*pmass(k)*(up(k,1)**2 + up(k,2)**2 + up(k,3)**2) + &
                 0.5d0*pinert(k)*(rp(k,1)**2 + rp(k,2)**2 + rp(k,3)**2)
        wcap = fcap(k,1)*up(k,1) + fcap(k,2)*up(k,2) + fcap(k,3)*up(k,3) + &
               tcap(k,1)*rp(k,1) + tcap(k,2)*rp(k,2) + tcap(k,3)*rp(k,3)
        if (abs(py(k) - (dble(yl) - 0.5d0)) < 0.1 .or. abs(py(k) - 1.5d0) < 0.1) then
          emov_p = emov_p*0.5d0
          wcap = wcap*0.5d0
          shb(k) = shb(k)*0.5d0
        end if
        if (abs(px(k) - (dble(xl) - 0.5d0)) < 0.1) then
          emov_p = emov_p*0.5d0
          wcap = wcap*0.5d0
          shb(k) = shb(k)*0.5d0
        end if
        if (abs(pz(k) - (dble(zl) - 0.5d0)) < 0.1) then
          emov_p = emov_p*0.5d0
          wcap = wcap*0.5d0
          shb(k) = shb(k)*0.5d0
        end if
        s_emov = s_emov + emov_p
        s_wcap = s_wcap + wcap
      end do
      do k = 1, npar
        s_shb = s_shb + shb(k)/(PI*Rad_c*Rad_c)*8.0d0
      end do
    end if
    emov_psum = s_emov/er*8.0d0
    wcap_sum = wcap_old + s_wcap/er*8.0d0
    wcap_old = wcap_sum
    if (tt == 0) then
      open(newunit=uo, file=trim(outdir)//'/Energy.dat', status='replace')
      write(uo, '(a)') ' VARIABLES=t,t*,Esur,Shb,Efree,Emov,Wcap,Emov_p,Esum'
    else
      open(newunit=uo, file=trim(outdir)//'/Energy.dat', status='unknown', position='append')
    end if
    write(uo, "(1(1xI8),8(1xF12.6))") tt, tt/tr, esur, s_shb, esur - epv, emov, wcap_sum, emov_psum, &
          (esur - epv + emov + emov_psum)
    close(uo)
  end subroutine output4

  ! output_results (S:898-924): Tecplot POINT zone, x fastest
  subroutine output_results(tt)
    integer, intent(in) :: tt
    integer :: x, y, z, uo
    character(len=50) :: filename
    write(filename, *) tt
    open(newunit=uo, file=trim(outdir)//'/OUT_'//trim(adjustl(filename))//'.dat', status='unknown')
    write(uo,*) 'TITLE="Contact_Angles"'
    write(uo,*) 'VARIABLES=X,Y,Z,WALL,RHO_N,U,V,W,RHO'
    write(uo,*) 'ZONE F=POINT I=', xl, 'J=', yl, 'K=', zl
    do z = 1, zl
      do y = 1, yl
        do x = 1, xl
          write(uo, "(4(1xI6),5(1xF12.8))") x, y, z, walls_all(z,y,x), rho_N(z,y,x), ux(z,y,x), uy(z,y,x), &
                uz(z,y,x), rho(z,y,x)
        end do
      end do
    end do
    close(uo)
  end subroutine output_results

  ! raw FP64 fields (rho, rho_R, rho_B, ux, uy, uz, rho_N), stream access
  subroutine dump_fields(tt)
    integer, intent(in) :: tt
    integer :: uo
    character(len=50) :: filename
    write(filename, *) tt
    open(newunit=uo, file=trim(outdir)//'/FIELDS_'//trim(adjustl(filename))//'.bin', access='stream', &
         form='unformatted', status='replace')
    write(uo) rho, rho_R, rho_B, ux, uy, uz, rho_N
    close(uo)
  end subroutine dump_fields
end program cgl_driver
