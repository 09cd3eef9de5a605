! ref_harness.f90 -- TEST INFRASTRUCTURE ONLY.
!
! Drives the unmodified reference (Source1.f90 of X-Sheen/Symmetry, compiled
! from where it lies by oracle/Makefile into oracle/_ref/) one subroutine at a
! time, so that the FP64 oracle (oracle/cgl_oracle.c) can be pinned against
! the reference stage by stage.  It owns the reference's state arrays the way
! `program main` does (S:144-179), hands their addresses to the caller, and
! calls the reference's external subroutines with them.  The reference's
! compile-time box (xl=yl=zl=152, par=12, S:9-24) is used as is.
module ref_harness
  use iso_c_binding
  use const
  implicit none
  integer, allocatable, target, dimension(:,:,:) :: walls, walls_old, walls_all, pin, pout
  integer, allocatable, target, dimension(:,:) :: next_x, next_y, next_z
  double precision, allocatable, target, dimension(:,:,:) :: rho, ux, uy, uz, rho_N, rho_R, rho_B
  double precision, allocatable, target, dimension(:,:,:) :: n_x, n_y, n_z, gradx, grady, gradz
  double precision, allocatable, target, dimension(:,:,:,:) :: fi, ri, bi, ffi
  logical, allocatable, target, dimension(:,:,:) :: flag
  double precision, allocatable, target, dimension(:,:) :: Lax, Lay, Laz, La_rhoN, La_nx, La_ny, La_nz, ds
  logical, allocatable, target, dimension(:,:) :: La_flag
  double precision, target :: center(par,3), center_old(par,3), center0(par,3)
  double precision, target :: Fpx(par), Fpy(par), Fpz(par), Tpx(par), Tpy(par), Tpz(par)
  double precision, target :: Frex(par), Frey(par), Frez(par)
  double precision, target :: tension(par,3), tension_t(par,3)
  double precision, target :: upx(par), upy(par), upz(par), rpx(par), rpy(par), rpz(par)
  double precision, target :: Shb(par)
  double precision, target :: Mmat(0:18,0:18), Minv(0:18,0:18)
  double precision, target :: scal(8)   ! Fg, mass, inertia, lamda_d, Wcap_old, tr, t
contains

  subroutine ref_alloc() bind(C, name="ref_alloc")
    if (.not. allocated(rho)) then
      allocate(next_x(18,xl), next_y(18,yl), next_z(18,zl))
      allocate(walls(zl,yl,xl), walls_all(zl,yl,xl), walls_old(zl,yl,xl), pin(zl,yl,xl), pout(zl,yl,xl))
      allocate(rho(zl,yl,xl), rho_R(zl,yl,xl), rho_B(zl,yl,xl), rho_N(zl,yl,xl))
      allocate(ux(zl,yl,xl), uy(zl,yl,xl), uz(zl,yl,xl), flag(zl,yl,xl))
      allocate(gradx(zl,yl,xl), grady(zl,yl,xl), gradz(zl,yl,xl))
      allocate(n_x(zl,yl,xl), n_y(zl,yl,xl), n_z(zl,yl,xl))
      allocate(fi(0:18,zl,yl,xl), ffi(0:18,zl,yl,xl), ri(0:18,zl,yl,xl), bi(0:18,zl,yl,xl))
      allocate(Lax(par,La_num), Lay(par,La_num), Laz(par,La_num), La_rhoN(par,La_num))
      allocate(La_nx(par,La_num), La_ny(par,La_num), La_nz(par,La_num), La_flag(par,La_num), ds(par,La_num))
    end if
    walls = 0; walls_all = 0; walls_old = 0; pin = 0; pout = 0
    rho = 0.0d0; rho_R = 0.0d0; rho_B = 0.0d0; rho_N = 0.0d0
    ux = 0.0d0; uy = 0.0d0; uz = 0.0d0; flag = .false.
    gradx = 0.0d0; grady = 0.0d0; gradz = 0.0d0
    n_x = 0.0d0; n_y = 0.0d0; n_z = 0.0d0
    fi = 0.0d0; ffi = 0.0d0; ri = 0.0d0; bi = 0.0d0
    center = 0.0d0; center_old = 0.0d0; center0 = 0.0d0
    Fpx = 0.0d0; Fpy = 0.0d0; Fpz = 0.0d0; Tpx = 0.0d0; Tpy = 0.0d0; Tpz = 0.0d0
    Frex = 0.0d0; Frey = 0.0d0; Frez = 0.0d0; tension = 0.0d0; tension_t = 0.0d0
    upx = 0.0d0; upy = 0.0d0; upz = 0.0d0; rpx = 0.0d0; rpy = 0.0d0; rpz = 0.0d0
    Shb = 0.0d0
  end subroutine ref_alloc

  ! Module parameters the reference's main assigns (S:57-89, S:132-142).  The
  ! D3Q19 tables are the lattice definition shared by every implementation.
  subroutine ref_setup(p_sigma, p_rho0, p_nu_r, p_nu_b, p_rad, p_rad_c, p_M) bind(C, name="ref_setup")
    real(c_double), value :: p_sigma, p_rho0, p_nu_r, p_nu_b, p_rad, p_rad_c
    real(c_double) :: p_M(0:18,0:18)
    integer :: i, j
    sigma = p_sigma; rho0 = p_rho0; nu_r = p_nu_r; nu_b = p_nu_b; Rad = p_rad; Rad_c = p_rad_c
    La_m = 2.0d0*pi/alpha
    La_n = pi/alpha
    La_num = La_m*(La_n+1)
    cix = (/0,1,-1,0,0,0,0,1,-1,1,-1,1,-1,1,-1,0,0,0,0/)
    ciy = (/0,0,0,1,-1,0,0,1,1,-1,-1,0,0,0,0,1,-1,1,-1/)
    ciz = (/0,0,0,0,0,1,-1,0,0,0,0,1,1,-1,-1,1,1,-1,-1/)
    do i = 0, 18
      do j = 0, 18
        if (cix(j) == -cix(i) .and. ciy(j) == -ciy(i) .and. ciz(j) == -ciz(i)) opc(i) = j
        if (cix(j) == -cix(i) .and. ciy(j) == ciy(i) .and. ciz(j) == ciz(i)) symx(i) = j
        if (cix(j) == cix(i) .and. ciy(j) == -ciy(i) .and. ciz(j) == ciz(i)) symy(i) = j
        if (cix(j) == cix(i) .and. ciy(j) == ciy(i) .and. ciz(j) == -ciz(i)) symz(i) = j
      end do
    end do
    wi(0) = 1.0d0/3.0d0
    wi(1:6) = 1.0d0/18.0d0
    wi(7:18) = 1.0d0/36.0d0
    Mmat = p_M
    do i = 0, 18
      do j = 0, 18
        Minv(i,j) = Mmat(j,i) / dot_product(Mmat(j,:), Mmat(j,:))
      end do
    end do
  end subroutine ref_setup

  function ref_dims(k) bind(C, name="ref_dims") result(v)
    integer(c_int), value :: k
    integer(c_int) :: v
    select case (k)
    case (0); v = xl
    case (1); v = yl
    case (2); v = zl
    case (3); v = par
    case (4); v = La_num
    case default; v = -1
    end select
  end function ref_dims

  function ref_ptr(k) bind(C, name="ref_ptr") result(p)
    integer(c_int), value :: k
    type(c_ptr) :: p
    select case (k)
    case (0); p = c_loc(fi)
    case (1); p = c_loc(ffi)
    case (2); p = c_loc(ri)
    case (3); p = c_loc(bi)
    case (4); p = c_loc(rho)
    case (5); p = c_loc(rho_R)
    case (6); p = c_loc(rho_B)
    case (7); p = c_loc(rho_N)
    case (8); p = c_loc(ux)
    case (9); p = c_loc(uy)
    case (10); p = c_loc(uz)
    case (11); p = c_loc(gradx)
    case (12); p = c_loc(grady)
    case (13); p = c_loc(gradz)
    case (14); p = c_loc(n_x)
    case (15); p = c_loc(n_y)
    case (16); p = c_loc(n_z)
    case (17); p = c_loc(flag)
    case (18); p = c_loc(walls)
    case (19); p = c_loc(walls_old)
    case (20); p = c_loc(walls_all)
    case (21); p = c_loc(pin)
    case (22); p = c_loc(pout)
    case (23); p = c_loc(center)
    case (24); p = c_loc(center_old)
    case (25); p = c_loc(center0)
    case (26); p = c_loc(Fpx)
    case (27); p = c_loc(Fpy)
    case (28); p = c_loc(Fpz)
    case (29); p = c_loc(Tpx)
    case (30); p = c_loc(Tpy)
    case (31); p = c_loc(Tpz)
    case (32); p = c_loc(Frex)
    case (33); p = c_loc(Frey)
    case (34); p = c_loc(Frez)
    case (35); p = c_loc(tension)
    case (36); p = c_loc(tension_t)
    case (37); p = c_loc(upx)
    case (38); p = c_loc(upy)
    case (39); p = c_loc(upz)
    case (40); p = c_loc(rpx)
    case (41); p = c_loc(rpy)
    case (42); p = c_loc(rpz)
    case (43); p = c_loc(Shb)
    case (44); p = c_loc(Mmat)
    case (45); p = c_loc(Minv)
    case (46); p = c_loc(scal)
    case (47); p = c_loc(next_x)
    case (48); p = c_loc(next_y)
    case (49); p = c_loc(next_z)
    case default; p = c_null_ptr
    end select
  end function ref_ptr

  ! One reference subroutine per id, with the argument lists of the calls in
  ! the reference's main (S:208-251).  scal = (Fg, mass, inertia, lamda_d, Wcap_old).
  subroutine ref_call(k) bind(C, name="ref_call")
    integer(c_int), value :: k
    select case (k)
    case (1); call calmacros(fi,ri,bi,rho,ux,uy,uz,rho_R,rho_B,rho_N,walls_all)
    case (2); call derivatives(walls_all,pin,pout,center,next_x,next_y,next_z,rho_N,gradx,grady,gradz,n_x,n_y,n_z,flag)
    case (3); call collision(walls_all,pout,next_x,next_y,next_z,flag,n_x,n_y,n_z,gradx,grady,gradz,rho,rho_N, &
                             ux,uy,uz,fi,ffi,Mmat,Minv)
    case (4); call Recolor(flag,fi,ri,bi,rho_R,rho_B,rho,n_x,n_y,n_z,walls_all)
    case (5); call boundary(next_x,next_y,next_z,ri,bi)
    case (6); call streaming(ri, next_x, next_y, next_z)
    case (7); call streaming(bi, next_x, next_y, next_z)
    case (8); call curve(walls,pout,ri,bi,upx,upy,upz,rpx,rpy,rpz,center,rho_R,rho_B)
    case (9); call set_walls(walls,walls_old,walls_all,pin,pout,center,next_x,next_y,next_z)
    case (10); call refilling(next_x,next_y,next_z,pout,walls_old,center_old,upx,upy,upz,rpx,rpy,rpz,rho,rho_N, &
                              rho_R,rho_B,ux,uy,uz,ri,bi,fi)
    case (11); call momentum(walls,pout,center,Fpx,Fpy,Fpz,Tpx,Tpy,Tpz,upx,upy,upz,rpx,rpy,rpz,ffi,fi)
    case (12); call Lagrange_points(walls,center,rho_N,gradx,grady,gradz,n_x,n_y,n_z,flag,Lax,Lay,Laz,La_rhoN, &
                                    La_nx,La_ny,La_nz,La_flag,ds,Shb)
    case (13); call capillary_force2(center,Lax,Lay,Laz,La_nx,La_ny,La_nz,La_rhoN,La_flag,ds,tension,tension_t)
    case (14); call repulsive(center,Frex,Frey,Frez,scal(1))
    case (15); call movement(center,center_old,center0,Fpx,Fpy,Fpz,Frex,Frey,Frez,Tpx,Tpy,Tpz,upx,upy,upz, &
                             rpx,rpy,rpz,scal(1),scal(2),scal(3),tension,tension_t)
    case (16); call initialization(scal(4),rho,rho_N,rho_R,rho_B,ux,uy,uz,fi,ffi,ri,bi)
    case (17); call initia_center(center,center0,scal(4))
    case (18); call tables(next_x, next_y, next_z)
    ! output4 with t = scal(7), tr = scal(6), mass = scal(2), inertia =
    ! scal(3), Wcap_old = scal(5) (updated); unit 400 closed afterwards so
    ! that Energy.dat holds this call's line
    case (19)
      call output4(rho_N,int(scal(7)),scal(6),walls_all,pout,center0,ux,uy,uz,rho,upx,upy,upz,rpx,rpy,rpz, &
                   tension,tension_t,scal(2),scal(3),scal(5),Shb)
      close(400)
    ! output3 with t = scal(7), tr = scal(6), lamda_d = scal(4); unit 300
    ! closed afterwards so that hmin.dat holds this call's line
    case (20)
      call output3(rho_N,int(scal(7)),scal(6),walls_all,pout,scal(4))
      close(300)
    end select
  end subroutine ref_call

  ! The loop body S:216-251 without its output block; particles toggles the
  ! particle block S:237-243 the way the reference's commented if(1==2) does.
  subroutine ref_step(nsteps, particles) bind(C, name="ref_step")
    integer(c_int), value :: nsteps, particles
    integer :: t
    do t = 1, nsteps
      call ref_call(1)
      if (particles /= 0) then
        call ref_call(12); call ref_call(13); call ref_call(11); call ref_call(14)
        call ref_call(15); call ref_call(9); call ref_call(10)
      end if
      call ref_call(2); call ref_call(3); call ref_call(4); call ref_call(5)
      call ref_call(6); call ref_call(7); call ref_call(8)
    end do
  end subroutine ref_step
end module ref_harness
