! box_driver.f90 -- Fortran host driving the HIP Maxwell step through the
! iso_c_binding module, on NekCEM's tests/3dboxper case: 4x4x8 periodic box
! on [0,2pi]^3, N=8, eps = mu = 1, upwind flux, dt = 2e-4, 50 steps, the
! (1,1,1) mode as initial condition, and the userchk rule of
! tests/3dboxper/3dboxper.usr:169-216 (istep <= 10 or mod(istep,50) == 0,
! L2 <= 5e-10 and Linf <= 5e-9 per field, else exit status 1).
!
! The arrays are built in the reference's COMMON layout (hn(lpts1,3) with
! lpts1 = lx1**3 * lelt above npts, as lelt = lelg + 3 in the case's SIZE
! makes it; face arrays (lx1*lz1*6*nelt) in Ed's face order, 1-based cemface
! and partner map) for straight-sided box elements; every array crosses the
! C-ABI at the COMMON stride (lpts1 or lxzfl1), and the padding rows beyond
! npts / nxzfl must come back untouched.  Straight-sided box elements have
! cofactor metrics rx = (hy/2)(hz/2), sy = (hx/2)(hz/2), tz = (hx/2)(hy/2),
! jac = hx hy hz/8, bm1 = jac w_i w_j w_k, area = face Jacobian x w_a w_b
! (src/nek5_coef.F:555-635, 1150-1237).
program box_driver
  use iso_c_binding
  use nekcem_hip
  implicit none
  integer, parameter :: N = 8, n1 = N + 1, nelx = 4, nely = 4, nelz = 8
  integer, parameter :: nelt = nelx*nely*nelz, n2 = n1*n1, n3 = n2*n1
  integer, parameter :: npts = n3*nelt, nxzfl = 6*n2*nelt
  ! COMMON capacities: lelt = lelg + 3 (tests/3dboxper/SIZE:15)
  integer, parameter :: lelt = nelt + 3, lpts1 = n3*lelt, lxzfl1 = 6*n2*lelt
  real(c_double), parameter :: pad = -7.25d0
  real(c_double), parameter :: pi = 3.141592653589793238462643383279502884d0
  real(c_double), target :: z(n1), w(n1), dxm1(n1, n1)
  real(c_double), allocatable, target :: xm1(:), ym1(:), zm1(:)
  real(c_double), allocatable, target :: rx(:), ry(:), rz(:), sx(:), sy(:), sz(:)
  real(c_double), allocatable, target :: tx(:), ty(:), tz(:), jac(:), bm1(:)
  real(c_double), allocatable, target :: unx(:), uny(:), unz(:), area(:)
  real(c_double), allocatable, target :: eps(:), mu(:), z0(:), z1(:), y0(:), y1(:)
  integer(c_int32_t), allocatable, target :: cemface(:), partner(:)
  real(c_double), allocatable :: hn(:,:), en(:,:), sh(:,:), se(:,:)
  type(nkc_desc) :: desc
  type(c_ptr) :: h
  real(c_double) :: hx, hy, hz, time, dt, l2, linf
  integer :: e, ex, ey, ez, i, j, k, f, q, a, b, istep, c, ne, nb(3), p
  integer :: vidx, axis, side, ef, opp
  integer, parameter :: eface_axis(6) = [2, 1, 2, 1, 3, 3]   ! Ed order: y-,x+,y+,x-,z-,z+
  integer, parameter :: eface_side(6) = [0, 1, 1, 0, 0, 1]

  call gll(n1, z, w)
  call dgll(n1, z, dxm1)
  hx = 2*pi/nelx; hy = 2*pi/nely; hz = 2*pi/nelz

  allocate(xm1(npts), ym1(npts), zm1(npts))
  allocate(rx(npts), ry(npts), rz(npts), sx(npts), sy(npts), sz(npts))
  allocate(tx(npts), ty(npts), tz(npts), jac(npts), bm1(npts))
  allocate(unx(nxzfl), uny(nxzfl), unz(nxzfl), area(nxzfl), cemface(nxzfl), partner(nxzfl))
  allocate(eps(npts), mu(npts), z0(nxzfl), z1(nxzfl), y0(nxzfl), y1(nxzfl))
  allocate(hn(lpts1,3), en(lpts1,3), sh(npts,3), se(npts,3))
  hn = pad; en = pad

  do e = 1, nelt
     ex = mod(e-1, nelx); ey = mod((e-1)/nelx, nely); ez = (e-1)/(nelx*nely)
     do k = 1, n1
        do j = 1, n1
           do i = 1, n1
              p = i + n1*(j-1) + n2*(k-1) + n3*(e-1)
              xm1(p) = hx*ex + 0.5d0*(z(i)+1)*hx
              ym1(p) = hy*ey + 0.5d0*(z(j)+1)*hy
              zm1(p) = hz*ez + 0.5d0*(z(k)+1)*hz
              rx(p) = 0.25d0*hy*hz; sy(p) = 0.25d0*hx*hz; tz(p) = 0.25d0*hx*hy
              ry(p) = 0; rz(p) = 0; sx(p) = 0; sz(p) = 0; tx(p) = 0; ty(p) = 0
              jac(p) = hx*hy*hz/8
              bm1(p) = jac(p)*w(i)*w(j)*w(k)
           end do
        end do
     end do
     ! faces in Ed's order, in-face a fast, b slow (skpdat)
     do ef = 1, 6
        axis = eface_axis(ef); side = eface_side(ef)
        nb = [ex, ey, ez]
        nb(axis) = nb(axis) + merge(1, -1, side == 1)
        nb(1) = modulo(nb(1), nelx); nb(2) = modulo(nb(2), nely); nb(3) = modulo(nb(3), nelz)
        ne = 1 + nb(1) + nelx*(nb(2) + nely*nb(3))
        do opp = 1, 6
           if (eface_axis(opp) == axis .and. eface_side(opp) /= side) exit
        end do
        do b = 1, n1
           do a = 1, n1
              q = a + n1*(b-1)
              p = q + n2*(ef-1) + 6*n2*(e-1)
              select case (axis)
              case (1)
                 i = merge(n1, 1, side == 1); j = a; k = b
                 area(p) = 0.25d0*hy*hz*w(a)*w(b)
              case (2)
                 i = a; j = merge(n1, 1, side == 1); k = b
                 area(p) = 0.25d0*hx*hz*w(a)*w(b)
              case default
                 i = a; j = b; k = merge(n1, 1, side == 1)
                 area(p) = 0.25d0*hx*hy*w(a)*w(b)
              end select
              vidx = i + n1*(j-1) + n2*(k-1) + n3*(e-1)
              cemface(p) = vidx
              unx(p) = 0; uny(p) = 0; unz(p) = 0
              if (axis == 1) unx(p) = merge(1d0, -1d0, side == 1)
              if (axis == 2) uny(p) = merge(1d0, -1d0, side == 1)
              if (axis == 3) unz(p) = merge(1d0, -1d0, side == 1)
              partner(p) = q + n2*(opp-1) + 6*n2*(ne-1)
           end do
        end do
     end do
  end do
  ! vacuum: Z_0 = (Z- + Z+)/2 = 1, Z_1 = Z+ = 1 (cem_maxwell_materials)
  eps = 1; mu = 1; z0 = 1; z1 = 1; y0 = 1; y1 = 1

  desc%N = N; desc%ifupwind = 1; desc%nelt = nelt
  desc%dxm1 = c_loc(dxm1); desc%wxm1 = c_loc(w)
  desc%rxmn = c_loc(rx); desc%rymn = c_loc(ry); desc%rzmn = c_loc(rz)
  desc%sxmn = c_loc(sx); desc%symn = c_loc(sy); desc%szmn = c_loc(sz)
  desc%txmn = c_loc(tx); desc%tymn = c_loc(ty); desc%tzmn = c_loc(tz)
  desc%jacm = c_loc(jac); desc%bmn = c_loc(bm1)
  desc%unxm = c_loc(unx); desc%unym = c_loc(uny); desc%unzm = c_loc(unz); desc%aream = c_loc(area)
  desc%cemface = c_loc(cemface); desc%face_partner = c_loc(partner); desc%ncempec = 0
  desc%permittivity = c_loc(eps); desc%permeability = c_loc(mu)
  desc%Z_0 = c_loc(z0); desc%Z_1 = c_loc(z1); desc%Y_0 = c_loc(y0); desc%Y_1 = c_loc(y1)
  call nkc_check(nkc_create(desc, h), 'nkc_create')

  time = 0
  dt = 2d-4
  call usersol(time, sh, se)                      ! userini
  hn(1:npts,:) = sh; en(1:npts,:) = se
  call nkc_check(nkc_set_state(h, hn, en, c_null_ptr, c_null_ptr, int(lpts1, c_int64_t)), 'nkc_set_state')
  do istep = 1, 50
     call nkc_check(nkc_step(h, time, dt, 1_c_int64_t), 'nkc_step')
     if (istep <= 10 .or. mod(istep, 50) == 0) then          ! userchk
        call nkc_check(nkc_get_state(h, hn, en, c_null_ptr, c_null_ptr, int(lpts1, c_int64_t)), &
             'nkc_get_state')
        call check_pad(hn, 'hn'); call check_pad(en, 'en')
        call usersol(time, sh, se)
        do c = 1, 3
           call cem_error(hn(:,c), sh(:,c), l2, linf)
           if (l2 > 5d-10 .or. linf > 5d-9) call fail(istep, c, l2, linf)
           call cem_error(en(:,c), se(:,c), l2, linf)
           if (l2 > 5d-10 .or. linf > 5d-9) call fail(istep, 3+c, l2, linf)
        end do
        write(6, '(i6,es12.4,a)') istep, time, '  userchk ok'
     end if
  end do
  ! The PML entry points from Fortran: a layer with sigma = 0 on the first
  ! two elements changes no field (pml_step adds (sa - sb) P - sd ... = 0),
  ! while B and D integrate the residual there; 10 more steps still pass the
  ! userchk, and the PML state comes back finite and nonzero.
  block
    integer(c_int32_t), target :: pmlptr(2)
    real(c_double), allocatable :: sigma(:,:), bd(:,:)
    allocate(sigma(lpts1,3), bd(lpts1,6))
    sigma = 0; sigma(npts+1:,:) = pad; bd = pad
    pmlptr = [1, 2]
    call nkc_check(nkc_set_pml(h, 2_c_int64_t, pmlptr, sigma, c_null_ptr, c_null_ptr, &
         int(lpts1, c_int64_t)), 'nkc_set_pml')
    call nkc_check(nkc_step(h, time, dt, 10_c_int64_t), 'nkc_step (PML)')
    call nkc_check(nkc_get_state(h, hn, en, c_null_ptr, c_null_ptr, int(lpts1, c_int64_t)), 'nkc_get_state')
    call check_pad(hn, 'hn'); call check_pad(en, 'en')
    call usersol(time, sh, se)
    do c = 1, 3
       call cem_error(hn(:,c), sh(:,c), l2, linf)
       if (l2 > 5d-10 .or. linf > 5d-9) call fail(60, c, l2, linf)
       call cem_error(en(:,c), se(:,c), l2, linf)
       if (l2 > 5d-10 .or. linf > 5d-9) call fail(60, 3+c, l2, linf)
    end do
    call nkc_check(nkc_get_pml(h, 0_c_int32_t, bd, int(lpts1, c_int64_t)), 'nkc_get_pml')
    if (.not. (maxval(abs(bd(1:2*n3, :))) > 0 .and. maxval(abs(bd(1:2*n3, :))) < 1d10)) then
       write(6, '(a,es12.4)') 'box_driver: PML state not as expected, max |B,D| =', maxval(abs(bd(1:2*n3, :)))
       error stop 1
    end if
    ! nodes outside the PML elements and the padding rows are not written
    if (any(bd(2*n3+1:, :) /= pad)) then
       write(6, '(a)') 'box_driver: PML state written outside the PML elements'
       error stop 1
    end if
    write(6, '(a,es12.4)') 'box_driver: PML entry points ok, max |B,D| =', maxval(abs(bd(1:2*n3, :)))
  end block
  ! The graphene-sheet entry points from Fortran, with the arrays of the
  ! .usr's /usergraph/ block: a Drude-only sheet (b_d = 1.5) on the x+ face
  ! of element 1 starting from zero; after 5 steps its Drude slot (fjn(:,:,2))
  ! has charged up, the critical-point slots 3..6 are still zero, and
  ! ngraph = 0 removes the sheet.
  block
    integer(c_int32_t), allocatable :: graphindex(:)
    real(c_double), allocatable :: graphparams(:,:), fjn(:,:,:)
    integer :: i
    allocate(graphindex(n2), graphparams(lxzfl1,12), fjn(lxzfl1,3,6))
    graphparams = 0; fjn = 0
    graphparams(nxzfl+1:,:) = pad; fjn(nxzfl+1:,:,:) = pad
    do i = 1, n2
       graphindex(i) = n2*(2-1) + i          ! element 1, face 2 (x+)
       graphparams(graphindex(i), 2) = 1.5d0
    end do
    call nkc_check(nkc_set_graphene(h, int(n2, c_int64_t), graphindex, graphparams, fjn, &
         int(lxzfl1, c_int64_t)), 'nkc_set_graphene')
    call nkc_check(nkc_step(h, time, dt, 5_c_int64_t), 'nkc_step (graphene)')
    call nkc_check(nkc_get_graphene(h, 0_c_int32_t, fjn, int(lxzfl1, c_int64_t)), 'nkc_get_graphene')
    if (.not. (maxval(abs(fjn(graphindex, :, 2))) > 0 .and. maxval(abs(fjn(graphindex, :, 2))) < 1d10) &
         .or. maxval(abs(fjn(1:nxzfl, :, 3:6))) /= 0 .or. any(fjn(nxzfl+1:, :, :) /= pad)) then
       write(6, '(a)') 'box_driver: graphene state not as expected'
       error stop 1
    end if
    call nkc_check(nkc_set_graphene(h, 0_c_int64_t, graphindex, graphparams, fjn, int(lxzfl1, c_int64_t)), &
         'nkc_set_graphene (remove)')
    write(6, '(a,es12.4)') 'box_driver: graphene entry points ok, max |fjn(:,:,2)| =', &
         maxval(abs(fjn(graphindex, :, 2)))
  end block
  call nkc_check(nkc_destroy(h), 'nkc_destroy')
  write(6, '(a)') 'box_driver: 3dboxper passed on the HIP path'

contains

  ! the rows npts+1..lpts1 of a COMMON field array hold the pad value still
  subroutine check_pad(u, what)
    real(c_double), intent(in) :: u(:,:)
    character(len=*), intent(in) :: what
    if (any(u(npts+1:, :) /= pad)) then
       write(6, '(a,a)') 'box_driver: padding rows written in ', what
       error stop 1
    end if
  end subroutine check_pad

  subroutine fail(istep, c, l2, linf)
    integer, intent(in) :: istep, c
    real(c_double), intent(in) :: l2, linf
    write(6, '(a,i0,a,i0,2es12.4)') 'box_driver: userchk failed at istep ', istep, ' field ', c, l2, linf
    error stop 1
  end subroutine fail

  ! usersol of tests/3dboxper/3dboxper.usr:45-86
  subroutine usersol(tt, hs, es)
    real(c_double), intent(in) :: tt
    real(c_double), intent(out) :: hs(npts,3), es(npts,3)
    real(c_double) :: om, th, te
    integer :: ii
    om = sqrt(3d0); th = sin(om*tt)/om; te = cos(om*tt)
    do ii = 1, npts
       hs(ii,1) = 2*cos(xm1(ii))*sin(ym1(ii))*cos(zm1(ii))*th
       hs(ii,2) = -sin(xm1(ii))*cos(ym1(ii))*cos(zm1(ii))*th
       hs(ii,3) = sin(xm1(ii))*sin(ym1(ii))*sin(zm1(ii))*th
       es(ii,1) = 0
       es(ii,2) = cos(xm1(ii))*sin(ym1(ii))*sin(zm1(ii))*te
       es(ii,3) = cos(xm1(ii))*cos(ym1(ii))*cos(zm1(ii))*te
    end do
  end subroutine usersol

  ! cem_error (src/cem_common.F:1335-1355)
  subroutine cem_error(u, ex_, l2, linf)
    real(c_double), intent(in) :: u(npts), ex_(npts)
    real(c_double), intent(out) :: l2, linf
    l2 = sum((ex_ - u)**2*bm1)/sum(bm1)
    if (l2 > 0) l2 = sqrt(l2)
    linf = maxval(abs(ex_ - u))
  end subroutine cem_error

  ! GLL nodes (roots of (1-x^2) P_N') by Newton with deflation, weights
  ! 2/(N(N+1) P_N^2)
  subroutine gll(m, zz, ww)
    integer, intent(in) :: m
    real(c_double), intent(out) :: zz(m), ww(m)
    real(c_double) :: x, p, dp, d2p, dx, defl
    integer :: ii, it, jj, nn
    nn = m - 1
    zz(1) = -1; zz(m) = 1
    do ii = 2, m-1
       x = -cos(pi*(ii-1)/nn)
       do it = 1, 100
          call legendre(x, nn, p, dp)
          d2p = (2*x*dp - nn*(nn+1)*p)/(1 - x*x)
          defl = 0
          do jj = 2, ii-1
             defl = defl + 1/(x - zz(jj))
          end do
          dx = dp/(d2p - dp*defl)
          x = x - dx
          if (abs(dx) < 1d-17) exit
       end do
       zz(ii) = x
    end do
    do ii = 1, m
       call legendre(zz(ii), nn, p, dp)
       ww(ii) = 2/(nn*(nn+1d0)*p*p)
    end do
  end subroutine gll

  subroutine legendre(x, nn, p, dp)
    real(c_double), intent(in) :: x
    integer, intent(in) :: nn
    real(c_double), intent(out) :: p, dp
    real(c_double) :: p0, p1, p2, d0, d1, d2
    integer :: kk
    p0 = 1; p1 = x; d0 = 0; d1 = 1
    do kk = 1, nn-1
       p2 = ((2*kk+1)*x*p1 - kk*p0)/(kk+1)
       d2 = d0 + (2*kk+1)*p1
       p0 = p1; p1 = p2; d0 = d1; d1 = d2
    end do
    p = p1; dp = d1
  end subroutine legendre

  ! DGLL (src/nek5_speclib.F:807-838), D(i,j) = dl_j/dr at z_i
  subroutine dgll(m, zz, dd)
    integer, intent(in) :: m
    real(c_double), intent(in) :: zz(m)
    real(c_double), intent(out) :: dd(m,m)
    real(c_double) :: pi_, pj, dpi, dpj
    integer :: ii, jj, nn
    nn = m - 1
    do jj = 1, m
       do ii = 1, m
          dd(ii,jj) = 0
          if (ii /= jj) then
             call legendre(zz(ii), nn, pi_, dpi)
             call legendre(zz(jj), nn, pj, dpj)
             dd(ii,jj) = pi_/(pj*(zz(ii) - zz(jj)))
          end if
       end do
    end do
    dd(1,1) = -nn*(nn+1)/4d0
    dd(m,m) = nn*(nn+1)/4d0
  end subroutine dgll

end program box_driver
