! nekcem_hip_mod.f90 -- iso_c_binding interface of libnekcem_hip.so for the
! Fortran host (NekCEM's driver).  Mirrors include/nekcem_hip.h one to one.
! A NekCEM build calls these in place of cem_maxwell_op_rk
! (src/cem_maxwell.F:327-345), with the COMMON arrays of src/EMWAVE, src/GEOM,
! src/INPUT passed as they are (REAL*8, 1-based INTEGER index arrays).
module nekcem_hip
  use iso_c_binding
  implicit none
  private

  integer(c_int32_t), parameter, public :: NKC_OK = 0, NKC_ERR_ARG = 1, &
       NKC_ERR_HIP = 2, NKC_ERR_UNSUPPORTED = 3, NKC_ERR_BLOWUP = 4, &
       NKC_ERR_NODEVICE = 5

  ! struct nkc_desc
  type, bind(C), public :: nkc_desc
     integer(c_int32_t) :: N = 0
     integer(c_int32_t) :: ifupwind = 1
     integer(c_int64_t) :: nelt = 0
     type(c_ptr) :: dxm1 = c_null_ptr, wxm1 = c_null_ptr
     type(c_ptr) :: rxmn = c_null_ptr, rymn = c_null_ptr, rzmn = c_null_ptr
     type(c_ptr) :: sxmn = c_null_ptr, symn = c_null_ptr, szmn = c_null_ptr
     type(c_ptr) :: txmn = c_null_ptr, tymn = c_null_ptr, tzmn = c_null_ptr
     type(c_ptr) :: jacm = c_null_ptr, bmn = c_null_ptr
     type(c_ptr) :: unxm = c_null_ptr, unym = c_null_ptr, unzm = c_null_ptr
     type(c_ptr) :: aream = c_null_ptr
     type(c_ptr) :: cemface = c_null_ptr, face_partner = c_null_ptr
     type(c_ptr) :: cempec = c_null_ptr
     integer(c_int64_t) :: ncempec = 0
     type(c_ptr) :: permittivity = c_null_ptr, permeability = c_null_ptr
     type(c_ptr) :: Z_0 = c_null_ptr, Z_1 = c_null_ptr
     type(c_ptr) :: Y_0 = c_null_ptr, Y_1 = c_null_ptr
  end type nkc_desc

  public :: nkc_create, nkc_destroy, nkc_set_state, nkc_get_state, nkc_step, &
       nkc_residual, nkc_set_source, nkc_set_source_time, nkc_set_drude, nkc_get_drude, &
       nkc_set_lorentz, nkc_get_lorentz, nkc_set_materials, &
       nkc_set_pml, nkc_set_pml_fields, nkc_get_pml, nkc_set_incident, &
       nkc_set_graphene, nkc_get_graphene, &
       nkc_adjoint_forward, nkc_adjoint_backward, nkc_adjoint_get_drude, &
       nkc_check_finite, nkc_sync, nkc_stream, nkc_info, &
       nkc_set_graph, nkc_set_keep_k, nkc_restart_write, nkc_restart_read, &
       nkc_last_error, nkc_check, nkc_error_message

  interface
     integer(c_int32_t) function nkc_create(d, h) bind(C, name='nkc_create')
       import :: c_int32_t, c_ptr, nkc_desc
       type(nkc_desc), intent(in) :: d
       type(c_ptr), intent(out) :: h
     end function nkc_create

     integer(c_int32_t) function nkc_destroy(h) bind(C, name='nkc_destroy')
       import :: c_int32_t, c_ptr
       type(c_ptr), value :: h
     end function nkc_destroy

     ! hn, en: (ld,3); khn, ken: c_loc of (ld,3) arrays or c_null_ptr
     integer(c_int32_t) function nkc_set_state(h, hn, en, khn, ken, ld) &
          bind(C, name='nkc_set_state')
       import :: c_int32_t, c_int64_t, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double), intent(in) :: hn(*), en(*)
       type(c_ptr), value :: khn, ken
       integer(c_int64_t), value :: ld
     end function nkc_set_state

     integer(c_int32_t) function nkc_get_state(h, hn, en, khn, ken, ld) &
          bind(C, name='nkc_get_state')
       import :: c_int32_t, c_int64_t, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double), intent(out) :: hn(*), en(*)
       type(c_ptr), value :: khn, ken
       integer(c_int64_t), value :: ld
     end function nkc_get_state

     integer(c_int32_t) function nkc_step(h, time, dt, nsteps) bind(C, name='nkc_step')
       import :: c_int32_t, c_int64_t, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double), intent(inout) :: time
       real(c_double), value :: dt
       integer(c_int64_t), value :: nsteps
     end function nkc_step

     integer(c_int32_t) function nkc_residual(h, rktime, reshn, resen, ld) &
          bind(C, name='nkc_residual')
       import :: c_int32_t, c_int64_t, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double), value :: rktime
       real(c_double), intent(out) :: reshn(*), resen(*)
       integer(c_int64_t), value :: ld
     end function nkc_residual

     ! amp: c_loc of npts bm1-weighted amplitudes, or c_null_ptr to remove
     integer(c_int32_t) function nkc_set_source(h, amp, dir, t0, tw) bind(C, name='nkc_set_source')
       import :: c_int32_t, c_ptr, c_double
       type(c_ptr), value :: h, amp
       real(c_double), intent(in) :: dir(3)
       real(c_double), value :: t0, tw
     end function nkc_set_source

     ! dindex(ndrude) 1-based volume nodes; params(2,ndrude) = (gamma,
     ! omega_p^2); J, kJ: c_loc of (ndrude,3) arrays or c_null_ptr
     integer(c_int32_t) function nkc_set_drude(h, ndrude, dindex, params, J, kJ) &
          bind(C, name='nkc_set_drude')
       import :: c_int32_t, c_int64_t, c_ptr, c_double
       type(c_ptr), value :: h
       integer(c_int64_t), value :: ndrude
       integer(c_int32_t), intent(in) :: dindex(*)
       real(c_double), intent(in) :: params(*)
       type(c_ptr), value :: J, kJ
     end function nkc_set_drude

     integer(c_int32_t) function nkc_get_drude(h, J, kJ, resJ) bind(C, name='nkc_get_drude')
       import :: c_int32_t, c_ptr
       type(c_ptr), value :: h, J, kJ, resJ
     end function nkc_get_drude

     ! source time factor: kind 0 Gaussian pulse, 1 sin(omega t + phase)
     integer(c_int32_t) function nkc_set_source_time(h, kind, omega, phase) &
          bind(C, name='nkc_set_source_time')
       import :: c_int32_t, c_ptr, c_double
       type(c_ptr), value :: h
       integer(c_int32_t), value :: kind
       real(c_double), value :: omega, phase
     end function nkc_set_source_time

     ! Lorentz ADE on lindex (1-based); params (a, b, c) per node; L, kL:
     ! c_loc of [6][nlor] arrays or c_null_ptr
     integer(c_int32_t) function nkc_set_lorentz(h, nlor, lindex, params, L, kL) &
          bind(C, name='nkc_set_lorentz')
       import :: c_int32_t, c_int64_t, c_ptr, c_double
       type(c_ptr), value :: h
       integer(c_int64_t), value :: nlor
       integer(c_int32_t), intent(in) :: lindex(*)
       real(c_double), intent(in) :: params(*)
       type(c_ptr), value :: L, kL
     end function nkc_set_lorentz

     integer(c_int32_t) function nkc_get_lorentz(h, L, kL, resL) bind(C, name='nkc_get_lorentz')
       import :: c_int32_t, c_ptr
       type(c_ptr), value :: h, L, kL, resL
     end function nkc_get_lorentz

     ! PML from the reference's arrays: pmlptr(maxpml), pmlsigma(lpts1,3),
     ! pmlbn / pmldn (lpts1,3) as c_loc or c_null_ptr, ld = lpts1
     integer(c_int32_t) function nkc_set_pml(h, maxpml, pmlptr, sigma, bn, dn, ld) &
          bind(C, name='nkc_set_pml')
       import :: c_int32_t, c_int64_t, c_ptr, c_double
       type(c_ptr), value :: h
       integer(c_int64_t), value :: maxpml, ld
       integer(c_int32_t), intent(in) :: pmlptr(*)
       real(c_double), intent(in) :: sigma(*)
       type(c_ptr), value :: bn, dn
     end function nkc_set_pml

     integer(c_int32_t) function nkc_set_pml_fields(h, bn, dn, ld) bind(C, name='nkc_set_pml_fields')
       import :: c_int32_t, c_int64_t, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double), intent(in) :: bn(*), dn(*)
       integer(c_int64_t), value :: ld
     end function nkc_set_pml_fields

     ! which: 0 (pmlbn, pmldn), 1 their RK register, 2 residual, 3 sigma
     integer(c_int32_t) function nkc_get_pml(h, which, out, ld) bind(C, name='nkc_get_pml')
       import :: c_int32_t, c_int64_t, c_ptr, c_double
       type(c_ptr), value :: h
       integer(c_int32_t), value :: which
       real(c_double), intent(out) :: out(*)
       integer(c_int64_t), value :: ld
     end function nkc_get_pml

     ! incident plane wave on faces (element, face) pairs, 1-based
     integer(c_int32_t) function nkc_set_incident(h, nface, faces, k, omega, hamp, eamp) &
          bind(C, name='nkc_set_incident')
       import :: c_int32_t, c_int64_t, c_ptr, c_double
       type(c_ptr), value :: h
       integer(c_int64_t), value :: nface
       integer(c_int32_t), intent(in) :: faces(*)
       real(c_double), intent(in) :: k(3), hamp(3), eamp(3)
       real(c_double), value :: omega
     end function nkc_set_incident

     ! graphene sheet current of userfsrc on the face points graphindex
     ! (1-based face numbering), graphparams(ld,12), fjn(ld,3,6) as the .usr
     ! common block /usergraph/ holds them; ngraph = 0 removes it
     integer(c_int32_t) function nkc_set_graphene(h, ngraph, graphindex, graphparams, fjn, ld) &
          bind(C, name='nkc_set_graphene')
       import :: c_int32_t, c_int64_t, c_ptr, c_double
       type(c_ptr), value :: h
       integer(c_int64_t), value :: ngraph, ld
       integer(c_int32_t), intent(in) :: graphindex(*)
       real(c_double), intent(in) :: graphparams(*), fjn(*)
     end function nkc_set_graphene

     ! sheet state (which = 0) or RK register (1) into fjn(ld,3,2:6)
     integer(c_int32_t) function nkc_get_graphene(h, which, fjn, ld) bind(C, name='nkc_get_graphene')
       import :: c_int32_t, c_int64_t, c_ptr, c_double
       type(c_ptr), value :: h
       integer(c_int32_t), value :: which
       real(c_double), intent(inout) :: fjn(*)
       integer(c_int64_t), value :: ld
     end function nkc_get_graphene

     ! per-node permittivity and permeability, npts each
     integer(c_int32_t) function nkc_set_materials(h, eps, mu) bind(C, name='nkc_set_materials')
       import :: c_int32_t, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double), intent(in) :: eps(*), mu(*)
     end function nkc_set_materials

     ! discrete adjoint: checkpointed forward, then dL/deps, dL/dmu per
     ! node from dL/dhn, dL/den at the final state (layout of hn/en, stride
     ! ld); geps/gmu: c_loc of npts arrays or c_null_ptr
     integer(c_int32_t) function nkc_adjoint_forward(h, time, dt, nsteps, every) &
          bind(C, name='nkc_adjoint_forward')
       import :: c_int32_t, c_int64_t, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double), intent(inout) :: time
       real(c_double), value :: dt
       integer(c_int64_t), value :: nsteps, every
     end function nkc_adjoint_forward

     integer(c_int32_t) function nkc_adjoint_backward(h, lamh, lame, ld, geps, gmu) &
          bind(C, name='nkc_adjoint_backward')
       import :: c_int32_t, c_int64_t, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double), intent(in) :: lamh(*), lame(*)
       integer(c_int64_t), value :: ld
       type(c_ptr), value :: geps, gmu
     end function nkc_adjoint_backward

     ! dL/dgamma, dL/d(omega_p^2) per listed Drude node (c_loc or c_null_ptr)
     integer(c_int32_t) function nkc_adjoint_get_drude(h, ga, gb) bind(C, name='nkc_adjoint_get_drude')
       import :: c_int32_t, c_ptr
       type(c_ptr), value :: h, ga, gb
     end function nkc_adjoint_get_drude

     integer(c_int32_t) function nkc_check_finite(h, maxabs) bind(C, name='nkc_check_finite')
       import :: c_int32_t, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double), intent(out) :: maxabs
     end function nkc_check_finite

     integer(c_int32_t) function nkc_sync(h) bind(C, name='nkc_sync')
       import :: c_int32_t, c_ptr
       type(c_ptr), value :: h
     end function nkc_sync

     integer(c_int32_t) function nkc_stream(h, stream) bind(C, name='nkc_stream')
       import :: c_int32_t, c_ptr
       type(c_ptr), value :: h
       type(c_ptr), intent(out) :: stream
     end function nkc_stream

     integer(c_int32_t) function nkc_info(h, out) bind(C, name='nkc_info')
       import :: c_int32_t, c_int64_t, c_ptr
       type(c_ptr), value :: h
       integer(c_int64_t), intent(out) :: out(8)
     end function nkc_info

     integer(c_int32_t) function nkc_set_graph(h, enable) bind(C, name='nkc_set_graph')
       import :: c_int32_t, c_ptr
       type(c_ptr), value :: h
       integer(c_int32_t), value :: enable
     end function nkc_set_graph

     ! restart file of cem_restart_out / restart_swap; path is NUL-terminated
     integer(c_int32_t) function nkc_restart_write(h, path, istep, dumpno, time, dt, single) &
          bind(C, name='nkc_restart_write')
       import :: c_int32_t, c_ptr, c_double, c_char
       type(c_ptr), value :: h
       character(kind=c_char), intent(in) :: path(*)
       integer(c_int32_t), value :: istep, dumpno, single
       real(c_double), value :: time, dt
     end function nkc_restart_write

     integer(c_int32_t) function nkc_restart_read(h, path, istep, dumpno, time, dt) &
          bind(C, name='nkc_restart_read')
       import :: c_int32_t, c_ptr, c_double, c_char
       type(c_ptr), value :: h
       character(kind=c_char), intent(in) :: path(*)
       integer(c_int32_t), intent(out) :: istep, dumpno
       real(c_double), intent(out) :: time, dt
     end function nkc_restart_read

     ! 1: khn/ken after nkc_step are rk4_upd's register (last stage writes it)
     integer(c_int32_t) function nkc_set_keep_k(h, keep) bind(C, name='nkc_set_keep_k')
       import :: c_int32_t, c_ptr
       type(c_ptr), value :: h
       integer(c_int32_t), value :: keep
     end function nkc_set_keep_k

     type(c_ptr) function nkc_last_error() bind(C, name='nkc_last_error')
       import :: c_ptr
     end function nkc_last_error
  end interface

contains

  ! Message of the last failed call as a Fortran string.
  function nkc_error_message() result(msg)
    character(len=512) :: msg
    character(kind=c_char), pointer :: s(:)
    type(c_ptr) :: p
    integer :: i
    msg = ' '
    p = nkc_last_error()
    if (.not. c_associated(p)) return
    call c_f_pointer(p, s, [512])
    do i = 1, 512
       if (s(i) == c_null_char) exit
       msg(i:i) = s(i)
    end do
  end function nkc_error_message

  ! Any nonzero status ends the run, as exitt(rc) does in the reference
  ! (src/nek5_comm_mpi.F:650-690).
  subroutine nkc_check(rc, where)
    integer(c_int32_t), intent(in) :: rc
    character(len=*), intent(in) :: where
    if (rc /= NKC_OK) then
       write(6, '(a,a,a,i0,a,a)') 'nekcem_hip: ', where, ' failed, status ', rc, ': ', &
            trim(nkc_error_message())
       error stop 1
    end if
  end subroutine nkc_check

end module nekcem_hip
