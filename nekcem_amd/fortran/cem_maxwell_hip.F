c     cem_maxwell_hip.F -- the drop-in for NekCEM's RK Maxwell step.
c
c     A NekCEM build compiles this file with the rest of its src/ (fixed
c     form, -fdefault-real-8, the case's SIZE on the include path) and
c     links libnekcem_hip.so plus the nekcem_hip module object; it then
c     drops the CPU body of cem_maxwell_op_rk (src/cem_maxwell.F:327-345)
c     for this one.  The COMMON arrays cem_maxwell_init leaves behind
c     (src/EMWAVE, src/GEOM, src/INPUT, /bdry1/) cross the C-ABI as they
c     are, at the COMMON strides lpts1 / lxzfl.
c
c       cem_maxwell_hip_init        once, after cem_maxwell_init, where
c                                   acc_copy_all_in sits
c                                   (src/cem_drive.F:397)
c       cem_maxwell_op_rk           one RK5 step on the device
c       cem_maxwell_hip_update_host the !$ACC UPDATE HOST of userchk and
c                                   cem_out (src/io.F:193-195)
c       cem_maxwell_hip_pml         the PML arrays of cem_maxwell_init,
c                                   once, when ifpml
c       cem_maxwell_hip_restart_out cem_restart_out (src/io.F:411-490)
c       cem_maxwell_hip_restart_swap restart_swap (src/io.F:637-788)
c       cem_maxwell_hip_end         acc_copy_all_out (src/cem_drive.F:482)
c
c     Any nonzero status ends the run through nkc_check (error stop 1),
c     the analogue of exitt(rc).

c     The device handle, kept between calls.
      module cem_hip_state
      use iso_c_binding
      implicit none
      type(c_ptr), save :: hnkc = c_null_ptr
      contains
c     c_loc of a COMMON array: the array is passed to a TARGET dummy
c     (sequence association), whose address is the COMMON member's.
      function cem_hip_dptr(a) result(p)
      real(c_double), target, intent(in) :: a(*)
      type(c_ptr) :: p
      p = c_loc(a)
      end function
      function cem_hip_iptr(a) result(p)
      integer(c_int32_t), target, intent(in) :: a(*)
      type(c_ptr) :: p
      p = c_loc(a)
      end function
      end module

      subroutine cem_maxwell_hip_init
      use iso_c_binding
      use nekcem_hip
      use cem_hip_state
      implicit none
      include 'SIZE'
      include 'TOTAL'
      include 'EMWAVE'
      common /bdry1/ cempec(6*lx1*lz1*2*ldim*lelt),ncempec
      integer cempec,ncempec
      common /scrns/ srflx(6*2*ldim*lx1*lz1*lelt)
      real srflx
      integer, save :: partner(2*ldim*lx1*lz1*lelt)
      integer i
      type(nkc_desc) :: d
c     The face partner map from the gs handle itself: summing the face
c     index over gsh_face gives i + partner(i) on paired nodes, i on
c     unpaired ones (one rank holds the whole mesh on this GPU).
      do i = 1,nxzfl
         srflx(i) = i
      enddo
      call gs_op_fields(gsh_face,srflx,nxzfl,1,1,1,0)
      do i = 1,nxzfl
         partner(i) = nint(srflx(i)) - i
      enddo
      d%N = nx1-1
      d%ifupwind = merge(1, 0, ifupwind)
      d%nelt = nelt
      d%dxm1 = cem_hip_dptr(dxm1)
      d%wxm1 = cem_hip_dptr(wxm1)
      d%rxmn = cem_hip_dptr(rxmn)
      d%rymn = cem_hip_dptr(rymn)
      d%rzmn = cem_hip_dptr(rzmn)
      d%sxmn = cem_hip_dptr(sxmn)
      d%symn = cem_hip_dptr(symn)
      d%szmn = cem_hip_dptr(szmn)
      d%txmn = cem_hip_dptr(txmn)
      d%tymn = cem_hip_dptr(tymn)
      d%tzmn = cem_hip_dptr(tzmn)
      d%jacm = cem_hip_dptr(jacm)
      d%bmn  = cem_hip_dptr(bmn)
      d%unxm = cem_hip_dptr(unxm)
      d%unym = cem_hip_dptr(unym)
      d%unzm = cem_hip_dptr(unzm)
      d%aream = cem_hip_dptr(aream)
      d%cemface = cem_hip_iptr(cemface)
      d%face_partner = cem_hip_iptr(partner)
      d%cempec = cem_hip_iptr(cempec)
      d%ncempec = ncempec
      d%permittivity = cem_hip_dptr(permittivity)
      d%permeability = cem_hip_dptr(permeability)
      d%Z_0 = cem_hip_dptr(Z_0)
      d%Z_1 = cem_hip_dptr(Z_1)
      d%Y_0 = cem_hip_dptr(Y_0)
      d%Y_1 = cem_hip_dptr(Y_1)
      call nkc_check(nkc_create(d, hnkc), 'nkc_create')
c     khn/ken read back after a step are then rk4_upd's register
      call nkc_check(nkc_set_keep_k(hnkc, 1), 'nkc_set_keep_k')
      call nkc_check(nkc_set_state(hnkc, hn, en, cem_hip_dptr(khn),
     $     cem_hip_dptr(ken), int(lpts1, c_int64_t)), 'nkc_set_state')
      return
      end

c     The new body of cem_maxwell_op_rk: one RK5 step on the device.
c     time itself is advanced by time_advancing_pde (src/cem_drive.F:639).
      subroutine cem_maxwell_op_rk
      use iso_c_binding
      use nekcem_hip
      use cem_hip_state
      implicit none
      include 'SIZE'
      include 'TOTAL'
      real(c_double) :: tnow
      tnow = time
      call nkc_check(nkc_step(hnkc, tnow, dt, 1_c_int64_t), 'nkc_step')
      return
      end

c     At the host-sync points: the fields and the RK register back into
c     COMMON.
      subroutine cem_maxwell_hip_update_host
      use iso_c_binding
      use nekcem_hip
      use cem_hip_state
      implicit none
      include 'SIZE'
      include 'TOTAL'
      include 'EMWAVE'
      call nkc_check(nkc_get_state(hnkc, hn, en, cem_hip_dptr(khn),
     $     cem_hip_dptr(ken), int(lpts1, c_int64_t)), 'nkc_get_state')
      return
      end

c     The PML arrays cem_maxwell_init leaves in the PML common blocks,
c     with userini's pmlbn / pmldn (src/cem_drive.F:444-448).
      subroutine cem_maxwell_hip_pml
      use iso_c_binding
      use nekcem_hip
      use cem_hip_state
      implicit none
      include 'SIZE'
      include 'TOTAL'
      include 'EMWAVE'
      include 'PML'
      if (.not. ifpml) return
      call nkc_check(nkc_set_pml(hnkc, int(maxpml, c_int64_t), pmlptr,
     $     pmlsigma, cem_hip_dptr(pmlbn), cem_hip_dptr(pmldn),
     $     int(lpts1, c_int64_t)), 'nkc_set_pml')
      return
      end

c     cem_restart_out (src/io.F:411-490) for the device state: the file
c     restart-mpi-binary-N1-t<dumpno>.vtk of getfilename_restart
c     (src/io_mpiutil.c:41-60) in the vtk directory, doubles unless
c     param(87) asks for floats.
      subroutine cem_maxwell_hip_restart_out(dumpno)
      use iso_c_binding
      use nekcem_hip
      use cem_hip_state
      implicit none
      include 'SIZE'
      include 'TOTAL'
      integer dumpno
      character(len=64) fname
      integer(c_int32_t) single
      write(fname,'(a,i5.5,a)') 'vtk/restart-mpi-binary-N1-t',
     $     dumpno,'.vtk'
      single = 0
      if (nint(param(87)).ne.0) single = 1
      call nkc_check(nkc_restart_write(hnkc, trim(fname)//c_null_char,
     $     int(istep, c_int32_t), int(dumpno, c_int32_t), time, dt,
     $     single), 'nkc_restart_write')
      return
      end

c     restart_swap (src/io.F:637-788): the state from the restart file of
c     dump param(84); the header's step goes to irststep (INPUT) and its
c     time is added to the run's time.
      subroutine cem_maxwell_hip_restart_swap
      use iso_c_binding
      use nekcem_hip
      use cem_hip_state
      implicit none
      include 'SIZE'
      include 'TOTAL'
      character(len=64) fname
      integer(c_int32_t) idump
      real(c_double) rtime, rdt
      write(fname,'(a,i5.5,a)') 'vtk/restart-mpi-binary-N1-t',
     $     nint(param(84)),'.vtk'
      call nkc_check(nkc_restart_read(hnkc, trim(fname)//c_null_char,
     $     irststep, idump, rtime, rdt), 'nkc_restart_read')
      time = time + rtime
      return
      end

      subroutine cem_maxwell_hip_end
      use iso_c_binding
      use nekcem_hip
      use cem_hip_state
      implicit none
      call nkc_check(nkc_destroy(hnkc), 'nkc_destroy')
      hnkc = c_null_ptr
      return
      end
