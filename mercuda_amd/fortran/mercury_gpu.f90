! mercury_gpu.f90 -- the MERCURY6 driver's GPU module, over libmercury_hip.
!
! The unchanged driver (mercury6_2.for) uses this module by name and calls the
! entities below (gpu_select/push/advance/pull/old/export/bcoord/find_events/
! central/limits/refresh_events/hybrid_begin/hybrid_finish/free and the raw
! encounter_enter/update/exit), so the public names, argument lists and the
! selection and failure rules are those of the CUDA binding it replaces
! (reference mercury_gpu.f90:1-248).  The library is the C ABI declared in
! include/mercury_hip.h.
!
! Backend requests come from mercury_support's set_backend: 0 cpu, 1 the
! device backend, 2 auto.  Step failures never fall back to the CPU; only a
! failed configure or first upload under `auto` does, and says so in info.out.
module mercury_gpu
  use iso_c_binding
  use mercury_support, only: backend_request, fail, force_calls, rejected_steps
  implicit none
  ! Public by default, as the module this replaces: the driver also reaches
  ! mercury_support's fail/force_calls/rejected_steps through `use mercury_gpu`.
  private :: loaded_n, nevents, info_path, event_list, note_fallback
  private :: hip_available, hip_configure, hip_reset, hip_limits, hip_export, hip_upload
  private :: hip_step, hip_download, hip_events, hip_hybrid_begin, hip_hybrid_finish

  logical :: gpu_enabled = .false.    ! the device backend integrates this run
  logical :: gpu_dirty = .true.       ! host arrays changed: upload before the next step
  logical :: host_current = .true.    ! host arrays hold the device state
  integer :: loaded_n = 0
  integer :: nevents = 0
  character(len=512) :: info_path = ''

  type, bind(C) :: hip_event
    integer(c_int) :: i, j, kind, now
    real(c_double) :: distance, time, xi(6), xj(6)
  end type
  type(hip_event), pointer :: event_list(:) => null()

  interface
    integer(c_int) function hip_available() bind(C, name='mercury_hip_available')
      import
    end function
    integer(c_int) function hip_configure(algorithm) bind(C, name='mercury_hip_configure')
      import
      integer(c_int), value :: algorithm
    end function
    subroutine hip_reset(flag) bind(C, name='mercury_hip_reset')
      import
      integer(c_int), value :: flag
    end subroutine
    integer(c_int) function hip_limits(rce) bind(C, name='mercury_hip_limits')
      import
      real(c_double), intent(in) :: rce(*)
    end function
    integer(c_int) function hip_export(h, physical, x, v) bind(C, name='mercury_hip_export')
      import
      real(c_double), value :: h
      integer(c_int), value :: physical
      real(c_double) :: x(*), v(*)
    end function
    integer(c_int) function hip_upload(n, nbig, pn, ngflag, m, x, v, ngf, jcen, rce, rphys) &
        bind(C, name='mercury_hip_upload')
      import
      integer(c_int), value :: n, nbig, pn, ngflag
      real(c_double), intent(in) :: m(*), x(*), v(*), ngf(*), jcen(*), rce(*), rphys(*)
    end function
    integer(c_int) function hip_step(time, h, hdid, tol, forces, rejected) bind(C, name='mercury_hip_step')
      import
      real(c_double), value :: time, tol
      real(c_double) :: h, hdid
      integer(c_int64_t) :: forces, rejected
    end function
    integer(c_int) function hip_download(x, v, previous) bind(C, name='mercury_hip_download')
      import
      real(c_double) :: x(*), v(*)
      integer(c_int), value :: previous
    end function
    integer(c_int) function hip_events(time, h, rcen, list, count) bind(C, name='mercury_hip_events')
      import
      real(c_double), value :: time, h, rcen
      type(c_ptr) :: list
      integer(c_int) :: count
    end function
    integer(c_int) function hip_hybrid_begin(h, crit, flag, cap, ce, nce, pi, pj, x, v, nf) &
        bind(C, name='mercury_hip_hybrid_begin')
      import
      real(c_double), value :: h
      integer(c_int), value :: flag, cap
      real(c_double) :: crit(*), x(*), v(*)
      integer(c_int) :: ce(*), nce, pi(*), pj(*)
      integer(c_int64_t) :: nf
    end function
    integer(c_int) function hip_hybrid_finish(h, m, x, v, nf) bind(C, name='mercury_hip_hybrid_finish')
      import
      real(c_double), value :: h
      real(c_double) :: m(*), x(*), v(*)
      integer(c_int64_t) :: nf
    end function
    integer(c_int) function encounter_enter(n, nb, m, x, v, crit, rce, rphys, np, pi, pj) &
        bind(C, name='mercury_hip_encounter_enter')
      import
      integer(c_int), value :: n, nb, np
      real(c_double) :: m(*), x(*), v(*), crit(*), rce(*), rphys(*)
      integer(c_int) :: pi(*), pj(*)
    end function
    integer(c_int) function encounter_update(m, x, v) bind(C, name='mercury_hip_encounter_update')
      import
      real(c_double) :: m(*), x(*), v(*)
    end function
    subroutine encounter_exit() bind(C, name='mercury_hip_encounter_exit')
    end subroutine
    subroutine gpu_free() bind(C, name='mercury_hip_free')
    end subroutine
  end interface

contains

  ! Decide the backend for this run and log it (reference mercury_gpu.f90:115-138).
  subroutine gpu_select(algor, n, nbig, m, opt, unit)
    integer, intent(in) :: algor, n, nbig, opt(8), unit
    real(8), intent(in) :: m(n)
    logical :: allowed, present
    allowed = (algor == 1 .or. algor == 2 .or. algor == 3 .or. algor == 4 .or. algor == 9 .or. &
               algor == 10) .and. opt(8) == 0
    present = hip_available() /= 0
    if (backend_request == 1 .and. .not. allowed) &
      call fail('HIP requires MVS/BS/BS2/RADAU/HYBRID and user-defined force = no')
    if (backend_request == 1 .and. .not. present) &
      call fail('HIP was requested but no HIP device/build is available')
    gpu_enabled = backend_request /= 0 .and. allowed .and. present
    if (backend_request == 2 .and. n - nbig < 4096) gpu_enabled = .false.
    gpu_dirty = .true.
    host_current = .true.
    loaded_n = 0
    inquire (unit=unit, name=info_path)
    if (.not. gpu_enabled) then
      write (unit, '(a)') ' Execution backend: CPU'
      return
    end if
    if (hip_configure(algor) /= 0) then
      if (backend_request /= 2) call fail('HIP algorithm initialization failed')
      gpu_enabled = .false.
      call gpu_free()
      call note_fallback(' Execution backend: CPU (HIP configuration failed; auto fallback)')
      return
    end if
    write (unit, '(a,i2)') ' Execution backend: HIP, algorithm ', algor
  end subroutine

  ! Append a fallback line to info.out beside the backend line.
  subroutine note_fallback(text)
    character(len=*), intent(in) :: text
    integer :: u
    logical :: isopen
    write (*, '(a)') text
    if (info_path == '') return
    inquire (file=info_path, opened=isopen, number=u)
    if (isopen) then
      write (u, '(a)') text
    else
      open (newunit=u, file=info_path, status='old', access='append')
      write (u, '(a)') text
      close (u)
    end if
  end subroutine

  ! Upload the host state when it changed (reference mercury_gpu.f90:139-152).
  subroutine gpu_push(n, nbig, m, x, v, ngf, jcen, rce, rphys, opt, ngflag)
    integer, intent(in) :: n, nbig, opt(8), ngflag
    real(8), intent(in) :: m(n), x(3,n), v(3,n), ngf(5,n), jcen(3), rce(n), rphys(n)
    if (.not. gpu_enabled .or. .not. gpu_dirty) return
    if (hip_upload(n, nbig, opt(7), ngflag, m, x, v, ngf, jcen, rce, rphys) /= 0) then
      if (backend_request == 2 .and. loaded_n == 0) then
        call note_fallback(' Execution backend: CPU (HIP initialization failed; auto fallback)')
        gpu_enabled = .false.
        call gpu_free()
        return
      end if
      call fail('HIP state upload failed')
    end if
    loaded_n = n
    gpu_dirty = .false.
    host_current = .true.
  end subroutine

  ! One step on the device; dtflag /= 2 discards the step history first.
  subroutine gpu_advance(time, h, hdid, tol, dtflag)
    real(8), intent(in) :: time, tol
    real(8), intent(inout) :: h
    real(8), intent(out) :: hdid
    integer, intent(inout) :: dtflag
    integer(c_int64_t) :: nf, nr
    call hip_reset(dtflag)
    if (hip_step(time, h, hdid, tol, nf, nr) /= 0) call fail('HIP integration step failed')
    force_calls = force_calls + nf
    rejected_steps = rejected_steps + nr
    dtflag = 2
    host_current = .false.
  end subroutine

  subroutine gpu_pull(x, v)
    real(8), intent(inout) :: x(3,*), v(3,*)
    if (.not. gpu_enabled .or. host_current) return
    if (hip_download(x, v, 0) /= 0) call fail('HIP download failed')
    host_current = .true.
  end subroutine

  ! State before the last step (central-impact redo).
  subroutine gpu_old(x, v)
    real(8), intent(out) :: x(3,*), v(3,*)
    if (hip_download(x, v, 1) /= 0) call fail('HIP previous state download failed')
  end subroutine

  subroutine gpu_export(h, physical, x, v)
    real(8), intent(in) :: h
    integer, intent(in) :: physical
    real(8), intent(out) :: x(3,*), v(3,*)
    if (hip_export(h, physical, x, v) /= 0) call fail('HIP coordinate export failed')
  end subroutine

  ! Output coordinates: the device corrector when the device holds the state,
  ! otherwise the driver's own transform.
  subroutine gpu_bcoord(time, jcen, n, nbig, h, m, x, v, xh, vh, ngf, ngflag, opt, bcoord)
    integer :: n, nbig, ngflag, opt(8)
    real(8) :: time, jcen(3), h, m(n), x(3,n), v(3,n), xh(3,n), vh(3,n), ngf(5,n)
    external bcoord
    if (gpu_enabled .and. .not. gpu_dirty) then
      call gpu_export(h, 1, xh, vh)
    else
      call bcoord(time, jcen, n, nbig, h, m, x, v, xh, vh, ngf, ngflag, opt)
    end if
  end subroutine

  subroutine gpu_limits(rce)
    real(8), intent(in) :: rce(*)
    if (.not. gpu_enabled .or. gpu_dirty) return
    if (hip_limits(rce) /= 0) call fail('HIP encounter-radius update failed')
  end subroutine

  ! Screen the last step; the event list stays library-owned until the next call.
  subroutine gpu_refresh_events(time, h, rcen)
    real(8) :: time, h, rcen
    type(c_ptr) :: list
    if (hip_events(time, h, rcen, list, nevents) /= 0) call fail('HIP event screening failed')
    nullify (event_list)
    if (nevents > 0) call c_f_pointer(list, event_list, [nevents])
  end subroutine

  ! Split the event list into mce_stat's encounter and collision lists
  ! (kind bit 1: encounter minimum; 2: near miss; 4: collision).
  subroutine gpu_find_events(time, h, rcen, cap, nclo, iclo, jclo, dclo, tclo, ixv, jxv, &
      nhit, ihit, jhit, chit, dhit, thit, thit1, nowflag)
    real(8), intent(in) :: time, h, rcen
    integer, intent(in) :: cap
    integer, intent(out) :: nclo, nhit, nowflag, iclo(cap), jclo(cap), ihit(cap), jhit(cap), chit(cap)
    real(8), intent(out) :: dclo(cap), tclo(cap), ixv(6,cap), jxv(6,cap), dhit(cap), thit(cap), thit1
    integer :: k
    call gpu_refresh_events(time, h, rcen)
    nclo = 0
    nhit = 0
    nowflag = 0
    thit1 = sign(9.9d29, h)
    do k = 1, nevents
      associate (e => event_list(k))
        if (iand(e%kind, 1) /= 0) then
          nclo = nclo + 1
          if (nclo > cap) call fail('Encounter buffer capacity exceeded')
          iclo(nclo) = e%i
          jclo(nclo) = e%j
          dclo(nclo) = e%distance
          tclo(nclo) = e%time
          ixv(:, nclo) = e%xi
          jxv(:, nclo) = e%xj
        end if
        if (iand(e%kind, 2) /= 0) then
          nhit = nhit + 1
          if (nhit > cap) call fail('Collision buffer capacity exceeded')
          ihit(nhit) = e%i
          jhit(nhit) = e%j
          chit(nhit) = merge(1, 0, iand(e%kind, 4) /= 0)
          dhit(nhit) = e%distance
          thit(nhit) = e%time
          nowflag = max(nowflag, e%now)
          if ((thit(nhit) - time) * h < (thit1 - time) * h) thit1 = thit(nhit)
        end if
      end associate
    end do
  end subroutine

  ! Central-body impacts of the last screen (kind bit 8).
  subroutine gpu_central(cap, nhit, jhit, thit, dhit)
    integer, intent(in) :: cap
    integer, intent(out) :: nhit, jhit(cap)
    real(8), intent(out) :: thit(cap), dhit(cap)
    integer :: k
    nhit = 0
    do k = 1, nevents
      if (iand(event_list(k)%kind, 8) == 0) cycle
      nhit = nhit + 1
      if (nhit > cap) call fail('Central collision buffer capacity exceeded')
      jhit(nhit) = event_list(k)%j
      thit(nhit) = event_list(k)%time
      dhit(nhit) = event_list(k)%distance
    end do
  end subroutine

  ! HYBRID: first half of the step and the changeover scan; only encounter
  ! members' x, v are written back.
  subroutine gpu_hybrid_begin(h, crit, flag, cap, ce, nce, pi, pj, x, v)
    real(8) :: h, crit(*), x(3,*), v(3,*)
    integer :: flag, cap, ce(*), nce, pi(*), pj(*)
    integer(c_int64_t) :: nf
    if (hip_hybrid_begin(h, crit, flag, cap, ce, nce, pi, pj, x, v, nf) /= 0) &
      call fail('HIP hybrid drift failed')
    force_calls = force_calls + nf
    flag = 2
  end subroutine

  subroutine gpu_hybrid_finish(h, m, x, v)
    real(8) :: h, m(*), x(3,*), v(3,*)
    integer(c_int64_t) :: nf
    if (hip_hybrid_finish(h, m, x, v, nf) /= 0) call fail('HIP hybrid kick failed')
    force_calls = force_calls + nf
    host_current = .false.
  end subroutine

end module
