!-----------------------------------------------------------------------
! tc_host: bou.in-driven Taylor-Couette driver on one MI355X.
!
! Keeps the reference's program structure and routine names (main.F90,
! TimeMarcher.F90, ReadInputFile.F90, ExplicitTerms*, ImplicitAndUpdate*,
! CalcLocalDivergence, SolvePressureCorrection, CorrectVelocity,
! CorrectPressure, CalcMaxCFL, CheckDivergence, CalcTorque,
! CalcVelocityStats); every routine forwards to the C-ABI of libtcamd.so
! (include/tc_abi.h), where the state lives on the device.
!
! Usage: tc_host [--fused] [--ic perturbed SEED AMP] [--restart FILE]
!                [--checkpoint FILE]
!   --fused      one fused device call per time step (tc_time_marcher)
!                instead of the routine-by-routine calls
!   --ic ...     seeded perturbed Couette start, projected divergence free;
!                without it NREAD=0 uses CreateInitialConditions and
!                NREAD=1 reads continua_{qt,qr,qz}.bin with
!                continua_master.txt, interpolating from another grid
!   --restart F  continue from a tc_write_restart checkpoint (bit for bit)
!   --checkpoint F  write a checkpoint of the final state
! Outputs: torquenu.out (time, inner, outer; e14.6 as CalcTorque.F90:59),
!          torquenu_full.out (same at full precision), vmax.out,
!          dissipnu.out (time, nute once time > TSTA),
!          continua_{qt,qr,qz,pr}.bin at exit (raw FP64, (nr,nzm,nthm)),
!          stats/<name>.bin (raw FP64, (nrm,nzm)) and
!          stats/averaging_time.txt on a normal exit; once time > TSTA the
!          slab dumps rcuts/, thcut/, stave/ every TFRAME2D and fullff/
!          every TFRAME3D (raw FP64, file names as the reference's .h5).
!-----------------------------------------------------------------------
module tc_abi
  use iso_c_binding
  implicit none
  type, bind(C) :: tc_params
    integer(c_int) :: nthm, nzm, nrm, nsst, idtv, nread
    integer(c_int) :: ntst
    real(c_double) :: tout, tmax
    integer(c_int) :: ireset
    real(c_double) :: walltimemax
    real(c_double) :: alx3, rint, rext
    integer(c_int) :: istr
    real(c_double) :: str, lamb
    real(c_double) :: virotint, virotext, sizepert, omegapert
    real(c_double) :: ren, romeg, resid, limitcfl
    real(c_double) :: tsta
    integer(c_int) :: starea
    real(c_double) :: dtmin, dtmax, limitvel, cfllim
    integer(c_int) :: nslabr, dpthcut, dpstave
    real(c_double) :: tframe2d
    integer(c_int) :: dpfullf
    real(c_double) :: tframe3d
    integer(c_int) :: minstrt, minstrz, multmin
  end type
  integer(c_int), parameter :: TC_QT = 0, TC_QR = 1, TC_QZ = 2, TC_PR = 3
  integer(c_int), parameter :: TC_LAYOUT_FORTRAN = 0
  integer(c_int), parameter :: TC_V_RM = 1
  interface
    integer(c_int) function tc_create(p, ctx) bind(C, name='tc_create')
      import
      type(tc_params), intent(in) :: p
      type(c_ptr), intent(out) :: ctx
    end function
    integer(c_int) function tc_destroy(ctx) bind(C, name='tc_destroy')
      import
      type(c_ptr), value :: ctx
    end function
    type(c_ptr) function tc_last_error() bind(C, name='tc_last_error')
      import
    end function
    integer(c_int) function tc_grid_vector(p, id, v, n) bind(C, name='tc_grid_vector')
      import
      type(tc_params), intent(in) :: p
      integer(c_int), value :: id, n
      real(c_double), intent(out) :: v(*)
    end function
    integer(c_int) function tc_upload_field(ctx, f, a, layout) bind(C, name='tc_upload_field')
      import
      type(c_ptr), value :: ctx
      integer(c_int), value :: f, layout
      real(c_double), intent(in) :: a(*)
    end function
    integer(c_int) function tc_download_field(ctx, f, a, layout) bind(C, name='tc_download_field')
      import
      type(c_ptr), value :: ctx
      integer(c_int), value :: f, layout
      real(c_double), intent(out) :: a(*)
    end function
    integer(c_int) function tc_set_time(ctx, t, dt) bind(C, name='tc_set_time')
      import
      type(c_ptr), value :: ctx
      real(c_double), value :: t, dt
    end function
    integer(c_int) function tc_set_substep(ctx, ns) bind(C, name='tc_set_substep')
      import
      type(c_ptr), value :: ctx
      integer(c_int), value :: ns
    end function
    integer(c_int) function tc_init_perturbed_couette(p, seed, amp, qt, qr, qz) &
        bind(C, name='tc_init_perturbed_couette')
      import
      type(tc_params), intent(in) :: p
      integer(c_long_long), value :: seed
      real(c_double), value :: amp
      real(c_double), intent(out) :: qt(*), qr(*), qz(*)
    end function
    integer(c_int) function tc_calc_max_cfl(ctx, v) bind(C, name='tc_calc_max_cfl')
      import
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: v
    end function
    integer(c_int) function tc_check_divergence(ctx, v) bind(C, name='tc_check_divergence')
      import
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: v
    end function
    integer(c_int) function tc_calc_torque(ctx, a, b) bind(C, name='tc_calc_torque')
      import
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: a, b
    end function
    integer(c_int) function tc_calc_velocity_stats(ctx, v, r) bind(C, name='tc_calc_velocity_stats')
      import
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: v(3), r(3)
    end function
    ! argument-less routine entries (one C function each)
    integer(c_int) function tc_begin_step(ctx) bind(C, name='tc_begin_step')
      import
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function tc_explicit_terms_qt(ctx) bind(C, name='tc_explicit_terms_qt')
      import
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function tc_explicit_terms_qr(ctx) bind(C, name='tc_explicit_terms_qr')
      import
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function tc_explicit_terms_qz(ctx) bind(C, name='tc_explicit_terms_qz')
      import
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function tc_implicit_and_update_qt(ctx) bind(C, name='tc_implicit_and_update_qt')
      import
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function tc_implicit_and_update_qr(ctx) bind(C, name='tc_implicit_and_update_qr')
      import
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function tc_implicit_and_update_qz(ctx) bind(C, name='tc_implicit_and_update_qz')
      import
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function tc_calc_local_divergence(ctx) bind(C, name='tc_calc_local_divergence')
      import
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function tc_solve_pressure_correction(ctx) bind(C, name='tc_solve_pressure_correction')
      import
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function tc_correct_velocity(ctx) bind(C, name='tc_correct_velocity')
      import
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function tc_correct_pressure(ctx) bind(C, name='tc_correct_pressure')
      import
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function tc_time_marcher(ctx) bind(C, name='tc_time_marcher')
      import
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function tc_project(ctx) bind(C, name='tc_project')
      import
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function tc_synchronize(ctx) bind(C, name='tc_synchronize')
      import
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function tc_calc_stats(ctx) bind(C, name='tc_calc_stats')
      import
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function tc_calc_dissipation(ctx, nute) bind(C, name='tc_calc_dissipation')
      import
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: nute
    end function
    integer(c_int) function tc_get_stats(ctx, a, nsamples) bind(C, name='tc_get_stats')
      import
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: a(*)
      integer(c_int), intent(out) :: nsamples
    end function
    integer(c_int) function tc_regrid_field(ctx, f, a, ntho, nzo, nro, istro, stro) &
        bind(C, name='tc_regrid_field')
      import
      type(c_ptr), value :: ctx
      integer(c_int), value :: f, ntho, nzo, nro, istro
      real(c_double), intent(in) :: a(*)
      real(c_double), value :: stro
    end function
    integer(c_int) function tc_dump_rcut(ctx, k, vt, vr, vz) bind(C, name='tc_dump_rcut')
      import
      type(c_ptr), value :: ctx
      integer(c_int), value :: k
      real(c_double), intent(out) :: vt(*), vr(*), vz(*)
    end function
    integer(c_int) function tc_dump_thcut(ctx, vt, vr, vz) bind(C, name='tc_dump_thcut')
      import
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: vt(*), vr(*), vz(*)
    end function
    integer(c_int) function tc_dump_streamwise_avg(ctx, vt, vr, vz) &
        bind(C, name='tc_dump_streamwise_avg')
      import
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: vt(*), vr(*), vz(*)
    end function
    integer(c_int) function tc_get_scalars(ctx, out8) bind(C, name='tc_get_scalars')
      import
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: out8(8)
    end function
    integer(c_int) function tc_write_restart(ctx, path, ntime) bind(C, name='tc_write_restart')
      import
      type(c_ptr), value :: ctx
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int), value :: ntime
    end function
    integer(c_int) function tc_read_restart(ctx, path, ntime) bind(C, name='tc_read_restart')
      import
      type(c_ptr), value :: ctx
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int), intent(out) :: ntime
    end function
  end interface
contains
  subroutine tc_check(ierr, where)
    integer(c_int), intent(in) :: ierr
    character(len=*), intent(in) :: where
    character(kind=c_char), pointer :: msg(:)
    integer :: n
    if (ierr >= 0) return
    call c_f_pointer(tc_last_error(), msg, [4096])
    n = 0
    do while (n < 4096)
      if (msg(n + 1) == c_null_char) exit
      n = n + 1
    end do
    write(6, '(a,a,a,i0,a)', advance='no') 'tc_host: ', where, ' failed (', ierr, '): '
    write(6, '(4096a)') msg(1:n)
    error stop 1
  end subroutine
end module tc_abi

! Host mirror of the param module (param.F90:4-88): bou.in values and the
! time-stepping scalars; the fields themselves live on the device.
module param
  use iso_c_binding
  use tc_abi, only: tc_params
  implicit none
  type(tc_params) :: bou
  type(c_ptr) :: ctx = c_null_ptr
  integer :: nthm, nzm, nrm, nr, nsst, ntst, idtv, nread
  real :: time = 0.d0, dt = 0.d0, tout, tmax, walltimemax, resid, limitcfl, limitvel
  real :: dtmin, dtmax
  integer :: ntime
  logical :: variabletstep = .true., readflow = .false., use_fused = .false.
  logical :: perturbed_ic = .false.
  character(len=512) :: restart_file = '', checkpoint_file = ''
  integer :: ntime0 = 0
  integer :: nslab = 0, kslab(9) = 1   ! DumpRCuts planes (NSLABR, InitDumpRCuts)
  integer(c_long_long) :: ic_seed = 1
  real :: ic_amp = 1.d-2
  real :: vmax(3)
end module param

! ReadInputFile.F90:17-39: eleven header/value line pairs
subroutine ReadInputFile
  use param
  implicit none
  character(len=4) :: dummy
  integer :: u, ireset, istr, starea, nslabr, dpthcut, dpstave, dpfullf
  integer :: minstrt, minstrz, multmin
  real :: alx3, rint, rext, str, lamb, virotint, virotext, sizepert, omegapert
  real :: ren, romeg, tsta, cfllim, tframe2d, tframe3d
  open(newunit=u, file='bou.in', status='old')
  read(u, '(a4)') dummy
  read(u, *) nthm, nzm, nrm, nsst, idtv, nread
  read(u, '(a4)') dummy
  read(u, *) ntst, tout, tmax, ireset, walltimemax
  read(u, '(a4)') dummy
  read(u, *) alx3, rint, rext, istr, str, lamb
  read(u, '(a4)') dummy
  read(u, *) virotint, virotext, sizepert, omegapert
  read(u, '(a4)') dummy
  read(u, *) ren, romeg, resid, limitcfl
  read(u, '(a4)') dummy
  read(u, *) tsta, starea
  read(u, '(a4)') dummy
  read(u, *) dtmin, dtmax, limitvel, cfllim
  read(u, '(a4)') dummy
  read(u, *) nslabr, dpthcut, dpstave, tframe2d
  read(u, '(a4)') dummy
  read(u, *) dpfullf, tframe3d
  read(u, '(a4)') dummy
  read(u, *) minstrt, minstrz, multmin
  close(u)
  nr = nrm + 1
  if (idtv == 0) variabletstep = .false.
  if (nread /= 0) readflow = .true.
  bou%nthm = nthm; bou%nzm = nzm; bou%nrm = nrm; bou%nsst = nsst
  bou%idtv = idtv; bou%nread = nread; bou%ntst = ntst; bou%tout = tout
  bou%tmax = tmax; bou%ireset = ireset; bou%walltimemax = walltimemax
  bou%alx3 = alx3; bou%rint = rint; bou%rext = rext; bou%istr = istr
  bou%str = str; bou%lamb = lamb; bou%virotint = virotint
  bou%virotext = virotext; bou%sizepert = sizepert; bou%omegapert = omegapert
  bou%ren = ren; bou%romeg = romeg; bou%resid = resid; bou%limitcfl = limitcfl
  bou%tsta = tsta; bou%starea = starea; bou%dtmin = dtmin; bou%dtmax = dtmax
  bou%limitvel = limitvel; bou%cfllim = cfllim; bou%nslabr = nslabr
  bou%dpthcut = dpthcut; bou%dpstave = dpstave; bou%tframe2d = tframe2d
  bou%dpfullf = dpfullf; bou%tframe3d = tframe3d; bou%minstrt = minstrt
  bou%minstrz = minstrz; bou%multmin = multmin
end subroutine ReadInputFile

! CreateInitialConditions.F90:11-49 (the reference's default start)
subroutine CreateInitialConditions
  use param
  use tc_abi
  implicit none
  real, allocatable :: qt(:,:,:), qr(:,:,:), qz(:,:,:), rm(:)
  real :: eps, xxx, yyy, pi, zm, dxz
  integer :: i, j, k, m
  allocate(qt(nr, nzm, nthm), qr(nr, nzm, nthm), qz(nr, nzm, nthm), rm(nr))
  m = tc_grid_vector(bou, TC_V_RM, rm, nr)
  call tc_check(m, 'tc_grid_vector')
  pi = 2.d0 * asin(1.d0)
  dxz = dble(nzm) / bou%alx3
  eps = 0.1d0
  qt = 0.d0; qr = 0.d0; qz = 0.d0
  do i = 1, nthm
    do j = 1, nzm
      zm = (dble(j - 1) + 0.5d0) / dxz
      do k = 1, nrm
        yyy = zm / bou%alx3
        qr(k, j, i) = eps * (sin(yyy*2*pi) + sin(yyy*4*pi) + sin(yyy*8*pi) + sin(yyy*6*pi))
        xxx = (rm(k) - bou%rint) / (bou%rext - bou%rint)
        qz(k, j, i) = (sin(xxx*2*pi)*(1-abs(sin(yyy*2*pi)))*sin(yyy*2*pi)*eps) + &
                      (sin(xxx*2*pi)*(1-abs(sin(yyy*4*pi)))*sin(yyy*4*pi)*eps) + &
                      (sin(xxx*2*pi)*(1-abs(sin(yyy*6*pi)))*sin(yyy*6*pi)*eps) + &
                      (sin(xxx*2*pi)*(1-abs(sin(yyy*8*pi)))*sin(yyy*8*pi)*eps)
      end do
    end do
  end do
  call tc_check(tc_upload_field(ctx, TC_QT, qt, TC_LAYOUT_FORTRAN), 'upload qt')
  call tc_check(tc_upload_field(ctx, TC_QR, qr, TC_LAYOUT_FORTRAN), 'upload qr')
  call tc_check(tc_upload_field(ctx, TC_QZ, qz, TC_LAYOUT_FORTRAN), 'upload qz')
end subroutine CreateInitialConditions

subroutine PerturbedCouetteStart
  use param
  use tc_abi
  implicit none
  real, allocatable :: qt(:,:,:), qr(:,:,:), qz(:,:,:)
  allocate(qt(nr, nzm, nthm), qr(nr, nzm, nthm), qz(nr, nzm, nthm))
  call tc_check(tc_init_perturbed_couette(bou, ic_seed, ic_amp, qt, qr, qz), 'initial field')
  call tc_check(tc_upload_field(ctx, TC_QT, qt, TC_LAYOUT_FORTRAN), 'upload qt')
  call tc_check(tc_upload_field(ctx, TC_QR, qr, TC_LAYOUT_FORTRAN), 'upload qr')
  call tc_check(tc_upload_field(ctx, TC_QZ, qz, TC_LAYOUT_FORTRAN), 'upload qz')
  call tc_check(tc_project(ctx), 'projection')
end subroutine PerturbedCouetteStart

! ReadFlowField.F90:12-152 with raw files: continua_master.txt holds the
! scalars of continua_master.h5 (nr nth nz time rint rext istr str alx3
! lamb, as WriteFlowField writes them) and continua_<field>.bin the
! (nr, nzm, nthm) FP64 fields of the old grid.  Another grid or stretching
! is interpolated on the device (tc_regrid_field, interp.F90).
subroutine ReadFlowField
  use param
  use tc_abi
  implicit none
  real, allocatable :: a(:,:,:)
  character(len=2), parameter :: names(3) = ['qt', 'qr', 'qz']
  integer(c_int), parameter :: ids(3) = [TC_QT, TC_QR, TC_QZ]
  integer :: q, u, ios, nro, ntho, nzo, istro
  real :: timeo, rinto, rexto, stro, alx3o, lambo
  open(newunit=u, file='continua_master.txt', status='old', iostat=ios)
  if (ios /= 0) then
    write(6, *) 'Continuation files not found'
    stop 3
  end if
  read(u, *) nro, ntho, nzo, timeo, rinto, rexto, istro, stro, alx3o, lambo
  close(u)
  if (nzo /= nzm + 1 .or. nro /= nr .or. ntho /= nthm + 1 .or. istro /= bou%istr .or. &
      stro /= bou%str) then
    write(6, *) 'Interpolating new grid'
    allocate(a(nro, nzo - 1, ntho - 1))
    do q = 1, 3
      open(newunit=u, file='continua_'//names(q)//'.bin', access='stream', form='unformatted', &
           status='old')
      read(u) a
      close(u)
      call tc_check(tc_regrid_field(ctx, ids(q), a, int(ntho, c_int), int(nzo, c_int), &
                                    int(nro, c_int), int(istro, c_int), stro), 'regrid '//names(q))
    end do
  else
    allocate(a(nr, nzm, nthm))
    do q = 1, 3
      open(newunit=u, file='continua_'//names(q)//'.bin', access='stream', form='unformatted', &
           status='old')
      read(u) a
      close(u)
      call tc_check(tc_upload_field(ctx, ids(q), a, TC_LAYOUT_FORTRAN), 'upload '//names(q))
    end do
  end if
  time = timeo
  if (bou%ireset /= 0) time = 0.d0  ! resetlogstime
  tmax = tmax + time
  bou%tmax = tmax
end subroutine ReadFlowField

subroutine WriteFlowField
  use param
  use tc_abi
  implicit none
  real, allocatable :: a(:,:,:)
  character(len=2), parameter :: names(4) = ['qt', 'qr', 'qz', 'pr']
  integer(c_int), parameter :: ids(4) = [TC_QT, TC_QR, TC_QZ, TC_PR]
  integer :: q, u
  allocate(a(nr, nzm, nthm))
  do q = 1, 4
    call tc_check(tc_download_field(ctx, ids(q), a, TC_LAYOUT_FORTRAN), 'download '//names(q))
    open(newunit=u, file='continua_'//names(q)//'.bin', access='stream', form='unformatted', status='replace')
    write(u) a
    close(u)
  end do
  ! the scalars of continua_master.h5, WriteFlowField.F90:24-49
  open(newunit=u, file='continua_master.txt', status='replace')
  write(u, *) nr, nthm + 1, nzm + 1, time, bou%rint, bou%rext, bou%istr, bou%str, bou%alx3, bou%lamb
  close(u)
end subroutine WriteFlowField

! ---- the hot-path routines: one device call each ----
subroutine ExplicitTermsQT
  use param
  use tc_abi
  call tc_check(tc_explicit_terms_qt(ctx), 'ExplicitTermsQT')
end subroutine
subroutine ExplicitTermsQR
  use param
  use tc_abi
  call tc_check(tc_explicit_terms_qr(ctx), 'ExplicitTermsQR')
end subroutine
subroutine ExplicitTermsQZ
  use param
  use tc_abi
  call tc_check(tc_explicit_terms_qz(ctx), 'ExplicitTermsQZ')
end subroutine
subroutine ImplicitAndUpdateQT
  use param
  use tc_abi
  call tc_check(tc_implicit_and_update_qt(ctx), 'ImplicitAndUpdateQT')
end subroutine
subroutine ImplicitAndUpdateQR
  use param
  use tc_abi
  call tc_check(tc_implicit_and_update_qr(ctx), 'ImplicitAndUpdateQR')
end subroutine
subroutine ImplicitAndUpdateQZ
  use param
  use tc_abi
  call tc_check(tc_implicit_and_update_qz(ctx), 'ImplicitAndUpdateQZ')
end subroutine
subroutine CalcLocalDivergence
  use param
  use tc_abi
  call tc_check(tc_calc_local_divergence(ctx), 'CalcLocalDivergence')
end subroutine
subroutine SolvePressureCorrection
  use param
  use tc_abi
  call tc_check(tc_solve_pressure_correction(ctx), 'SolvePressureCorrection')
end subroutine
subroutine CorrectVelocity
  use param
  use tc_abi
  call tc_check(tc_correct_velocity(ctx), 'CorrectVelocity')
end subroutine
subroutine CorrectPressure
  use param
  use tc_abi
  call tc_check(tc_correct_pressure(ctx), 'CorrectPressure')
end subroutine

subroutine CalcMaxCFL(cflm)
  use param
  use tc_abi
  real, intent(out) :: cflm
  call tc_check(tc_calc_max_cfl(ctx, cflm), 'CalcMaxCFL')
end subroutine
subroutine CheckDivergence(qmax)
  use param
  use tc_abi
  real, intent(out) :: qmax
  call tc_check(tc_check_divergence(ctx, qmax), 'CheckDivergence')
end subroutine
subroutine CalcVelocityStats
  use param
  use tc_abi
  real :: rms(3)
  call tc_check(tc_calc_velocity_stats(ctx, vmax, rms), 'CalcVelocityStats')
  write(95, '(4(1x,e14.6))') time, vmax(1), vmax(2), vmax(3)
end subroutine
subroutine CalcTorque
  use param
  use tc_abi
  real :: anusslow, anussup
  call tc_check(tc_calc_torque(ctx, anusslow, anussup), 'CalcTorque')
  write(97, '(3(1x,e14.6))') time, anusslow, anussup
  write(96, '(3(1x,es24.16))') time, anusslow, anussup
end subroutine

! StatRoutines.F90:12-71 (with CalcVorticity) on the device
subroutine CalcStats
  use param
  use tc_abi
  call tc_check(tc_calc_stats(ctx), 'CalcStats')
end subroutine
! CalcDissipation.F90:11-153 on the device; dissipnu.out as :150
subroutine CalcDissipation
  use param
  use tc_abi
  real :: nute
  call tc_check(tc_calc_dissipation(ctx, nute), 'CalcDissipation')
  write(92, *) time, nute
end subroutine
! StatRoutines.F90:73-171: one raw (nrm, nzm) file per accumulator under
! stats/, added to the existing files when STAREA /= 0
! (StatReadReduceWrite.F90:27-31), and the sample count in
! stats/averaging_time.txt
subroutine WriteStats
  use param
  use tc_abi
  implicit none
  character(len=9), parameter :: names(18) = [character(len=9) :: 'vt_mean', 'vr_mean', &
      'vz_mean', 'wt_mean', 'wr_mean', 'wz_mean', 'vt_rms', 'vr_rms', 'vz_rms', 'wt_rms', &
      'wr_rms', 'wz_rms', 'vtvr_mean', 'vtvz_mean', 'vrvz_mean', 'pr_mean', 'pr_rms', 'disste']
  real, allocatable :: a(:,:,:), old(:,:)
  integer(c_int) :: ns
  integer :: q, u, ios, nold
  allocate(a(nrm, nzm, 18), old(nrm, nzm))
  call tc_check(tc_get_stats(ctx, a, ns), 'WriteStats')
  nold = 0
  if (bou%starea /= 0) then
    open(newunit=u, file='stats/averaging_time.txt', status='old', iostat=ios)
    if (ios == 0) then
      read(u, *) nold
      close(u)
    end if
  end if
  do q = 1, 18
    if (bou%starea /= 0) then
      open(newunit=u, file='stats/'//trim(names(q))//'.bin', access='stream', &
           form='unformatted', status='old', iostat=ios)
      if (ios == 0) then
        read(u) old
        close(u)
        a(:, :, q) = a(:, :, q) + old
      end if
    end if
    open(newunit=u, file='stats/'//trim(names(q))//'.bin', access='stream', &
         form='unformatted', status='replace', iostat=ios)
    if (ios /= 0) then
      write(6, *) 'WriteStats: cannot write stats/ (create the directory first)'
      return
    end if
    write(u) a(:, :, q)
    close(u)
  end do
  open(newunit=u, file='stats/averaging_time.txt', status='replace')
  write(u, *) ns + nold
  close(u)
end subroutine

! ---- slab dumps (raw FP64 files in place of the reference's HDF5) ----

! time stamp of the dump files, DumpRCuts.F90:17-19
function dump_stamp() result(ipfi)
  use param
  implicit none
  character(len=8) :: ipfi
  write(ipfi, '(i8.8)') nint(1000. * time + 0.005)
end function

subroutine write_raw(name, a, n)
  implicit none
  character(len=*), intent(in) :: name
  integer, intent(in) :: n
  real, intent(in) :: a(n)
  integer :: u, ios
  open(newunit=u, file=name, access='stream', form='unformatted', status='replace', iostat=ios)
  if (ios /= 0) then
    write(6, *) 'cannot write ', trim(name), ' (create the directory first)'
    return
  end if
  write(u) a
  close(u)
end subroutine

! InitDumpRCuts (DumpRCuts.F90:61-95): at most 9 radii from rcuts.in; each
! takes the plane k = 2..nrm with rm(k-1) < r < rm(k) (1 if none)
subroutine InitDumpRCuts
  use param
  use tc_abi
  implicit none
  character(len=4) :: dummy
  real :: rslab(9)
  real, allocatable :: rm(:)
  integer :: u, k, j, m
  if (nslab > 9) nslab = 9
  open(newunit=u, file='rcuts.in', status='old')
  read(u, '(a4)') dummy
  read(u, *) (rslab(j), j = 1, nslab)
  close(u)
  allocate(rm(nr))
  m = tc_grid_vector(bou, TC_V_RM, rm, nr)
  call tc_check(m, 'tc_grid_vector')
  kslab = 1
  do k = 2, nrm
    do j = 1, nslab
      if (rm(k) > rslab(j) .and. rm(k - 1) < rslab(j)) kslab(j) = k
    end do
  end do
  open(newunit=u, file='rcuts/rcutslocs.out', status='replace', iostat=k)
  if (k == 0) then
    write(u, *) (kslab(j), j = 1, nslab)
    close(u)
  end if
end subroutine

! DumpRCuts.F90:11-57
subroutine DumpRCuts
  use param
  use tc_abi
  implicit none
  interface
    function dump_stamp() result(ipfi)
      character(len=8) :: ipfi
    end function
  end interface
  real, allocatable :: vt(:), vr(:), vz(:)
  character(len=8) :: ipfi
  character(len=1) :: charm
  integer :: m
  allocate(vt(nzm * nthm), vr(nzm * nthm), vz(nzm * nthm))
  ipfi = dump_stamp()
  do m = 1, nslab
    call tc_check(tc_dump_rcut(ctx, int(kslab(m), c_int), vt, vr, vz), 'DumpRCuts')
    write(charm, '(i1.1)') m
    call write_raw('rcuts/slab'//charm//'vt_'//ipfi//'.bin', vt, nzm * nthm)
    call write_raw('rcuts/slab'//charm//'vr_'//ipfi//'.bin', vr, nzm * nthm)
    call write_raw('rcuts/slab'//charm//'vz_'//ipfi//'.bin', vz, nzm * nthm)
  end do
end subroutine

! DumpThCuts.F90:11-71 (the reference names the vr file slabvt_ as well;
! here it is slabvr_)
subroutine DumpThCuts
  use param
  use tc_abi
  implicit none
  interface
    function dump_stamp() result(ipfi)
      character(len=8) :: ipfi
    end function
  end interface
  real, allocatable :: vt(:), vr(:), vz(:)
  character(len=8) :: ipfi
  allocate(vt(nrm * nzm), vr(nrm * nzm), vz(nrm * nzm))
  call tc_check(tc_dump_thcut(ctx, vt, vr, vz), 'DumpThCuts')
  ipfi = dump_stamp()
  call write_raw('thcut/slabvt_'//ipfi//'.bin', vt, nrm * nzm)
  call write_raw('thcut/slabvr_'//ipfi//'.bin', vr, nrm * nzm)
  call write_raw('thcut/slabvz_'//ipfi//'.bin', vz, nrm * nzm)
end subroutine

! DumpStreamwiseAvg.F90:11-71
subroutine DumpStreamwiseAvg
  use param
  use tc_abi
  implicit none
  interface
    function dump_stamp() result(ipfi)
      character(len=8) :: ipfi
    end function
  end interface
  real, allocatable :: vt(:), vr(:), vz(:)
  character(len=8) :: ipfi
  allocate(vt(nrm * nzm), vr(nrm * nzm), vz(nrm * nzm))
  call tc_check(tc_dump_streamwise_avg(ctx, vt, vr, vz), 'DumpStreamwiseAvg')
  ipfi = dump_stamp()
  call write_raw('stave/slabvt_ax_'//ipfi//'.bin', vt, nrm * nzm)
  call write_raw('stave/slabvr_ax_'//ipfi//'.bin', vr, nrm * nzm)
  call write_raw('stave/slabvz_ax_'//ipfi//'.bin', vz, nrm * nzm)
end subroutine

! DumpFlowField.F90:12-50: qt, qr, qz as (nr, nzm, nthm) and the master
! scalars as text
subroutine DumpFlowField
  use param
  use tc_abi
  implicit none
  interface
    function dump_stamp() result(ipfi)
      character(len=8) :: ipfi
    end function
  end interface
  real, allocatable :: a(:)
  character(len=2), parameter :: names(3) = ['qt', 'qr', 'qz']
  integer(c_int), parameter :: ids(3) = [TC_QT, TC_QR, TC_QZ]
  character(len=8) :: ipfi
  integer :: q, u, ios
  allocate(a(nr * nzm * nthm))
  ipfi = dump_stamp()
  do q = 1, 3
    call tc_check(tc_download_field(ctx, ids(q), a, TC_LAYOUT_FORTRAN), 'DumpFlowField')
    call write_raw('fullff/'//names(q)//'_'//ipfi//'.bin', a, nr * nzm * nthm)
  end do
  open(newunit=u, file='fullff/ms_'//ipfi//'.txt', status='replace', iostat=ios)
  if (ios == 0) then
    write(u, *) time, bou%rint, bou%rext, bou%istr, bou%str, bou%alx3, bou%lamb
    close(u)
  end if
end subroutine

! Whole-state checkpoint of the library (fields, pressure, RK histories,
! time, step index): the run continues bit for bit from step ntime0
subroutine RestartFromCheckpoint
  use param
  use tc_abi
  implicit none
  integer(c_int) :: n
  real(c_double) :: s(8)
  call tc_check(tc_read_restart(ctx, trim(restart_file)//c_null_char, n), 'restart')
  call tc_check(tc_get_scalars(ctx, s), 'tc_get_scalars')
  ntime0 = n
  time = s(1)
  write(6, '(a,a,a,i0,a,es12.5)') 'restart from ', trim(restart_file), ' at step ', n, ' time ', time
end subroutine

! TimeMarcher.F90:12-57; the halo updates have no device counterpart
subroutine TimeMarcher
  use param
  use tc_abi
  implicit none
  integer :: ns
  call tc_check(tc_set_time(ctx, time, dt), 'tc_set_time')
  if (use_fused) then
    call tc_check(tc_time_marcher(ctx), 'TimeMarcher')
    return
  end if
  ! beta = dt/ren*0.5, InnerCylVelocity, SlipBCs
  call tc_check(tc_begin_step(ctx), 'tc_begin_step')
  do ns = 1, nsst
    call tc_check(tc_set_substep(ctx, ns), 'tc_set_substep')
    call ExplicitTermsQT
    call ExplicitTermsQR
    call ExplicitTermsQZ
    call ImplicitAndUpdateQT
    call ImplicitAndUpdateQR
    call ImplicitAndUpdateQZ
    call CalcLocalDivergence
    call SolvePressureCorrection
    call CorrectVelocity
    call CorrectPressure
  end do
end subroutine TimeMarcher

! main.F90:18-263
program tc_host
  use param
  use tc_abi
  implicit none
  integer :: errorcode, narg, a
  character(len=64) :: arg
  real :: instcfl, dmax, t0, t1
  integer(8) :: c0, c1, crate

  narg = command_argument_count()
  a = 1
  do while (a <= narg)
    call get_command_argument(a, arg)
    if (trim(arg) == '--fused') then
      use_fused = .true.
    else if (trim(arg) == '--restart') then
      call get_command_argument(a + 1, restart_file)
      a = a + 1
    else if (trim(arg) == '--checkpoint') then
      call get_command_argument(a + 1, checkpoint_file)
      a = a + 1
    else if (trim(arg) == '--ic') then
      call get_command_argument(a + 1, arg)
      if (trim(arg) /= 'perturbed') error stop 'usage: --ic perturbed SEED AMP'
      call get_command_argument(a + 2, arg); read(arg, *) ic_seed
      call get_command_argument(a + 3, arg); read(arg, *) ic_amp
      perturbed_ic = .true.
      a = a + 3
    else
      write(6, *) 'unknown argument ', trim(arg)
      error stop 2
    end if
    a = a + 1
  end do

  call ReadInputFile
  call tc_check(tc_create(bou, ctx), 'tc_create')
  nslab = bou%nslabr
  if (nslab /= 0) call InitDumpRCuts  ! main.F90:82
  call system_clock(c0, crate)
  open(97, file='torquenu.out', status='replace')
  open(96, file='torquenu_full.out', status='replace')
  open(95, file='vmax.out', status='replace')
  open(92, file='dissipnu.out', status='replace')

  time = 0.d0
  if (len_trim(restart_file) > 0) then
    call RestartFromCheckpoint
  else if (perturbed_ic) then
    call PerturbedCouetteStart
  else if (readflow) then
    call ReadFlowField
  else
    call CreateInitialConditions
  end if
  call CheckDivergence(dmax)
  write(6, *) ' initial divg dmax  ', dmax

  errorcode = 0
  do ntime = ntime0, ntst
    call CalcMaxCFL(instcfl)
    if (variabletstep) then
      if (ntime > 1) then
        if (instcfl < 1.0d-8) then
          dt = dtmax
        else
          dt = limitcfl / instcfl
        end if
        if (dt > dtmax) dt = dtmax
      else
        dt = dtmin
      end if
      if (dt < dtmin) errorcode = 166
    else
      ! fixed time step: DTMAX (main.F90:63-65 states this intent)
      dt = dtmax
      instcfl = instcfl * dt
      if (instcfl > limitcfl) errorcode = 165
    end if

    call TimeMarcher

    time = time + dt

    if (ntime == 1 .or. mod(time, tout) < dt) then
      call CalcVelocityStats
      if (vmax(1) > limitvel .and. vmax(2) > limitvel) errorcode = 266
      call CalcMaxCFL(instcfl)
      call CheckDivergence(dmax)
      call CalcTorque
      if (.not. variabletstep) instcfl = instcfl * dt  ! main.F90:201
      if (abs(dmax) > resid) errorcode = 169
      if (time > bou%tsta) then  ! main.F90:188-199
        call CalcStats
        call CalcDissipation
        if (mod(time, bou%tframe2d) < dt) then
          if (nslab /= 0) call DumpRCuts
          if (bou%dpthcut /= 0) call DumpThCuts
          if (bou%dpstave /= 0) call DumpStreamwiseAvg
        end if
        if (mod(time, bou%tframe3d) < dt) then
          if (bou%dpfullf /= 0) call DumpFlowField
        end if
      end if
    end if

    if (time > tmax) errorcode = 333
    call system_clock(c1)
    if (dble(c1 - c0) / dble(crate) > walltimemax) errorcode = 334
    if (ntime == ntst) errorcode = 555
    if (errorcode /= 0) exit
  end do

  call tc_check(tc_synchronize(ctx), 'tc_synchronize')
  call system_clock(c1)
  write(6, '(a,i0,a,i0,a,es12.5,a,f10.3,a)') 'exit code ', errorcode, ' after step ', ntime, &
      ' time ', time, ' wall ', dble(c1 - c0) / dble(crate), ' s'
  ! QuitRoutine.F90:28-30: statistics and fields on a normal exit
  if (errorcode == 333 .or. errorcode == 334 .or. errorcode == 555) call WriteStats
  call WriteFlowField
  if (len_trim(checkpoint_file) > 0) then
    ! the device holds the start time of the last step; store the time reached
    call tc_check(tc_set_time(ctx, time, dt), 'tc_set_time')
    call tc_check(tc_write_restart(ctx, trim(checkpoint_file)//c_null_char, int(ntime + 1, c_int)), &
                  'checkpoint')
  end if
  close(92); close(95); close(96); close(97)
  call tc_check(tc_destroy(ctx), 'tc_destroy')
  if (errorcode /= 333 .and. errorcode /= 334 .and. errorcode /= 555 .and. errorcode /= 0) stop 3
end program tc_host
