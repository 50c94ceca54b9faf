!=============================================================================
! sac_vac.f90 -- Fortran host of the MI355X perturbation-MHD solver.
!
! A drop-in for the serial reference program `vac` (src/vac.t:5-158) on its
! 'cd4' + Runge-Kutta + gravity/hyperdiffusion path: the same parameter
! input (par/DEFAULT, then the namelists of src/vacio.t:25-53 on standard
! input), the same ini files (ASCII and binary, src/vacio.t:591-752), the
! same output and log files (src/vacio.t:901-1149) and the same main-loop
! control (save -> dt -> finish -> advance).  The time step itself runs on
! the GPU through libsac_amd.so (include/sac_amd.h); the host owns w.
!
! Configurations outside the GPU path stop with a message, as the
! reference's die() does: other schemes or splittings, a resistivity given
! by the user's specialeta (eqpar(eta_)<0), smallfix, generalised
! coordinates, axial symmetry and the boundary type special.
!=============================================================================
module host_state
  use iso_c_binding
  implicit none
  integer, parameter :: dp = kind(1.d0)
  integer, parameter :: lentype = 10, lenname = 400
  integer, parameter :: nfile = 2, filelog_ = 1, fileout_ = 2, nsavehi = 100, nhiB = 10
  integer, parameter :: unitini = 10
  real(dp), parameter :: bigdouble = 1.d99, smalldouble = 1.d-99
  integer, parameter :: biginteger = 10000000
  ! user parameters of the gravity+hyperdiffusion module
  ! (sac/src/vacusrpar.t.stuart:3-5): eqpar(grav0_+d) is gravity along d,
  ! eqpar(nu_) switches hyperdiffusion on; neqpar = 2 (gamma, eta)
  integer, parameter :: neqpar = 2, grav0_ = neqpar + 1, nu_ = neqpar + 4, nspecialpar = 4

  ! problem size (from the ini file)
  integer :: ndim = 0, ndir = 0, nw = 0, nevolve = 0
  integer :: nx(3) = 1, ixGhi(3) = 1, ixMlo(3) = 1, ixMhi(3) = 1
  integer(c_long) :: ncell = 0

  ! namelist variables (names and meaning as in src/vacio.t:25-53)
  character(len=lenname) :: teststr, filenameini, filename(nfile), varnames, wnames
  character(len=lenname) :: fileheadini, fileheadout
  character(len=lentype) :: typefileini, typefileout, typefilelog
  character(len=lentype) :: typeadvance, typelimited, typefct, typetvd, typeaxial
  character(len=lentype) :: typepoisson, typeconstrain, typedimsplit, typesourcesplit
  character(len=lentype) :: typeimplinit, typeimpliter, typeimplmat
  character(len=lentype), allocatable :: typefull(:), typepred(:), typeimpl(:), typefilter(:)
  character(len=lentype), allocatable :: typelimiter(:), typeentropy(:), typeB(:,:), typeBscalar(:)
  integer :: ixtest1, ixtest2, ixtest3, iwtest, idimtest
  integer :: snapshotini, snapshotout
  integer :: dixBmin1 = 2, dixBmin2 = 2, dixBmin3 = 2, dixBmax1 = 2, dixBmax2 = 2, dixBmax3 = 2
  integer :: itsave(nsavehi, nfile), ditsave(nfile), itmax, it, acmwidth, nB
  integer :: slowsteps, implmrpcpar, implrestart, implrestart2, impliter, impliternr
  integer :: nproc(nfile + 2)
  integer, allocatable :: ixBmin(:,:), ixBmax(:,:), idimB(:), ipairB(:)
  real(dp) :: tsave(nsavehi, nfile), dtsave(nfile), tmax, dtmin, residmin, residmax, t
  real(dp) :: cputimemax, constraincoef, muscleta, musclomega, acmexpo
  real(dp) :: smallp, smallpcoeff, smallrho, smallrhocoeff, vacuumrho
  real(dp) :: courantpar, dtpar, dtdiffpar, implpar, impldiffpar, implerror, implrelax
  real(dp) :: impldwlimit, procpar(3)
  real(dp), allocatable :: eqpar(:), entropycoef(:), acmcoef(:)
  logical :: fullgridini, fullgridout, tmaxexact, useprimitive, acmnolim, fourthorder
  logical :: implmrpc, dimsplit, sourcesplit, sourceunsplit, divbfix, divbwave
  logical :: divbconstrain, angmomfix, compactres, smallfix, extraB, dtcantgrow
  logical :: implnewton, implconserv, implcentered, implnewmat, implpred, impl3level
  logical :: impljacfast, implsource
  logical, allocatable :: artcomp(:), upperB(:)

  namelist /testlist/ teststr, ixtest1, ixtest2, ixtest3, iwtest, idimtest
  namelist /filelist/ filenameini, filename, varnames, typefileini, typefileout, typefilelog, &
       snapshotini, snapshotout, fullgridini, fullgridout, &
       dixBmin1, dixBmin2, dixBmin3, dixBmax1, dixBmax2, dixBmax3
  namelist /savelist/ tsave, itsave, dtsave, ditsave
  namelist /stoplist/ itmax, tmax, tmaxexact, dtmin, residmin, residmax, t, it, cputimemax
  namelist /methodlist/ wnames, fileheadout, eqpar, typeadvance, typefull, typepred, typeimpl, &
       typefilter, typelimiter, typeentropy, entropycoef, artcomp, typelimited, typefct, &
       typetvd, typeaxial, typepoisson, typeconstrain, constraincoef, useprimitive, muscleta, &
       musclomega, acmwidth, acmnolim, acmcoef, acmexpo, fourthorder, implmrpc, dimsplit, &
       typedimsplit, sourcesplit, typesourcesplit, sourceunsplit, divbfix, divbwave, &
       divbconstrain, angmomfix, compactres, smallfix, smallp, smallpcoeff, smallrho, &
       smallrhocoeff, vacuumrho, nproc, procpar
  namelist /boundlist/ nB, typeB, ixBmin, ixBmax, idimB, upperB, extraB, typeBscalar, ipairB
  namelist /paramlist/ courantpar, dtpar, dtdiffpar, dtcantgrow, slowsteps, implmrpcpar, &
       implpar, impldiffpar, implerror, implrelax, impldwlimit, implrestart, implrestart2, &
       impliter, impliternr, typeimplinit, typeimpliter, typeimplmat, implnewton, &
       implconserv, implcentered, implnewmat, implpred, impl3level, impljacfast, implsource

  ! state and grid
  real(dp), allocatable :: w(:,:,:,:), x(:,:,:,:), dxg(:,:,:,:)
  real(dp) :: volume = 0.0_dp, dt = 0.0_dp
  ! save bookkeeping (src/vac.t:608-638)
  integer :: isavet(nfile), isaveit(nfile), itsavelast(nfile), isaveout = 0
  real(dp) :: tsavelast(nfile)
  integer :: itmin = 0
  logical :: verbose = .true.

contains

  subroutine die(message)
    character(*), intent(in) :: message
    interface
      subroutine c_exit(code) bind(C, name='exit')
        import :: c_int
        integer(c_int), value :: code
      end subroutine c_exit
    end interface
    write(*, '(a)') trim(message)
    flush(6)
    call c_exit(1_c_int)
  end subroutine die

  double precision function cputime()
    integer(8) :: clock, rate
    call system_clock(clock, rate)
    cputime = real(clock, dp) / real(rate, dp)
  end function cputime

end module host_state

!=============================================================================
module host_io
  use host_state
  implicit none
contains

  ! -- ini-file header peek: ndim and nw decide the problem size --------
  subroutine peek_ini(fname, ftype, ndimini, nwini, nxini)
    character(*), intent(in) :: fname
    character(len=lentype), intent(inout) :: ftype
    integer, intent(out) :: ndimini, nwini, nxini(3)
    character(len=91) :: fhead
    character(len=lenname) :: head
    integer :: u, ios, itini, neq
    real(dp) :: tini
    logical :: ex
    inquire(file=fname, exist=ex)
    if (.not. ex) call die('Stop: file does not exist, filenameini=' // trim(fname))
    if (ftype == 'auto') then
      ! auto-detection of src/vacio.t:548-564
      open(newunit=u, file=fname, status='old', access='stream', form='unformatted', iostat=ios)
      fhead = ' '
      read(u, iostat=ios) fhead
      close(u)
      if (ichar(fhead(1:1)) /= 0 .and. ichar(fhead(2:2)) /= 0 .and. &
          ichar(fhead(3:3)) /= 0 .and. ichar(fhead(4:4)) /= 0) then
        ftype = 'ascii'
      else if (ichar(fhead(89:89)) == 0 .and. ichar(fhead(90:90)) == 0 .and. &
               (ichar(fhead(88:88)) == 24 .or. ichar(fhead(91:91)) == 24)) then
        ftype = 'binary'
      else
        ! record marker before a 400-byte header (LENNAME=400)
        ftype = 'binary'
      end if
    end if
    nxini = 1
    if (ftype == 'ascii') then
      open(newunit=u, file=fname, status='old')
      read(u, '(a)') head
      read(u, *) itini, tini, ndimini, neq, nwini
      read(u, *) nxini(1:abs(ndimini))
      close(u)
    else if (ftype == 'binary') then
      open(newunit=u, file=fname, status='old', form='unformatted')
      read(u) head
      read(u) itini, tini, ndimini, neq, nwini
      read(u) nxini(1:abs(ndimini))
      close(u)
    else
      call die('Error in VAC: Unknown typefileini=' // trim(ftype))
    end if
    if (ndimini < 0) call die('Error: input file contains general grid (gencoord is not built)')
  end subroutine peek_ini

  ! -- readfileini (src/vacio.t:533-752), last snapshot or snapshotini ----
  subroutine readfileini()
    integer :: u, ios, ndimini, neqparini, nwini, snapshot, i1, i2, i3, iw, idim, ieq
    integer :: lo(3), hi(3), neqin, nwin
    real(dp) :: extra
    character(len=79) :: vn79
    character(len=lenname) :: vn
    real(dp), allocatable :: xr(:), wr(:)
    lo = 1
    hi = 1
    do idim = 1, ndim
      if (fullgridini) then
        lo(idim) = 1
        hi(idim) = ixGhi(idim)
      else
        lo(idim) = ixMlo(idim)
        hi(idim) = ixMhi(idim)
      end if
    end do
    snapshot = 0
    if (typefileini == 'ascii') then
      open(newunit=u, file=filenameini, status='old')
      allocate(xr(ndim), wr(nw))
      do
        read(u, '(a)', iostat=ios) fileheadini
        if (ios /= 0) exit
        read(u, *, iostat=ios) it, t, ndimini, neqparini, nwini
        read(u, *, iostat=ios) nx(1:ndim)
        neqin = min(neqparini, neqpar + nspecialpar)
        nwin = min(nwini, nw)
        read(u, *, iostat=ios) (eqpar(ieq), ieq = 1, neqin), (extra, ieq = neqin + 1, neqparini)
        read(u, '(a)', iostat=ios) vn
        if (varnames == 'default') varnames = vn
        do i3 = lo(3), hi(3)
          do i2 = lo(2), hi(2)
            do i1 = lo(1), hi(1)
              read(u, *, iostat=ios) xr(1:ndim), wr(1:nwin), (extra, iw = nwin + 1, nwini)
              x(i1, i2, i3, 1:ndim) = xr(1:ndim)
              w(i1, i2, i3, 1:nwin) = wr(1:nwin)
            end do
          end do
        end do
        if (ios /= 0) call die('Error in reading file')
        snapshot = snapshot + 1
        if (snapshot == snapshotini) exit
      end do
      close(u)
    else
      open(newunit=u, file=filenameini, status='old', form='unformatted')
      do
        read(u, iostat=ios) fileheadini
        if (ios /= 0) exit
        read(u, iostat=ios) it, t, ndimini, neqparini, nwini
        read(u, iostat=ios) nx(1:ndim)
        neqin = min(neqparini, neqpar + nspecialpar)
        nwin = min(nwini, nw)
        read(u, iostat=ios) (eqpar(ieq), ieq = 1, neqin), (extra, ieq = neqin + 1, neqparini)
        read(u, iostat=ios) vn79
        if (varnames == 'default') varnames = vn79
        read(u, iostat=ios) ((((x(i1, i2, i3, idim), i1 = lo(1), hi(1)), i2 = lo(2), hi(2)), &
             i3 = lo(3), hi(3)), idim = 1, ndim)
        do iw = 1, nwin
          read(u, iostat=ios) (((w(i1, i2, i3, iw), i1 = lo(1), hi(1)), i2 = lo(2), hi(2)), &
               i3 = lo(3), hi(3))
        end do
        if (ios /= 0) call die('Error in ReadFileIni: iostat')
        snapshot = snapshot + 1
        if (snapshot == snapshotini) exit
      end do
      close(u)
    end if
    if (snapshot == 0) call die('Error in reading file: no snapshot in ' // trim(filenameini))
  end subroutine readfileini

  ! -- savefile (src/vacio.t:901-958) -----------------------------------
  subroutine savefile(ifile)
    integer, intent(in) :: ifile
    integer :: lo(3), hi(3), d
    character(len=10) :: itstring
    character(len=lenname) :: fname
    logical :: op
    lo = 1
    hi = 1
    do d = 1, ndim
      lo(d) = ixMlo(d)
      hi(d) = ixMhi(d)
      if (fullgridout .and. ifile == fileout_) then
        lo(d) = 1
        hi(d) = ixGhi(d)
      end if
    end do
    select case (ifile)
    case (fileout_)
      fname = filename(fileout_)
      if (snapshotout > 0 .and. isaveout > 0) then
        if (isaveout == snapshotout * (isaveout / snapshotout)) then
          close(unitini + ifile)
          write(itstring, '(i10)') isaveout + 1
          fname = fname(1:index(fname, ' ') - 1) // '_' // &
               itstring(10 - int(log10(isaveout + 1.5)):10)
        end if
      end if
      isaveout = isaveout + 1
      inquire(unitini + ifile, opened=op)
      select case (typefileout)
      case ('ascii')
        if (.not. op) open(unitini + ifile, file=fname, status='unknown')
        call savefileout_asc(unitini + ifile, lo, hi)
      case ('binary')
        if (.not. op) open(unitini + ifile, file=fname, status='unknown', form='unformatted')
        call savefileout_bin(unitini + ifile, lo, hi)
      case default
        call die('Error in SaveFile: Unknown typefileout:' // trim(typefileout))
      end select
    case (filelog_)
      if (typefilelog /= 'default') call die('Error in SaveFile: Unknown typefilelog:' // &
           trim(typefilelog))
      call savefilelog_default(unitini + ifile)
    end select
  end subroutine savefile

  subroutine savefileout_asc(u, lo, hi)
    integer, intent(in) :: u, lo(3), hi(3)
    integer :: i1, i2, i3
    real(dp) :: qw(nw)
    write(u, '(a)') fileheadout
    write(u, '(i7,1pe13.5,3i3)') it, t, ndim, neqpar + nspecialpar, nw
    write(u, '(3i4)') hi(1:ndim) - lo(1:ndim) + 1
    write(u, '(100(1pe13.5))') eqpar
    write(u, '(a)') varnames
    do i3 = lo(3), hi(3)
      do i2 = lo(2), hi(2)
        do i1 = lo(1), hi(1)
          ! magnitudes up to 5e-16 are written as zero (src/vacio.t:990-995)
          where (abs(w(i1, i2, i3, 1:nw)) > 5.0d-16)
            qw = w(i1, i2, i3, 1:nw)
          elsewhere
            qw = 0.0_dp
          end where
          write(u, '(100(1pe18.10))') x(i1, i2, i3, 1:ndim), qw
        end do
      end do
    end do
    flush(u)
  end subroutine savefileout_asc

  subroutine savefileout_bin(u, lo, hi)
    integer, intent(in) :: u, lo(3), hi(3)
    integer :: iw, idim
    write(u) fileheadout
    write(u) it, t, ndim, neqpar + nspecialpar, nw
    write(u) hi(1:ndim) - lo(1:ndim) + 1
    write(u) eqpar
    write(u) varnames
    write(u) (x(lo(1):hi(1), lo(2):hi(2), lo(3):hi(3), idim), idim = 1, ndim)
    do iw = 1, nw
      write(u) w(lo(1):hi(1), lo(2):hi(2), lo(3):hi(3), iw)
    end do
    flush(u)
  end subroutine savefileout_bin

  subroutine savefilelog_default(u)
    integer, intent(in) :: u
    logical :: op
    integer :: iw
    real(dp) :: wmean(nw)
    inquire(u, opened=op)
    if (.not. op) then
      open(u, file=filename(filelog_), status='unknown')
      write(u, '(a)') fileheadout
      if (residmin > 0.0_dp .or. residmax < bigdouble) then
        write(u, '(a15,a55,a9)') 'it   t   dt    ', wnames, ' residual'
      else
        write(u, '(a15,a64)') 'it   t   dt    ', wnames
      end if
    end if
    do iw = 1, nw
      wmean(iw) = sum(dxg(ixMlo(1):ixMhi(1), ixMlo(2):ixMhi(2), ixMlo(3):ixMhi(3), 1) * &
           w(ixMlo(1):ixMhi(1), ixMlo(2):ixMhi(2), ixMlo(3):ixMhi(3), iw)) / volume
    end do
    write(u, '(i7,100(1pe13.5))') it, t, dt, wmean
    flush(u)
  end subroutine savefilelog_default

end module host_io

!=============================================================================
program sac_vac
  use iso_c_binding
  use host_state
  use host_io
  use sac_bind
  implicit none

  integer :: us, ios, ifile, i, iw, iB, idim, ndimini, nwini, nxini(3), d, j, k
  integer :: i1, i2, i3, adv, rc
  character(len=lenname) :: line
  character(len=lentype) :: typefull1, typepred1, ftype
  logical :: ex
  real(dp) :: time0, time1, dtold, dtc, cmax
  real(dp), allocatable :: qx(:), dxaxis(:)
  integer(c_int), allocatable :: iws(:), bcodes(:)
  integer(c_int) :: nxc(3), ready
  type(c_ptr) :: h
  ! asynchronous saves: the snapshot of the files due at one iteration is
  ! copied off the device while the next steps run and written once it has
  ! arrived, with the it, t and dt of the iteration that asked for it
  ! (SAC_SYNC_SAVES=1 in the environment downloads and writes at once)
  logical :: asyncsave, savepending, savedue(nfile)
  integer :: itsnap
  real(dp) :: tsnap, dtsnap
  character(len=8) :: envval

  time0 = cputime()

  ! ---- standard input to a scratch file (namelists are read twice) ----
  open(newunit=us, status='scratch', form='formatted', action='readwrite')
  do
    read(*, '(a)', iostat=ios) line
    if (ios /= 0) exit
    write(us, '(a)') trim(line)
  end do

  ! ---- size of the problem from the ini header ----
  call set_scalar_defaults()
  rewind(us)
  read(us, testlist, iostat=ios)
  rewind(us)
  read(us, filelist, iostat=ios)
  if (ios /= 0) call die('Error reading &filelist')
  ftype = typefileini
  call peek_ini(filenameini, ftype, ndimini, nwini, nxini)
  ndim = ndimini
  if (ndim < 1 .or. ndim > 3) call die('Incompatible dimensionalities')
  if (mod(nwini - 4, 3) /= 0 .or. nwini < 7) call die('ini file nw does not describe MHD')
  ndir = (nwini - 4) / 3
  nw = nwini
  nevolve = 2 + 2 * ndir
  call allocate_arrays()
  call set_array_defaults()

  ! ---- par/DEFAULT, then standard input (src/vacio.t:119-530) ----
  inquire(file='par/DEFAULT', exist=ex)
  if (ex) then
    open(unitini - 1, file='par/DEFAULT', status='old')
    read(unitini - 1, testlist)
    read(unitini - 1, filelist)
    read(unitini - 1, savelist)
    read(unitini - 1, stoplist)
    read(unitini - 1, methodlist)
    read(unitini - 1, boundlist)
    read(unitini - 1, paramlist)
    close(unitini - 1)
  end if
  eqpar = 0.0_dp
  rewind(us)
  read(us, testlist)
  varnames = 'default'
  read(us, filelist)
  if (typefileini == 'auto') typefileini = ftype
  if (typefileout == 'auto') typefileout = typefileini
  call setgrid(nxini)
  call readfileini()
  fileheadout = fileheadini
  read(us, savelist)
  read(us, stoplist)
  read(us, methodlist)
  read(us, boundlist, iostat=ios)
  read(us, paramlist, iostat=ios)
  close(us)
  call harmonize_and_check()

  ! ---- grid: gridsetup1 (src/vacgrid.t:578-661) ----
  call gridsetup1()

  ! ---- device set-up ----
  allocate(iws(nevolve), bcodes(nw * 2 * ndim))
  do iw = 1, nevolve
    iws(iw) = iw
  end do
  do iw = 1, nw
    do iB = 1, 2 * ndim
      bcodes((iw - 1) * 2 * ndim + iB) = bc_code(typeB(iw, iB))
    end do
  end do
  nxc = 1
  nxc(1:ndim) = nx(1:ndim)
  select case (typeadvance)
  case ('onestep')
    adv = SAC_ADV_ONESTEP
  case ('twostep')
    adv = SAC_ADV_TWOSTEP
  case ('threestep')
    adv = SAC_ADV_THREESTEP
  case ('sterck')
    adv = SAC_ADV_STERCK
  case ('jameson')
    adv = SAC_ADV_JAMESON
  case default
    adv = SAC_ADV_FOURSTEP
  end select
  rc = sac_dev_init(int(ndim, c_int), int(ndir, c_int), nxc, int(nw, c_int), iws, &
       int(nevolve, c_int), eqpar, int(neqpar + nspecialpar, c_int), int(grav0_, c_int), &
       int(nu_, c_int), dxaxis, bcodes, merge(1_c_int, 0_c_int, fullgridini), adv, &
       SAC_PREC_FP64, h)
  call check(rc)
  ! Powell's div B sources in addsource (src/vacphys.mhd0.t:256, ndim > 1)
  call check(sac_dev_set_divbfix(h, merge(1_c_int, 0_c_int, divbfix)))
  ! resistive sources (src/vacphys.mhd0.t:240-249) and the common dtdiffpar
  call check(sac_dev_set_resistive(h, merge(1_c_int, 0_c_int, compactres), &
       real(dtdiffpar, c_double)))
  call check(sac_dev_upload(h, w))

  ! startup (src/vac.t:161-221)
  itmin = it
  do ifile = 1, nfile
    tsavelast(ifile) = t
    itsavelast(ifile) = it
  end do
  isaveout = 0

  if (verbose) write(*, '(a,f10.3,a)') 'Start Advance  ', cputime() - time0, ' sec'
  call check(sac_dev_getboundary(h))
  call check(sac_dev_download(h, w))
  dtold = 0.0_dp
  call get_environment_variable('SAC_SYNC_SAVES', envval)
  asyncsave = trim(envval) /= '1'
  savepending = .false.
  ! main loop (src/vac.t:55-106)
  do
    do ifile = 1, nfile
      savedue(ifile) = timetosave(ifile)
    end do
    if (any(savedue)) then
      if (asyncsave) then
        call finish_save()
        call check(sac_dev_save_begin(h))
        savepending = .true.
        itsnap = it
        tsnap = t
        dtsnap = dt
      else
        call check(sac_dev_download(h, w))
        do ifile = 1, nfile
          if (savedue(ifile)) call savefile(ifile)
        end do
      end if
    end if
    if (dtpar > 0.0_dp) then
      dt = dtpar
    else
      if (courantpar > 0.0_dp) then
        call check(sac_dev_getdt_courant(h, courantpar, dtc, cmax))
        dt = dtc
      end if
      ! getdt (the resistive limit when eta > 0), then getdt_special (viscous
      ! and gravity limits), in one call
      call check(sac_dev_getdt_special(h, dt))
      if (dtcantgrow .and. it > itmin) dt = min(dt, dtold)
      dtold = dt
    end if
    if (timetofinish()) exit
    if (slowsteps == 1 .and. it == itmin) dtpar = -1.0_dp
    if (slowsteps > it - itmin + 1) dt = dt * (1.0_dp - (1.0_dp - (it - itmin + 1.0_dp) / &
         slowsteps)**2)
    if (tmaxexact) dt = min(dt, tmax - t + smalldouble)
    call check(sac_dev_advance(h, dt))
    it = it + 1
    t = t + dt
    if (savepending) then
      call check(sac_dev_save_ready(h, ready))
      if (ready == 1) call finish_save()
    end if
  end do
  call finish_save()
  time1 = cputime() - time0
  call check(sac_dev_download(h, w))
  do ifile = 1, nfile
    if (itsavelast(ifile) < it) call savefile(ifile)
    close(unitini + ifile)
  end do
  if (verbose) write(*, '(a,f10.3,a)') 'Finish Advance ', time1, ' sec'
  if (dt < dtmin) write(*, *) 'Warning: dt<dtmin !!!'
  if (time1 > cputimemax) write(*, *) 'Warning: cputimemax exceeded !!!'
  rc = sac_dev_free(h)
  if (verbose) write(*, '(a,f10.3,a)') 'Finished VAC   ', cputime() - time0, ' sec'

contains

  subroutine check(code)
    integer(c_int), intent(in) :: code
    if (code /= 0) call die('libsac_amd: ' // trim(sac_error_message()))
  end subroutine check

  ! write the files of the pending asynchronous save (waits for its copy)
  subroutine finish_save()
    integer :: itnow
    real(dp) :: tnow, dtnow
    if (.not. savepending) return
    savepending = .false.
    call check(sac_dev_save_end(h, w))
    itnow = it
    tnow = t
    dtnow = dt
    it = itsnap
    t = tsnap
    dt = dtsnap
    do ifile = 1, nfile
      if (savedue(ifile)) call savefile(ifile)
    end do
    it = itnow
    t = tnow
    dt = dtnow
  end subroutine finish_save

  integer(c_int) function bc_code(s)
    character(len=*), intent(in) :: s
    select case (trim(s))
    case ('cont')
      bc_code = SAC_BC_CONT
    case ('fixed')
      bc_code = SAC_BC_FIXED
    case ('periodic')
      bc_code = SAC_BC_PERIODIC
    case ('symm')
      bc_code = SAC_BC_SYMM
    case ('symm0')
      bc_code = SAC_BC_SYMM0
    case ('asymm')
      bc_code = SAC_BC_ASYMM
    case ('cont1')
      bc_code = SAC_BC_CONT1
    case ('fixed1')
      bc_code = SAC_BC_FIXED1
    case ('grad1')
      bc_code = SAC_BC_GRAD1
    case ('contCD4')
      bc_code = SAC_BC_CONTCD4
    case ('zero')
      bc_code = SAC_BC_ZERO
    case default
      call die('Error in GetBoundary: boundary type ' // trim(s) // ' is not implemented!')
      bc_code = 0
    end select
  end function bc_code

  subroutine set_scalar_defaults()
    teststr = ' '
    ixtest1 = 1; ixtest2 = 1; ixtest3 = 1; iwtest = 1; idimtest = 1
    filenameini = ' '; filename = ' '; varnames = 'default'; wnames = ' '
    typefileini = 'auto'; typefileout = 'auto'; typefilelog = 'default'
    snapshotini = 0; snapshotout = 0; fullgridini = .false.; fullgridout = .false.
    tsave = bigdouble; itsave = biginteger; dtsave = bigdouble; ditsave = biginteger
    isavet = 1; isaveit = 1
    itmax = 1000000; tmax = 1.d99; tmaxexact = .false.; dtmin = 1.d-99
    residmin = 0.0_dp; residmax = 1.d99; t = 0.0_dp; it = 0; cputimemax = bigdouble
    typeadvance = 'twostep'; typefct = 'etbfct'; typetvd = 'roe'; typeaxial = 'slab'
    typelimited = 'original'; useprimitive = .false.; acmwidth = 1; acmnolim = .false.
    acmexpo = 0.0_dp; fourthorder = .false.; dimsplit = .true.; typedimsplit = 'default'
    sourcesplit = .false.; sourceunsplit = .true.; typesourcesplit = 'default'
    divbfix = .true.; divbwave = .true.; typepoisson = 'default'; divbconstrain = .false.
    typeconstrain = 'EfieldCD'; angmomfix = .false.; muscleta = 0.0_dp; musclomega = 0.0_dp
    smallfix = .false.; smallpcoeff = 1.d-6; smallrhocoeff = 1.d-6; compactres = .false.
    implmrpc = .false.; constraincoef = 1.0_dp
    smallp = -1.0_dp; smallrho = -1.0_dp; vacuumrho = -1.0_dp
    nproc = 0; procpar = -1.0_dp
    extraB = .false.; nB = 0
    courantpar = 0.8_dp; dtcantgrow = .false.; dtdiffpar = 0.5_dp; dtpar = -1.0_dp
    slowsteps = 0; implmrpcpar = 4; implrelax = -0.5_dp; impldiffpar = 1.d-12
    impldwlimit = 1.d99; impliternr = 20; impliter = 100; implrestart = 20; implrestart2 = 5
    typeimpliter = 'vac_bicg'; impljacfast = .false.; implnewton = .false.
    implnewmat = .false.; implpred = .false.; implconserv = .false.; implcentered = .false.
    implpar = -1.0_dp; implerror = 1.d-3; impl3level = .false.; implsource = .false.
    typeimplinit = 'unused'; typeimplmat = 'unused'
  end subroutine set_scalar_defaults

  subroutine allocate_arrays()
    allocate(typefull(nw), typepred(nw), typeimpl(nw), typefilter(nw), typelimiter(nw))
    allocate(typeentropy(nw), typeB(nw, nhiB), typeBscalar(nhiB), artcomp(nw))
    allocate(entropycoef(nw), acmcoef(nw), eqpar(neqpar + nspecialpar))
    allocate(ixBmin(ndim, nhiB), ixBmax(ndim, nhiB), idimB(nhiB), ipairB(nhiB), upperB(nhiB))
  end subroutine allocate_arrays

  subroutine set_array_defaults()
    typepred = 'default'; typefull = 'tvdlf'; typeimpl = 'nul'; typefilter = 'nul'
    typelimiter = 'minmod'; typeentropy = 'nul'; artcomp = .false.
    acmcoef = -1.0_dp; entropycoef = -1.0_dp; eqpar = 0.0_dp
    typeB = 'cont'; typeBscalar = ' '; ipairB = 0; idimB = 0; upperB = .false.
    ixBmin = 0; ixBmax = 0
    nB = 2 * ndim
  end subroutine set_array_defaults

  ! setixGixMix (src/vacio.t:786-841) with dixB = 2
  subroutine setgrid(nxi)
    integer, intent(in) :: nxi(3)
    integer :: dd
    nx = 1
    ixGhi = 1
    ixMlo = 1
    ixMhi = 1
    do dd = 1, ndim
      nx(dd) = nxi(dd)
      if (fullgridini) nx(dd) = nx(dd) - 4
      ixMlo(dd) = 3
      ixMhi(dd) = nx(dd) + 2
      ixGhi(dd) = nx(dd) + 4
    end do
    allocate(w(ixGhi(1), ixGhi(2), ixGhi(3), nw), x(ixGhi(1), ixGhi(2), ixGhi(3), ndim))
    w = 0.0_dp
    x = 0.0_dp
    ncell = int(ixGhi(1), c_long) * ixGhi(2) * ixGhi(3)
  end subroutine setgrid

  ! method lists and checks of src/vacio.t:249-530, limited to the GPU path
  subroutine harmonize_and_check()
    integer :: jw, nfull
    typefull1 = 'nul'
    typepred1 = 'nul'
    nfull = 0
    do jw = 1, nw
      if (typefull(jw) /= 'nul') then
        typefull1 = typefull(jw)
        nfull = nfull + 1
        if (jw /= nfull) call die('typefull: only the evolving variables may use cd4')
      end if
      if (typepred(jw) /= 'nul') typepred1 = typepred(jw)
      if (typeimpl(jw) /= 'nul') call die('implicit methods are not built (typeimpl)')
      if (typefilter(jw) /= 'nul') call die('filters are not built (typefilter)')
    end do
    if (nfull /= nevolve) call die('typefull: cd4 on all evolving variables is required')
    if (typefull1 /= 'cd4') call die('Error in Advect1:' // trim(typefull1) // &
         ' is unknown or switched off!')
    if (typepred1 == 'default') typepred1 = 'nul'
    if (typedimsplit == 'default' .and. dimsplit) typedimsplit = 'xyyx'
    if (typedimsplit == 'default' .and. .not. dimsplit) typedimsplit = 'unsplit'
    if (typesourcesplit == 'default' .and. sourcesplit) typesourcesplit = 'sfs'
    if (typesourcesplit == 'default' .and. .not. sourcesplit) typesourcesplit = 'unsplit'
    dimsplit = typedimsplit /= 'unsplit'
    sourcesplit = typesourcesplit /= 'unsplit'
    if (sourcesplit) sourceunsplit = .false.
    if (typeadvance == 'twostep' .and. typepred1 == 'nul') typeadvance = 'onestep'
    if (typeadvance == 'twostep' .and. typepred1 /= 'cd4') &
         call die('typepred=' // trim(typepred1) // ' is not built (twostep needs cd4 or nul)')
    select case (typeadvance)
    case ('onestep', 'twostep', 'threestep', 'fourstep', 'sterck', 'jameson')
    case default
      call die('typeadvance=' // trim(typeadvance) // ' is not built')
    end select
    if (dimsplit) call die('dimsplit=T is not built: set dimsplit=F')
    if (sourcesplit .or. .not. sourceunsplit) call die('only sourceunsplit=T is built')
    if (smallfix) call die('smallfix=T is not built')
    if (typeaxial /= 'slab') call die('typeaxial=' // trim(typeaxial) // ' is not built')
    if (extraB) call die('extraB=T (special boundaries) is not built')
    if (eqpar(2) < 0.0_dp) &
         call die('Resistive MHD with eqpar(eta_)<0 (specialeta) is not built')
    if (nB /= 2 * ndim) call die('only the default boundary regions are built (nB=2*ndim)')
    if (residmin > 0.0_dp .or. residmax < bigdouble) &
         call die('steady-state residuals (residmin/residmax) are not built')
    ! startup (src/vac.t:175): a fixed dtpar switches the courant limit off
    if (dtpar > 0.0_dp .and. slowsteps /= 1) courantpar = -1.0_dp
    call setheaderstrings()
  end subroutine harmonize_and_check

  ! setheaderstrings (src/vacio.t:844-898)
  subroutine setheaderstrings()
    character(len=lentype) :: physics
    integer :: ii
    character(len=*), parameter :: eqparname = 'gamma eta'
    character(len=17) :: specialparname
    write(physics, '(a,i1,i1)') 'mhd', ndim, ndir
    ii = index(fileheadout, '_')
    if (ii >= 1) then
      if (physics /= fileheadout(ii + 1:ii + lentype)) then
        write(*, *) 'This code is configured to ', physics
        call die('Error:  physics in file is ' // fileheadout(ii + 1:ii + lentype))
      end if
    else
      ii = lenname - 3 - 3
      do
        if (fileheadout(ii:ii) /= ' ' .or. ii == 1) exit
        ii = ii - 1
      end do
      fileheadout = fileheadout(1:ii) // '_' // physics
      write(*, *) 'Warning: incomplete input headline.', ' Added to output headline _', physics
    end if
    if (varnames /= 'default' .and. index(varnames, eqparname) <= 0) then
      ii = lenname - len(eqparname) - 3
      do
        if (varnames(ii:ii) /= ' ' .or. ii == 1) exit
        ii = ii - 1
      end do
      varnames = varnames(1:ii) // '   ' // eqparname
    end if
    select case (ndim)
    case (1)
      specialparname = 'grav1'
    case (2)
      specialparname = 'grav1 grav2'
    case default
      specialparname = 'grav1 grav2 grav3'
    end select
    if (varnames /= 'default' .and. index(varnames, trim(specialparname)) <= 0) then
      ii = lenname - len_trim(specialparname) - 3
      do
        if (varnames(ii:ii) /= ' ' .or. ii == 1) exit
        ii = ii - 1
      end do
      varnames = varnames(1:ii) // '   ' // trim(specialparname)
    end if
  end subroutine setheaderstrings

  ! gridsetup1: linear extrapolation of x into the ghost layers, then
  ! dx=(x(i+1)-x(i-1))/2; the grid must be a tensor product (dx along d
  ! depends on index d only) for the per-axis widths the device uses
  subroutine gridsetup1()
    integer :: dd, ii, n, lo, off, ixe
    real(dp) :: tol
    allocate(dxaxis(sum(ixGhi(1:ndim))))
    allocate(dxg(ixGhi(1), ixGhi(2), ixGhi(3), 1))
    off = 0
    do dd = 1, ndim
      n = ixGhi(dd)
      allocate(qx(0:n + 1))
      do ii = 1, n
        select case (dd)
        case (1)
          qx(ii) = x(ii, ixMlo(2), ixMlo(3), 1)
        case (2)
          qx(ii) = x(ixMlo(1), ii, ixMlo(3), 2)
        case (3)
          qx(ii) = x(ixMlo(1), ixMlo(2), ii, 3)
        end select
      end do
      ! upper layers
      lo = ixMhi(dd) + 1
      if (fullgridini) lo = n + 1
      ixe = lo - 1
      do ii = lo, n + 1
        qx(ii) = (1 + abs(ixe - ii)) * qx(ixe) - abs(ixe - ii) * qx(ixe - 1)
      end do
      ! lower layers
      lo = ixMlo(dd) - 1
      if (fullgridini) lo = 0
      ixe = lo + 1
      do ii = 0, lo
        qx(ii) = (1 + abs(ixe - ii)) * qx(ixe) - abs(ixe - ii) * qx(ixe + 1)
      end do
      do ii = 1, n
        dxaxis(off + ii) = 0.5_dp * (qx(ii + 1) - qx(ii - 1))
      end do
      ! tensor-grid check on the mesh
      tol = 1.d-12 * maxval(abs(dxaxis(off + 1:off + n)))
      do i3 = ixMlo(3), ixMhi(3)
        do i2 = ixMlo(2), ixMhi(2)
          do i1 = ixMlo(1), ixMhi(1)
            select case (dd)
            case (1)
              ii = i1
            case (2)
              ii = i2
            case (3)
              ii = i3
            end select
            if (ii > 1 .and. ii < n) then
              if (abs(0.5_dp * (x(min(i1 + merge(1, 0, dd == 1), ixGhi(1)), &
                   min(i2 + merge(1, 0, dd == 2), ixGhi(2)), &
                   min(i3 + merge(1, 0, dd == 3), ixGhi(3)), dd) - &
                   x(max(i1 - merge(1, 0, dd == 1), 1), max(i2 - merge(1, 0, dd == 2), 1), &
                   max(i3 - merge(1, 0, dd == 3), 1), dd)) - dxaxis(off + ii)) > tol .and. &
                   ii > ixMlo(dd) .and. ii < ixMhi(dd)) &
                   call die('gridsetup: only tensor-product grids are built')
            end if
          end do
        end do
      end do
      ! x in the ghost layers as the reference extrapolates them
      do ii = 1, n
        if (ii < ixMlo(dd) .or. ii > ixMhi(dd)) then
          select case (dd)
          case (1)
            x(ii, :, :, 1) = qx(ii)
          case (2)
            x(:, ii, :, 2) = qx(ii)
          case (3)
            x(:, :, ii, 3) = qx(ii)
          end select
        end if
      end do
      off = off + n
      deallocate(qx)
    end do
    ! dvolume = product of dx (slab), volume over the mesh
    dxg = 1.0_dp
    off = 0
    do dd = 1, ndim
      do i3 = 1, ixGhi(3)
        do i2 = 1, ixGhi(2)
          do i1 = 1, ixGhi(1)
            select case (dd)
            case (1)
              ii = i1
            case (2)
              ii = i2
            case (3)
              ii = i3
            end select
            dxg(i1, i2, i3, 1) = dxg(i1, i2, i3, 1) * dxaxis(off + ii)
          end do
        end do
      end do
      off = off + ixGhi(dd)
    end do
    volume = sum(dxg(ixMlo(1):ixMhi(1), ixMlo(2):ixMhi(2), ixMlo(3):ixMhi(3), 1))
  end subroutine gridsetup1

  ! timetosave (src/vac.t:608-638)
  logical function timetosave(jf)
    integer, intent(in) :: jf
    logical :: oksave
    oksave = .false.
    if (t >= tsave(isavet(jf), jf)) then
      oksave = .true.
      isavet(jf) = isavet(jf) + 1
    end if
    if (it == itsave(isaveit(jf), jf)) then
      oksave = .true.
      isaveit(jf) = isaveit(jf) + 1
    end if
    if (it == itsavelast(jf) + ditsave(jf)) oksave = .true.
    if (t >= tsavelast(jf) + dtsave(jf)) oksave = .true.
    if (oksave) then
      tsavelast(jf) = t
      itsavelast(jf) = it
    end if
    timetosave = oksave
  end function timetosave

  ! timetofinish (src/vac.t:585-605)
  logical function timetofinish()
    logical :: okfinish
    okfinish = it >= itmax .or. t >= tmax .or. dt < dtmin
    if (cputimemax < bigdouble .and. .not. okfinish) okfinish = cputimemax <= cputime() - time0
    timetofinish = okfinish
  end function timetofinish

end program sac_vac
