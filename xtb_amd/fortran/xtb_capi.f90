! Drop-in implementation of xtb's C API (include/xtb.h, API version 2.0.0) for
! the GFN2-xTB single point on MI355X.
!
! The Fortran host layer owns the handle objects (environment with its message
! log, molecule, calculator, results with the cached wavefunction) and calls
! the HIP engine through its C ABI (include/xtb_amd_batch.h) via ISO_C_BINDING.
! Behaviour follows the reference API modules:
!   environment  src/api/environment.f90:46-258, src/type/environment.f90:162-346
!   molecule     src/api/molecule.f90:45-150, src/api/utils.f90:63-72
!   calculator   src/api/calculator.f90:55-700
!   single point src/api/interface.f90:40-156
!   results      src/api/results.f90:53-501
! Calls outside the GFN2-xTB molecular single point (GFN0/GFN1/GFN-FF,
! solvation, external charges, CPCM-X, Hessian) leave an error in the
! environment log.
module xtbamd_capi
   use, intrinsic :: iso_c_binding
   use, intrinsic :: iso_fortran_env, only : output_unit, error_unit
   implicit none
   private

   integer, parameter :: wp = c_double
   integer(c_int), parameter :: api_version = 20000_c_int
   real(wp), parameter :: autoev = 27.21138505_wp
   real(wp), parameter :: evtoau = 1.0_wp / autoev
   integer, parameter :: max_elem = 86

   !> engine result record (xtbamd_result of include/xtb_amd_batch.h)
   type, bind(C) :: engine_result
      integer(c_int) :: natoms, nshells, nao, iterations, converged, failed, ihomo
      real(c_double) :: energy, gnorm, hl_gap_ev, e_elec, e_es, e_aes, e_axc, e_disp, e_rep, e_atm
      real(c_double) :: e_solv
      real(c_double) :: dipole(3)
      type(c_ptr) :: gradient, charges, qsh, dipm, qp, wbo, emo, focc, coeff, density
   end type engine_result

   interface
      function engine_new(nmol, natoms, numbers, positions, charges, uhf, accuracy, etemp, maxiter, guess) &
            & result(handle) bind(C, name="xtbamd_batch_new")
         import :: c_int, c_double, c_ptr
         integer(c_int), value :: nmol
         integer(c_int), intent(in) :: natoms(*), numbers(*), uhf(*)
         real(c_double), intent(in) :: positions(*), charges(*)
         real(c_double), value :: accuracy, etemp
         integer(c_int), value :: maxiter, guess
         type(c_ptr) :: handle
      end function engine_new
      function engine_set_guess(handle, m, qsh, dipm, qp) result(rc) bind(C, name="xtbamd_batch_set_guess")
         import :: c_int, c_double, c_ptr
         type(c_ptr), value :: handle
         integer(c_int), value :: m
         real(c_double), intent(in) :: qsh(*), dipm(*), qp(*)
         integer(c_int) :: rc
      end function engine_set_guess
      function engine_setup(handle) result(rc) bind(C, name="xtbamd_batch_setup")
         import :: c_int, c_ptr
         type(c_ptr), value :: handle
         integer(c_int) :: rc
      end function engine_setup
      function engine_run(handle) result(rc) bind(C, name="xtbamd_batch_run")
         import :: c_int, c_ptr
         type(c_ptr), value :: handle
         integer(c_int) :: rc
      end function engine_run
      function engine_get(handle, m, res) result(rc) bind(C, name="xtbamd_batch_get")
         import :: c_int, c_ptr, engine_result
         type(c_ptr), value :: handle
         integer(c_int), value :: m
         type(engine_result), intent(inout) :: res
         integer(c_int) :: rc
      end function engine_get
      function engine_set_solvent(handle, solvent, kernel, alpb, state, temperature) result(rc) &
            & bind(C, name="xtbamd_batch_set_solvent")
         import :: c_int, c_double, c_ptr, c_char
         type(c_ptr), value :: handle
         character(kind=c_char), intent(in) :: solvent(*)
         integer(c_int), value :: kernel, alpb, state
         real(c_double), value :: temperature
         integer(c_int) :: rc
      end function engine_set_solvent
      function engine_solvent_known(solvent, kernel) result(ok) bind(C, name="xtbamd_solvent_known")
         import :: c_int, c_char
         character(kind=c_char), intent(in) :: solvent(*)
         integer(c_int), value :: kernel
         integer(c_int) :: ok
      end function engine_solvent_known
      function engine_set_positions(handle, m, positions) result(rc) bind(C, name="xtbamd_batch_set_positions")
         import :: c_int, c_double, c_ptr
         type(c_ptr), value :: handle
         integer(c_int), value :: m
         real(c_double), intent(in) :: positions(*)
         integer(c_int) :: rc
      end function engine_set_positions
      function engine_keep_wfn(handle, on) result(rc) bind(C, name="xtbamd_batch_keep_wavefunction")
         import :: c_int, c_ptr
         type(c_ptr), value :: handle
         integer(c_int), value :: on
         integer(c_int) :: rc
      end function engine_keep_wfn
      subroutine engine_free(handle) bind(C, name="xtbamd_batch_free")
         import :: c_ptr
         type(c_ptr), value :: handle
      end subroutine engine_free
      function engine_last_error() result(msg) bind(C, name="xtbamd_last_error")
         import :: c_ptr
         type(c_ptr) :: msg
      end function engine_last_error
      function c_strlen(s) result(n) bind(C, name="strlen")
         import :: c_ptr, c_size_t
         type(c_ptr), value :: s
         integer(c_size_t) :: n
      end function c_strlen
   end interface

   type :: message_t
      logical :: error = .false.
      character(len=:), allocatable :: text
   end type message_t

   type :: environment_t
      type(message_t), allocatable :: log(:)
      integer :: nlog = 0
      integer :: verbosity = 1
      integer :: unit = output_unit
   end type environment_t

   type :: molecule_t
      integer :: n = 0
      integer(c_int), allocatable :: at(:)
      real(wp), allocatable :: xyz(:, :)
      real(wp) :: chrg = 0.0_wp
      integer(c_int) :: uhf = 0
   end type molecule_t

   type :: calculator_t
      logical :: loaded = .false.
      real(wp) :: accuracy = 1.0_wp
      real(wp) :: etemp = 300.0_wp
      integer :: maxiter = 250
      !> implicit solvent (src/api/calculator.f90:341-440): name, Born kernel
      !> (1 Still / GBSA parameters, 2 P16 / ALPB parameters), ALPB shape term,
      !> solution state and temperature; unallocated name: gas phase
      character(len=:), allocatable :: solvent
      integer(c_int) :: solv_kernel = 1, solv_alpb = 0, solv_state = 1
      real(wp) :: solv_temp = 298.15_wp
   end type calculator_t

   type :: results_t
      !> cached wavefunction (restart guess of the next single point)
      logical :: has_wfn = .false.
      integer :: n = 0, nsh = 0, nao = 0
      real(wp), allocatable :: q(:), qsh(:), dipm(:, :), qp(:, :), wbo(:, :), emo(:), focc(:), coeff(:, :)
      !> single point results
      logical :: has_energy = .false., has_gradient = .false., has_sigma = .false., has_dipole = .false.
      real(wp) :: energy = 0.0_wp, egap = 0.0_wp
      real(wp), allocatable :: gradient(:, :)
      real(wp) :: sigma(3, 3) = 0.0_wp
      real(wp) :: dipole(3) = 0.0_wp
      !> device-resident engine of the last single point and what it was built
      !> for: while the atoms, charge, spin, calculator settings and solvent
      !> stay the same, the next single point reuses its device arrays and
      !> starts from the wavefunction it left on the device
      type(c_ptr) :: engine = c_null_ptr
      integer(c_int), allocatable :: eng_at(:)
      real(wp) :: eng_chrg = 0.0_wp, eng_acc = 0.0_wp, eng_etemp = 0.0_wp, eng_stemp = 0.0_wp
      integer :: eng_uhf = 0, eng_maxiter = 0, eng_kernel = 0, eng_alpb = 0, eng_state = 0
      character(len=:), allocatable :: eng_solvent
   end type results_t

contains

   ! --------------------------------------------------------------------------
   ! message log (src/type/environment.f90:162-346)

   subroutine env_push(env, error, text, source)
      type(environment_t), intent(inout) :: env
      logical, intent(in) :: error
      character(len=*), intent(in) :: text, source
      type(message_t), allocatable :: tmp(:)
      if (.not. allocated(env%log)) allocate(env%log(8))
      if (env%nlog >= size(env%log)) then
         allocate(tmp(2 * size(env%log)))
         tmp(:env%nlog) = env%log(:env%nlog)
         call move_alloc(tmp, env%log)
      end if
      env%nlog = env%nlog + 1
      env%log(env%nlog)%error = error
      env%log(env%nlog)%text = source // ": " // text
   end subroutine env_push

   subroutine env_error(env, text, source)
      type(environment_t), intent(inout) :: env
      character(len=*), intent(in) :: text, source
      call env_push(env, .true., text, source)
   end subroutine env_error

   subroutine env_warning(env, text, source)
      type(environment_t), intent(inout) :: env
      character(len=*), intent(in) :: text, source
      call env_push(env, .false., text, source)
   end subroutine env_warning

   logical function env_failed(env)
      type(environment_t), intent(in) :: env
      integer :: i
      env_failed = .false.
      do i = 1, env%nlog
         if (env%log(i)%error) env_failed = .true.
      end do
   end function env_failed

   !> newest message first, "-k- text" lines, then the log is emptied
   subroutine env_take_log(env, text)
      type(environment_t), intent(inout) :: env
      character(len=:), allocatable, intent(out) :: text
      character(len=512) :: line
      integer :: i
      text = ''
      do i = env%nlog, 1, -1
         write(line, '("-", i0, "-", 1x, a)') i, env%log(i)%text
         if (len_trim(text) > 0) then
            text = text // new_line('a') // trim(line)
         else
            text = trim(line)
         end if
      end do
      env%nlog = 0
   end subroutine env_take_log

   subroutine c_to_f_string(cstr, fstr)
      character(kind=c_char), intent(in) :: cstr(*)
      character(len=:), allocatable, intent(out) :: fstr
      integer :: i, n
      n = 0
      do while (cstr(n + 1) /= c_null_char)
         n = n + 1
      end do
      allocate(character(len=n) :: fstr)
      do i = 1, n
         fstr(i:i) = cstr(i)
      end do
   end subroutine c_to_f_string

   subroutine f_to_c_string(fstr, cstr, maxlen)
      character(len=*), intent(in) :: fstr
      character(kind=c_char), intent(inout) :: cstr(*)
      integer, intent(in) :: maxlen
      integer :: i, n
      n = min(maxlen - 1, len_trim(fstr))
      do i = 1, n
         cstr(i) = fstr(i:i)
      end do
      cstr(n + 1) = c_null_char
   end subroutine f_to_c_string

   ! --------------------------------------------------------------------------
   ! API version and environment

   function getAPIVersion_api() result(version) bind(C, name="xtb_getAPIVersion")
      integer(c_int) :: version
      version = api_version
   end function getAPIVersion_api

   function newEnvironment_api() result(venv) bind(C, name="xtb_newEnvironment")
      type(c_ptr) :: venv
      type(environment_t), pointer :: env
      allocate(env)
      env%verbosity = 1
      venv = c_loc(env)
   end function newEnvironment_api

   subroutine delEnvironment_api(venv) bind(C, name="xtb_delEnvironment")
      type(c_ptr), intent(inout) :: venv
      type(environment_t), pointer :: env
      call releaseOutput_api(venv)
      if (c_associated(venv)) then
         call c_f_pointer(venv, env)
         deallocate(env)
         venv = c_null_ptr
      end if
   end subroutine delEnvironment_api

   function checkEnvironment_api(venv) result(status) bind(C, name="xtb_checkEnvironment")
      type(c_ptr), value :: venv
      integer(c_int) :: status
      type(environment_t), pointer :: env
      status = 0
      if (c_associated(venv)) then
         call c_f_pointer(venv, env)
         if (env_failed(env)) status = 1
      end if
   end function checkEnvironment_api

   subroutine showEnvironment_api(venv, charptr) bind(C, name="xtb_showEnvironment")
      type(c_ptr), value :: venv
      character(kind=c_char), intent(in), optional :: charptr(*)
      type(environment_t), pointer :: env
      character(len=:), allocatable :: message
      integer :: i
      if (present(charptr)) then
         call c_to_f_string(charptr, message)
      else
         message = ''
      end if
      if (.not. c_associated(venv)) return
      call c_f_pointer(venv, env)
      if (env%nlog <= 0) return
      if (env_failed(env)) then
         write(env%unit, '(72("#"), /, "[ERROR]", 1x, a)') message
      else
         write(env%unit, '(72("#"), /, "[WARNING]", 1x, a)') message
      end if
      do i = env%nlog, 1, -1
         write(env%unit, '("-", i0, "-", 1x, a)') i, env%log(i)%text
      end do
      write(env%unit, '(72("#"))')
      env%nlog = 0
   end subroutine showEnvironment_api

   subroutine getError_api(venv, charptr, buffersize) bind(C, name="xtb_getError")
      type(c_ptr), value :: venv
      character(kind=c_char), intent(inout) :: charptr(*)
      integer(c_int), intent(in), optional :: buffersize
      type(environment_t), pointer :: env
      character(len=:), allocatable :: text
      integer :: maxlen
      if (.not. c_associated(venv)) return
      call c_f_pointer(venv, env)
      if (present(buffersize)) then
         maxlen = buffersize
      else
         maxlen = huge(maxlen) - 2
      end if
      if (env_failed(env)) then
         call env_take_log(env, text)
         if (len(text) > 0) call f_to_c_string(text, charptr, maxlen)
      end if
   end subroutine getError_api

   subroutine setOutput_api(venv, charptr) bind(C, name="xtb_setOutput")
      type(c_ptr), value :: venv
      character(kind=c_char), intent(in) :: charptr(*)
      character(len=*), parameter :: source = 'xtb_api_setOutput'
      type(environment_t), pointer :: env
      character(len=:), allocatable :: filename
      integer :: stat
      character(len=512) :: msg
      call c_to_f_string(charptr, filename)
      call releaseOutput_api(venv)
      if (.not. c_associated(venv)) return
      call c_f_pointer(venv, env)
      if (filename == '-' .or. filename == 'STDOUT') then
         env%unit = output_unit
      else if (filename == 'STDERR') then
         env%unit = error_unit
      else
         open(file=filename, newunit=env%unit, iostat=stat, iomsg=msg)
         if (stat /= 0) then
            env%unit = output_unit
            call env_error(env, trim(msg), source)
         end if
      end if
   end subroutine setOutput_api

   subroutine releaseOutput_api(venv) bind(C, name="xtb_releaseOutput")
      type(c_ptr), value :: venv
      character(len=*), parameter :: source = 'xtb_api_releaseOutput'
      type(environment_t), pointer :: env
      integer :: stat
      character(len=512) :: msg
      if (.not. c_associated(venv)) return
      call c_f_pointer(venv, env)
      if (env%unit /= output_unit .and. env%unit /= error_unit .and. env%unit /= -1) then
         close(unit=env%unit, iostat=stat, iomsg=msg)
         if (stat /= 0) call env_error(env, trim(msg), source)
         env%unit = output_unit
      end if
   end subroutine releaseOutput_api

   subroutine setVerbosity_api(venv, verbosity) bind(C, name="xtb_setVerbosity")
      type(c_ptr), value :: venv
      integer(c_int), value :: verbosity
      type(environment_t), pointer :: env
      if (.not. c_associated(venv)) return
      call c_f_pointer(venv, env)
      env%verbosity = verbosity
   end subroutine setVerbosity_api

   ! --------------------------------------------------------------------------
   ! molecule

   function newMolecule_api(venv, natoms, numbers, positions, charge, uhf, lattice, periodic) result(vmol) &
         & bind(C, name="xtb_newMolecule")
      type(c_ptr), value :: venv
      integer(c_int), intent(in) :: natoms
      integer(c_int), intent(in) :: numbers(natoms)
      real(c_double), intent(in) :: positions(3, natoms)
      real(c_double), intent(in), optional :: charge
      integer(c_int), intent(in), optional :: uhf
      real(c_double), intent(in), optional :: lattice(3, 3)
      logical(c_bool), intent(in), optional :: periodic(3)
      type(c_ptr) :: vmol
      character(len=*), parameter :: source = 'xtb_api_newMolecule'
      type(environment_t), pointer :: env
      type(molecule_t), pointer :: mol
      vmol = c_null_ptr
      if (.not. c_associated(venv)) return
      call c_f_pointer(venv, env)
      if (natoms <= 0) then
         call env_error(env, "Number of atoms must be positive", source)
         return
      end if
      if (any(numbers <= 0) .or. any(numbers > 118)) then
         call env_error(env, "Invalid atomic number present", source)
         return
      end if
      if (present(periodic)) then
         if (any(logical(periodic))) then
            call env_error(env, "Periodic systems are not supported by this GFN2-xTB implementation", source)
            return
         end if
      end if
      allocate(mol)
      mol%n = natoms
      mol%at = numbers
      mol%xyz = positions
      if (present(charge)) mol%chrg = charge
      if (present(uhf)) mol%uhf = uhf
      if (overlapping_atoms(mol) /= 0) then
         deallocate(mol)
         call env_error(env, "Could not generate molecular structure", source)
         return
      end if
      vmol = c_loc(mol)
   end function newMolecule_api

   !> number of atom pairs closer than 1e-9 bohr (verifyMolecule, src/api/utils.f90:63-72)
   integer function overlapping_atoms(mol) result(status)
      type(molecule_t), intent(in) :: mol
      integer :: i, j
      status = 0
      do i = 1, mol%n
         do j = 1, i - 1
            if (norm2(mol%xyz(:, i) - mol%xyz(:, j)) < 1.0e-9_wp) status = status + 1
         end do
      end do
   end function overlapping_atoms

   subroutine delMolecule_api(vmol) bind(C, name="xtb_delMolecule")
      type(c_ptr), intent(inout) :: vmol
      type(molecule_t), pointer :: mol
      if (c_associated(vmol)) then
         call c_f_pointer(vmol, mol)
         deallocate(mol)
         vmol = c_null_ptr
      end if
   end subroutine delMolecule_api

   subroutine updateMolecule_api(venv, vmol, positions, lattice) bind(C, name="xtb_updateMolecule")
      type(c_ptr), value :: venv, vmol
      real(c_double), intent(in) :: positions(3, *)
      real(c_double), intent(in), optional :: lattice(3, 3)
      character(len=*), parameter :: source = 'xtb_api_updateMolecule'
      type(environment_t), pointer :: env
      type(molecule_t), pointer :: mol
      if (.not. c_associated(venv)) return
      call c_f_pointer(venv, env)
      if (.not. c_associated(vmol)) then
         call env_error(env, "Molecular structure data is not allocated", source)
         return
      end if
      call c_f_pointer(vmol, mol)
      if (mol%n <= 0 .or. .not. allocated(mol%xyz)) then
         call env_error(env, "Invalid molecular structure data provided", source)
         return
      end if
      mol%xyz(:, :) = positions(1:3, 1:mol%n)
      if (overlapping_atoms(mol) /= 0) then
         call env_error(env, "Could not update molecular structure", source)
      end if
   end subroutine updateMolecule_api

   ! --------------------------------------------------------------------------
   ! calculator

   function newCalculator_api() result(vcalc) bind(C, name="xtb_newCalculator")
      type(c_ptr) :: vcalc
      type(calculator_t), pointer :: calc
      allocate(calc)
      vcalc = c_loc(calc)
   end function newCalculator_api

   subroutine delCalculator_api(vcalc) bind(C, name="xtb_delCalculator")
      type(c_ptr), intent(inout) :: vcalc
      type(calculator_t), pointer :: calc
      if (c_associated(vcalc)) then
         call c_f_pointer(vcalc, calc)
         deallocate(calc)
         vcalc = c_null_ptr
      end if
   end subroutine delCalculator_api

   !> common argument checks of the load routines (src/api/calculator.f90:278-336)
   logical function load_checks(venv, vmol, vcalc, source, env, mol, calc) result(ok)
      type(c_ptr), value :: venv, vmol, vcalc
      character(len=*), intent(in) :: source
      type(environment_t), pointer, intent(out) :: env
      type(molecule_t), pointer, intent(out) :: mol
      type(calculator_t), pointer, intent(out) :: calc
      ok = .false.
      nullify(env, mol, calc)
      if (.not. c_associated(venv)) return
      call c_f_pointer(venv, env)
      if (.not. c_associated(vmol)) then
         call env_error(env, "Molecular structure data is not allocated", source)
         return
      end if
      call c_f_pointer(vmol, mol)
      if (.not. c_associated(vcalc)) then
         call env_error(env, "Singlepoint calculator is not allocated", source)
         return
      end if
      call c_f_pointer(vcalc, calc)
      ok = .true.
   end function load_checks

   subroutine loadGFN2xTB_api(venv, vmol, vcalc, charptr) bind(C, name="xtb_loadGFN2xTB")
      type(c_ptr), value :: venv, vmol, vcalc
      character(kind=c_char), intent(in), optional :: charptr(*)
      character(len=*), parameter :: source = 'xtb_api_loadGFN2xTB'
      type(environment_t), pointer :: env
      type(molecule_t), pointer :: mol
      type(calculator_t), pointer :: calc
      ! GFN2-xTB parameters are built in; a parameter file name is not read
      if (.not. load_checks(venv, vmol, vcalc, source, env, mol, calc)) return
      if (any(mol%at > max_elem)) then
         call env_error(env, "Element not parametrised in GFN2-xTB", 'xtb_calculator_newXTBCalculator')
         call env_error(env, "Could not construct GFN2-xTB calculator", source)
         return
      end if
      calc%loaded = .true.
      calc%accuracy = 1.0_wp
      calc%etemp = 300.0_wp
      calc%maxiter = 250
   end subroutine loadGFN2xTB_api

   subroutine unsupported_load(venv, vmol, vcalc, method, source)
      type(c_ptr), value :: venv, vmol, vcalc
      character(len=*), intent(in) :: method, source
      type(environment_t), pointer :: env
      type(molecule_t), pointer :: mol
      type(calculator_t), pointer :: calc
      if (.not. load_checks(venv, vmol, vcalc, source, env, mol, calc)) return
      call env_error(env, method // " is not supported by this GFN2-xTB implementation", source)
      call env_error(env, "Could not construct " // method // " calculator", source)
   end subroutine unsupported_load

   subroutine loadGFN0xTB_api(venv, vmol, vcalc, charptr) bind(C, name="xtb_loadGFN0xTB")
      type(c_ptr), value :: venv, vmol, vcalc
      character(kind=c_char), intent(in), optional :: charptr(*)
      call unsupported_load(venv, vmol, vcalc, "GFN0-xTB", 'xtb_api_loadGFN0xTB')
   end subroutine loadGFN0xTB_api

   subroutine loadGFN1xTB_api(venv, vmol, vcalc, charptr) bind(C, name="xtb_loadGFN1xTB")
      type(c_ptr), value :: venv, vmol, vcalc
      character(kind=c_char), intent(in), optional :: charptr(*)
      call unsupported_load(venv, vmol, vcalc, "GFN1-xTB", 'xtb_api_loadGFN1xTB')
   end subroutine loadGFN1xTB_api

   subroutine loadGFNFF_api(venv, vmol, vcalc, charptr) bind(C, name="xtb_loadGFNFF")
      type(c_ptr), value :: venv, vmol, vcalc
      character(kind=c_char), intent(in), optional :: charptr(*)
      call unsupported_load(venv, vmol, vcalc, "GFN-FF", 'xtb_api_loadGFNFF')
   end subroutine loadGFNFF_api

   !> calculator argument check of the setters; returns the loaded calculator
   logical function calc_checks(venv, vcalc, source, env, calc) result(ok)
      type(c_ptr), value :: venv, vcalc
      character(len=*), intent(in) :: source
      type(environment_t), pointer, intent(out) :: env
      type(calculator_t), pointer, intent(out) :: calc
      ok = .false.
      nullify(env, calc)
      if (.not. c_associated(venv)) return
      call c_f_pointer(venv, env)
      if (.not. c_associated(vcalc)) then
         call env_error(env, "Singlepoint calculator is not allocated", source)
         return
      end if
      call c_f_pointer(vcalc, calc)
      ok = .true.
   end function calc_checks

   subroutine setSolvent_api(venv, vcalc, charptr, state, temperature, grid, charptr2) bind(C, name="xtb_setSolvent")
      type(c_ptr), value :: venv, vcalc
      character(kind=c_char), intent(in) :: charptr(*)
      integer(c_int), intent(in), optional :: state, grid
      real(c_double), intent(in), optional :: temperature
      character(kind=c_char), intent(in), optional :: charptr2(*)
      character(len=*), parameter :: source = 'xtb_api_setSolvent'
      type(environment_t), pointer :: env
      type(calculator_t), pointer :: calc
      character(len=:), allocatable :: solvent, model
      integer(c_int) :: gstate, kernel, alpb, nang
      real(wp) :: temp
      if (.not. calc_checks(venv, vcalc, source, env, calc)) return
      if (.not. calc%loaded) then
         call env_error(env, "No calculator loaded to add solvation model", source)
         return
      end if
      gstate = 1
      if (present(state)) gstate = state
      temp = 298.15_wp
      if (present(temperature)) temp = temperature
      nang = 230
      if (present(grid)) nang = grid
      call c_to_f_string(charptr, solvent)
      model = 'gbsa'
      if (present(charptr2)) call c_to_f_string(charptr2, model)
      select case(model)
      case('gbsa')
         kernel = 1
         alpb = 0
      case('alpb')
         kernel = 2
         alpb = 1
      case('cosmo', 'tmcosmo', 'cpcmx')
         call env_error(env, "Solvation model '" // model // "' is not supported by this GFN2-xTB implementation", &
            & source)
         return
      case default
         call env_error(env, "Unknown solvation model", source)
         return
      end select
      if (nang /= 230) then
         call env_error(env, "Only the 230-point angular grid is supported by this GFN2-xTB implementation", source)
         return
      end if
      if (gstate < 1 .or. gstate > 3) then
         call env_error(env, "Unknown solution state", source)
         return
      end if
      if (engine_solvent_known(solvent // c_null_char, kernel) == 0) then
         call env_error(env, "solvent: '" // solvent // "' is not parametrized", 'solv_model_loadInternalParam')
         call env_error(env, "Could not add solvation model for '" // solvent // "'", source)
         return
      end if
      calc%solvent = solvent
      calc%solv_kernel = kernel
      calc%solv_alpb = alpb
      calc%solv_state = gstate
      calc%solv_temp = temp
   end subroutine setSolvent_api

   subroutine releaseSolvent_api(venv, vcalc) bind(C, name="xtb_releaseSolvent")
      type(c_ptr), value :: venv, vcalc
      type(environment_t), pointer :: env
      type(calculator_t), pointer :: calc
      if (.not. calc_checks(venv, vcalc, 'xtb_api_releaseSolvent', env, calc)) return
      if (allocated(calc%solvent)) deallocate(calc%solvent)
   end subroutine releaseSolvent_api

   subroutine setExternalCharges_api(venv, vcalc, n, numbers, charges, positions) &
         & bind(C, name="xtb_setExternalCharges")
      type(c_ptr), value :: venv, vcalc
      integer(c_int), intent(in) :: n
      integer(c_int), intent(in) :: numbers(*)
      real(c_double), intent(in) :: charges(*), positions(3, *)
      character(len=*), parameter :: source = 'xtb_api_setExternalCharges'
      type(environment_t), pointer :: env
      type(calculator_t), pointer :: calc
      if (.not. c_associated(venv)) return
      call c_f_pointer(venv, env)
      if (n < 0) then
         call env_error(env, "Negative number of point charges provided", source)
         return
      end if
      if (.not. calc_checks(venv, vcalc, source, env, calc)) return
      if (.not. calc%loaded) then
         call env_error(env, "No calculator loaded to add external potential", source)
         return
      end if
      call env_error(env, "External potentials are not supported by this GFN2-xTB implementation", source)
   end subroutine setExternalCharges_api

   subroutine releaseExternalCharges_api(venv, vcalc) bind(C, name="xtb_releaseExternalCharges")
      type(c_ptr), value :: venv, vcalc
      type(environment_t), pointer :: env
      type(calculator_t), pointer :: calc
      if (.not. calc_checks(venv, vcalc, 'xtb_api_releaseExternalCharges', env, calc)) return
   end subroutine releaseExternalCharges_api

   subroutine setAccuracy_api(venv, vcalc, accuracy) bind(C, name="xtb_setAccuracy")
      type(c_ptr), value :: venv, vcalc
      real(c_double), value :: accuracy
      character(len=*), parameter :: source = 'xtb_api_setAccuracy'
      type(environment_t), pointer :: env
      type(calculator_t), pointer :: calc
      if (.not. calc_checks(venv, vcalc, source, env, calc)) return
      if (.not. calc%loaded) then
         call env_error(env, "Setting accuracy not possible, no calculator loaded", source)
         return
      end if
      if (accuracy < 1.e-4_c_double) then
         call env_warning(env, "We cannot provide this level of accuracy, resetted accuracy to 0.0001", source)
         calc%accuracy = 1.e-4_wp
      else if (accuracy > 1.e+3_c_double) then
         call env_warning(env, "We cannot provide this level of accuracy, resetted accuracy to 1000", source)
         calc%accuracy = 1.e+3_wp
      else
         calc%accuracy = accuracy
      end if
   end subroutine setAccuracy_api

   subroutine setMaxIter_api(venv, vcalc, maxiter) bind(C, name="xtb_setMaxIter")
      type(c_ptr), value :: venv, vcalc
      integer(c_int), value :: maxiter
      character(len=*), parameter :: source = 'xtb_api_setMaxIter'
      type(environment_t), pointer :: env
      type(calculator_t), pointer :: calc
      if (.not. calc_checks(venv, vcalc, source, env, calc)) return
      if (.not. calc%loaded) then
         call env_error(env, "Setting accuracy not possible, no calculator loaded", source)
         return
      end if
      calc%maxiter = max(maxiter, 1)
   end subroutine setMaxIter_api

   subroutine setElectronicTemp_api(venv, vcalc, temperature) bind(C, name="xtb_setElectronicTemp")
      type(c_ptr), value :: venv, vcalc
      real(c_double), value :: temperature
      character(len=*), parameter :: source = 'xtb_api_setElectronicTemp'
      type(environment_t), pointer :: env
      type(calculator_t), pointer :: calc
      if (.not. calc_checks(venv, vcalc, source, env, calc)) return
      if (.not. calc%loaded) then
         call env_error(env, "Setting accuracy not possible, no calculator loaded", source)
         return
      end if
      calc%etemp = max(temperature, 1.0e-6_c_double)
   end subroutine setElectronicTemp_api

   ! --------------------------------------------------------------------------
   ! single point (src/api/interface.f90:40-156)

   subroutine singlepoint_api(venv, vmol, vcalc, vres) bind(C, name="xtb_singlepoint")
      type(c_ptr), value :: venv, vmol, vcalc, vres
      character(len=*), parameter :: source = 'xtb_api_singlepoint'
      type(environment_t), pointer :: env
      type(molecule_t), pointer :: mol
      type(calculator_t), pointer :: calc
      type(results_t), pointer :: res
      type(c_ptr) :: handle
      type(engine_result), target :: er
      integer(c_int) :: natoms(1), uhf(1), rc, guess
      real(c_double) :: chrg(1)
      integer :: nsh, nao, n
      logical :: reuse
      real(wp), allocatable, target :: qsh(:), dipm(:, :), qp(:, :), grad(:, :), q(:), wbo(:, :), emo(:), &
         & focc(:), coeff(:, :)

      if (.not. c_associated(venv)) return
      call c_f_pointer(venv, env)
      if (.not. c_associated(vmol)) then
         call env_error(env, "Molecular structure data is not allocated", source)
         return
      end if
      call c_f_pointer(vmol, mol)
      if (.not. c_associated(vcalc)) then
         call env_error(env, "Singlepoint calculator is not allocated", source)
         return
      end if
      call c_f_pointer(vcalc, calc)
      if (.not. calc%loaded) then
         call env_error(env, "No calculator loaded for single point", source)
         return
      end if
      if (.not. c_associated(vres)) then
         call env_error(env, "Calculation results are not allocated", source)
         return
      end if
      call c_f_pointer(vres, res)

      n = mol%n
      natoms(1) = n
      uhf(1) = mol%uhf
      chrg(1) = mol%chrg
      ! the cached wavefunction seeds the SCC unless the dimensions changed
      guess = 0
      if (res%has_wfn .and. res%n /= n) res%has_wfn = .false.
      reuse = engine_matches(res, mol, calc)
      if (reuse) then
         ! the results object's engine: new positions, device arrays kept
         handle = res%engine
         rc = engine_set_positions(handle, 0_c_int, mol%xyz)
         if (rc /= 0) then
            call env_error(env, engine_message(), 'xtb_calculator_singlepoint')
            call drop_engine(res)
            return
         end if
      else
         call drop_engine(res)
         handle = engine_new(1_c_int, natoms, mol%at, mol%xyz, chrg, uhf, calc%accuracy, calc%etemp, &
            & int(calc%maxiter, c_int), guess)
         if (.not. c_associated(handle)) then
            call env_error(env, engine_message(), 'xtb_calculator_singlepoint')
            return
         end if
         if (.not. apply_solvent(env, calc, handle)) return
      end if
      ! sizes
      er%gradient = c_null_ptr; er%charges = c_null_ptr; er%qsh = c_null_ptr; er%dipm = c_null_ptr
      er%qp = c_null_ptr; er%wbo = c_null_ptr; er%emo = c_null_ptr; er%focc = c_null_ptr
      er%coeff = c_null_ptr; er%density = c_null_ptr
      rc = engine_get(handle, 0_c_int, er)
      nsh = er%nshells
      nao = er%nao
      if (res%has_wfn .and. (res%nsh /= nsh .or. res%nao /= nao)) res%has_wfn = .false.
      if (reuse) then
         ! the wavefunction of the last single point is still on the device
         rc = engine_keep_wfn(handle, merge(1_c_int, 0_c_int, res%has_wfn))
      else if (res%has_wfn) then
         rc = engine_set_guess(handle, 0_c_int, res%qsh, res%dipm, res%qp)
      end if
      rc = engine_run(handle)
      if (rc /= 0) then
         call env_error(env, engine_message(), 'xtb_calculator_singlepoint')
         if (reuse) then
            call drop_engine(res)
         else
            call engine_free(handle)
         end if
         return
      end if
      allocate(qsh(nsh), dipm(3, n), qp(6, n), grad(3, n), q(n), wbo(n, n), emo(nao), focc(nao), coeff(nao, nao))
      er%gradient = c_loc(grad); er%charges = c_loc(q); er%qsh = c_loc(qsh); er%dipm = c_loc(dipm)
      er%qp = c_loc(qp); er%wbo = c_loc(wbo); er%emo = c_loc(emo); er%focc = c_loc(focc)
      er%coeff = c_loc(coeff); er%density = c_null_ptr
      rc = engine_get(handle, 0_c_int, er)
      if (rc /= 0) then
         call env_error(env, engine_message(), 'xtb_calculator_singlepoint')
         if (reuse) then
            call drop_engine(res)
         else
            call engine_free(handle)
         end if
         return
      end if
      ! the engine stays with the results object (freed by xtb_delResults)
      if (.not. reuse) call keep_engine(res, handle, mol, calc)

      ! cache the wavefunction; coefficients are stored AO x MO (column = orbital)
      res%has_wfn = .true.
      res%n = n
      res%nsh = nsh
      res%nao = nao
      call move_alloc(q, res%q)
      call move_alloc(qsh, res%qsh)
      call move_alloc(dipm, res%dipm)
      call move_alloc(qp, res%qp)
      call move_alloc(wbo, res%wbo)
      call move_alloc(emo, res%emo)
      call move_alloc(focc, res%focc)
      res%coeff = transpose(coeff)
      if (allocated(res%gradient)) then
         if (any(shape(res%gradient) /= [3, n])) then
            call env_warning(env, "Shape missmatch in gradient, reallocating", source)
            deallocate(res%gradient)
         end if
      end if
      call move_alloc(grad, res%gradient)
      res%energy = er%energy
      res%egap = er%hl_gap_ev
      res%sigma = 0.0_wp
      res%dipole = er%dipole
      res%has_energy = .true.
      res%has_gradient = .true.
      res%has_sigma = .true.
      res%has_dipole = .true.
      if (env%verbosity >= 1) then
         write(env%unit, '(a, f20.12, a, f12.6, a, i0)') "  TOTAL ENERGY ", er%energy, " Eh   GAP ", &
            & er%hl_gap_ev, " eV   SCC iterations ", er%iterations
      end if
      if (er%failed /= 0) then
         call env_error(env, failure_message(er%failed), 'scc')
      else if (er%converged == 0) then
         call env_error(env, "Self consistent charge iterator did not converge", 'scf')
      end if
   end subroutine singlepoint_api

   !> true when the results object holds an engine built for the same atoms,
   !> charge, spin, calculator settings and solvent
   logical function engine_matches(res, mol, calc) result(ok)
      type(results_t), intent(in) :: res
      type(molecule_t), intent(in) :: mol
      type(calculator_t), intent(in) :: calc
      ok = c_associated(res%engine) .and. allocated(res%eng_at)
      if (.not. ok) return
      ok = size(res%eng_at) == mol%n
      if (.not. ok) return
      ok = all(res%eng_at == mol%at(1:mol%n)) .and. res%eng_chrg == mol%chrg .and. res%eng_uhf == mol%uhf &
         & .and. res%eng_acc == calc%accuracy .and. res%eng_etemp == calc%etemp &
         & .and. res%eng_maxiter == calc%maxiter
      if (.not. ok) return
      if (allocated(calc%solvent) .neqv. allocated(res%eng_solvent)) then
         ok = .false.
      else if (allocated(calc%solvent)) then
         ok = calc%solvent == res%eng_solvent .and. calc%solv_kernel == res%eng_kernel &
            & .and. calc%solv_alpb == res%eng_alpb .and. calc%solv_state == res%eng_state &
            & .and. calc%solv_temp == res%eng_stemp
      end if
   end function engine_matches

   !> hand an engine to the results object, with what it was built for
   subroutine keep_engine(res, handle, mol, calc)
      type(results_t), intent(inout) :: res
      type(c_ptr), intent(in) :: handle
      type(molecule_t), intent(in) :: mol
      type(calculator_t), intent(in) :: calc
      res%engine = handle
      res%eng_at = mol%at(1:mol%n)
      res%eng_chrg = mol%chrg
      res%eng_uhf = mol%uhf
      res%eng_acc = calc%accuracy
      res%eng_etemp = calc%etemp
      res%eng_maxiter = calc%maxiter
      if (allocated(res%eng_solvent)) deallocate(res%eng_solvent)
      if (allocated(calc%solvent)) then
         res%eng_solvent = calc%solvent
         res%eng_kernel = calc%solv_kernel
         res%eng_alpb = calc%solv_alpb
         res%eng_state = calc%solv_state
         res%eng_stemp = calc%solv_temp
      end if
   end subroutine keep_engine

   !> free the results object's engine, if any
   subroutine drop_engine(res)
      type(results_t), intent(inout) :: res
      if (c_associated(res%engine)) call engine_free(res%engine)
      res%engine = c_null_ptr
      if (allocated(res%eng_at)) deallocate(res%eng_at)
   end subroutine drop_engine

   !> message of an engine failure code (xtbamd_result%failed): the solver
   !> that gave up, in the reference's wording where it has one
   !> (src/scc_core.f90:478, src/scf_module.F90:642)
   function failure_message(code) result(msg)
      integer(c_int), intent(in) :: code
      character(len=:), allocatable :: msg
      select case (code)
      case (2)
         msg = "Orthonormalization of the overlap matrix failed"
      case (3)
         msg = "Diagonalization of Hamiltonian failed"
      case (4)
         msg = "EEQ guess charges failed"
      case (7)
         msg = "Solvation surface setup failed"
      case default
         msg = "Self consistent charge iterator terminated"
      end select
   end function failure_message

   !> pass the calculator's implicit solvent to an engine batch; frees the
   !> batch and logs the error on failure
   function apply_solvent(env, calc, handle) result(ok)
      type(environment_t), intent(inout) :: env
      type(calculator_t), intent(in) :: calc
      type(c_ptr), intent(inout) :: handle
      logical :: ok
      integer(c_int) :: rc
      ok = .true.
      if (.not. allocated(calc%solvent)) return
      rc = engine_set_solvent(handle, calc%solvent // c_null_char, calc%solv_kernel, calc%solv_alpb, &
         & calc%solv_state, calc%solv_temp)
      if (rc /= 0) then
         call env_error(env, engine_message(), 'xtb_calculator_singlepoint')
         call engine_free(handle)
         handle = c_null_ptr
         ok = .false.
      end if
   end function apply_solvent

   function engine_message() result(text)
      character(len=:), allocatable :: text
      type(c_ptr) :: p
      character(kind=c_char), pointer :: chars(:)
      integer :: i, n
      p = engine_last_error()
      n = int(c_strlen(p))
      call c_f_pointer(p, chars, [n])
      allocate(character(len=n) :: text)
      do i = 1, n
         text(i:i) = chars(i)
      end do
   end function engine_message

   subroutine cpcmx_calc_api(venv, vmol, vcalc, vres) bind(C, name="xtb_cpcmx_calc")
      type(c_ptr), value :: venv, vmol, vcalc, vres
      type(environment_t), pointer :: env
      if (.not. c_associated(venv)) return
      call c_f_pointer(venv, env)
      call env_error(env, "CPCM-X is not supported by this GFN2-xTB implementation", 'xtb_api_cpcmx_calc')
   end subroutine cpcmx_calc_api

   !> Numerical Hessian by central differences of analytic gradients
   !> (hessian: src/type/calculator.f90:113-228, xtb_hessian:
   !> src/api/interface.f90:310-454): all 6 x len(list) displaced single points
   !> run as one batch on the GPU; each starts from the results object's cached
   !> wavefunction, or from the EEQ guess at the undisplaced geometry.
   !> Arguments follow the reference's binding: step (default 0.005 bohr), then
   !> the 1-based atom list (NULL: all atoms; else read up to natoms entries or
   !> the first entry <= 0).  The polarizability gradient is zero for GFN2-xTB.
   subroutine hessian_api(venv, vmol, vcalc, vres, hess_out, c_step, c_list, c_dipgrad, c_polgrad) &
         & bind(C, name="xtb_hessian")
      type(c_ptr), value :: venv, vmol, vcalc, vres
      real(c_double), intent(inout) :: hess_out(*)
      real(c_double), intent(in), optional :: c_step
      integer(c_int), intent(in), optional :: c_list(*)
      real(c_double), intent(inout), optional :: c_dipgrad(*)
      real(c_double), intent(inout), optional :: c_polgrad(*)
      character(len=*), parameter :: source = 'xtb_api_hessian'
      type(environment_t), pointer :: env
      type(molecule_t), pointer :: mol
      type(calculator_t), pointer :: calc
      type(results_t), pointer :: res
      type(c_ptr) :: handle
      type(engine_result), target :: er
      real(wp) :: step, step2
      real(wp), allocatable, target :: xyz(:, :, :), hess(:, :), dipgrad(:, :), grad(:, :, :), dip(:, :)
      real(wp), allocatable, target :: qsh0(:), dipm0(:, :), qp0(:, :)
      integer(c_int), allocatable :: natoms(:), numbers(:), uhf(:)
      real(c_double), allocatable :: chrg(:)
      integer, allocatable :: list(:)
      integer :: n, nl, nmol, i, j, k, kat, ic, m, ii, nsh
      integer(c_int) :: rc
      logical :: conv

      if (.not. c_associated(venv)) return
      call c_f_pointer(venv, env)
      if (.not. c_associated(vmol)) then
         call env_error(env, "Molecular structure data is not allocated", source)
         return
      end if
      call c_f_pointer(vmol, mol)
      if (.not. c_associated(vcalc)) then
         call env_error(env, "Singlepoint calculator is not allocated", source)
         return
      end if
      call c_f_pointer(vcalc, calc)
      if (.not. calc%loaded) then
         call env_error(env, "No calculator loaded for single point", source)
         return
      end if
      if (.not. c_associated(vres)) then
         call env_error(env, "Calculation results are not allocated", source)
         return
      end if
      call c_f_pointer(vres, res)
      n = mol%n
      step = 0.005_wp
      if (present(c_step)) step = c_step
      step2 = 0.5_wp / step
      if (present(c_list)) then
         nl = 0
         do i = 1, n
            if (c_list(i) <= 0) exit
            nl = nl + 1
         end do
         allocate(list(nl))
         list = c_list(1:nl)
         if (any(list > n)) then
            call env_error(env, "Invalid atom index in the displacement list", source)
            return
         end if
      else
         nl = n
         list = [(i, i = 1, n)]
      end if

      ! the starting wavefunction of every displaced point
      if (res%has_wfn .and. res%n /= n) res%has_wfn = .false.
      if (res%has_wfn) then
         qsh0 = res%qsh
         dipm0 = res%dipm
         qp0 = res%qp
         nsh = res%nsh
      else
         natoms = [int(n, c_int)]
         uhf = [mol%uhf]
         chrg = [mol%chrg]
         handle = engine_new(1_c_int, natoms, mol%at, mol%xyz, chrg, uhf, calc%accuracy, calc%etemp, &
            & int(calc%maxiter, c_int), 0_c_int)
         if (.not. c_associated(handle)) then
            call env_error(env, engine_message(), 'xtb_calculator_newWavefunction')
            return
         end if
         rc = engine_setup(handle)
         call clear_result(er)
         rc = engine_get(handle, 0_c_int, er)
         nsh = er%nshells
         allocate(qsh0(nsh), dipm0(3, n), qp0(6, n))
         er%qsh = c_loc(qsh0)
         rc = engine_get(handle, 0_c_int, er)
         call engine_free(handle)
         if (rc /= 0) then
            call env_error(env, engine_message(), 'xtb_calculator_newWavefunction')
            return
         end if
         dipm0 = 0.0_wp
         qp0 = 0.0_wp
      end if

      ! 6 x nl displaced geometries: (+step, -step) for each atom and direction
      nmol = 6 * nl
      if (allocated(natoms)) deallocate(natoms)
      if (allocated(uhf)) deallocate(uhf)
      if (allocated(chrg)) deallocate(chrg)
      allocate(xyz(3, n, nmol), natoms(nmol), uhf(nmol), chrg(nmol), numbers(n * nmol))
      do kat = 1, nl
         do ic = 1, 3
            do k = 1, 2
               m = 6 * (kat - 1) + 2 * (ic - 1) + k
               xyz(:, :, m) = mol%xyz
               xyz(ic, list(kat), m) = mol%xyz(ic, list(kat)) + merge(step, -step, k == 1)
               numbers((m - 1) * n + 1:m * n) = mol%at
            end do
         end do
      end do
      natoms = int(n, c_int)
      uhf = mol%uhf
      chrg = mol%chrg
      handle = engine_new(int(nmol, c_int), natoms, numbers, xyz, chrg, uhf, calc%accuracy, calc%etemp, &
         & int(calc%maxiter, c_int), 0_c_int)
      if (.not. c_associated(handle)) then
         call env_error(env, engine_message(), 'xtb_calculator_singlepoint')
         return
      end if
      if (.not. apply_solvent(env, calc, handle)) return
      do m = 1, nmol
         rc = engine_set_guess(handle, int(m - 1, c_int), qsh0, dipm0, qp0)
      end do
      rc = engine_run(handle)
      if (rc /= 0) then
         call env_error(env, engine_message(), 'xtb_calculator_singlepoint')
         call engine_free(handle)
         return
      end if
      allocate(grad(3, n, nmol), dip(3, nmol))
      conv = .true.
      do m = 1, nmol
         call clear_result(er)
         er%gradient = c_loc(grad(1, 1, m))
         rc = engine_get(handle, int(m - 1, c_int), er)
         dip(:, m) = er%dipole
         conv = conv .and. er%converged /= 0 .and. er%failed == 0
      end do
      call engine_free(handle)

      allocate(hess(3 * n, 3 * n), dipgrad(3, 3 * n))
      hess = 0.0_wp
      dipgrad = 0.0_wp
      do kat = 1, nl
         do ic = 1, 3
            ii = 3 * (list(kat) - 1) + ic
            m = 6 * (kat - 1) + 2 * (ic - 1) + 1
            dipgrad(:, ii) = (dip(:, m) - dip(:, m + 1)) * step2
            do j = 1, n
               do k = 1, 3
                  hess(3 * (j - 1) + k, ii) = hess(3 * (j - 1) + k, ii) + (grad(k, j, m) - grad(k, j, m + 1)) * step2
               end do
            end do
         end do
      end do
      do i = 1, 3 * n
         do j = i + 1, 3 * n
            hess(i, j) = 0.5_wp * (hess(i, j) + hess(j, i))
            hess(j, i) = hess(i, j)
         end do
      end do
      hess_out(1:9 * n * n) = reshape(hess, [9 * n * n])
      if (present(c_dipgrad)) c_dipgrad(1:9 * n) = reshape(dipgrad, [9 * n])
      if (present(c_polgrad)) c_polgrad(1:18 * n) = 0.0_wp
      if (.not. conv) call env_error(env, "Self consistent charge iterator did not converge", 'scf')
   end subroutine hessian_api

   subroutine clear_result(er)
      type(engine_result), intent(inout) :: er
      er%gradient = c_null_ptr; er%charges = c_null_ptr; er%qsh = c_null_ptr; er%dipm = c_null_ptr
      er%qp = c_null_ptr; er%wbo = c_null_ptr; er%emo = c_null_ptr; er%focc = c_null_ptr
      er%coeff = c_null_ptr; er%density = c_null_ptr
   end subroutine clear_result

   ! --------------------------------------------------------------------------
   ! results (src/api/results.f90:53-501)

   function newResults_api() result(vres) bind(C, name="xtb_newResults")
      type(c_ptr) :: vres
      type(results_t), pointer :: res
      allocate(res)
      vres = c_loc(res)
   end function newResults_api

   subroutine delResults_api(vres) bind(C, name="xtb_delResults")
      type(c_ptr), intent(inout) :: vres
      type(results_t), pointer :: res
      if (c_associated(vres)) then
         call c_f_pointer(vres, res)
         call drop_engine(res)
         deallocate(res)
         vres = c_null_ptr
      end if
   end subroutine delResults_api

   function copyResults_api(vold) result(vres) bind(C, name="xtb_copyResults")
      type(c_ptr), value :: vold
      type(c_ptr) :: vres
      type(results_t), pointer :: old, res
      vres = c_null_ptr
      if (.not. c_associated(vold)) return
      call c_f_pointer(vold, old)
      allocate(res)
      res = old
      ! the copy keeps the wavefunction on the host; the engine stays with the original
      res%engine = c_null_ptr
      if (allocated(res%eng_at)) deallocate(res%eng_at)
      vres = c_loc(res)
   end function copyResults_api

   logical function res_checks(venv, vres, source, env, res) result(ok)
      type(c_ptr), value :: venv, vres
      character(len=*), intent(in) :: source
      type(environment_t), pointer, intent(out) :: env
      type(results_t), pointer, intent(out) :: res
      ok = .false.
      nullify(env, res)
      if (.not. c_associated(venv)) return
      call c_f_pointer(venv, env)
      if (.not. c_associated(vres)) then
         call env_error(env, "Results object is not allocated", source)
         return
      end if
      call c_f_pointer(vres, res)
      ok = .true.
   end function res_checks

   subroutine getEnergy_api(venv, vres, dptr) bind(C, name="xtb_getEnergy")
      type(c_ptr), value :: venv, vres
      real(c_double), intent(inout) :: dptr
      character(len=*), parameter :: source = "xtb_api_getEnergy"
      type(environment_t), pointer :: env
      type(results_t), pointer :: res
      if (.not. res_checks(venv, vres, source, env, res)) return
      if (.not. res%has_energy) then
         call env_error(env, "Energy is not available in results", source)
         return
      end if
      dptr = res%energy
   end subroutine getEnergy_api

   subroutine getSolvationEnergy_api(venv, vres, dptr) bind(C, name="xtb_getSolvationEnergy")
      type(c_ptr), value :: venv, vres
      real(c_double), intent(inout) :: dptr
      character(len=*), parameter :: source = "xtb_api_getSolvationEnergy"
      type(environment_t), pointer :: env
      type(results_t), pointer :: res
      if (.not. res_checks(venv, vres, source, env, res)) return
      call env_error(env, "Solvation energy is not available in results", source)
   end subroutine getSolvationEnergy_api

   subroutine getGradient_api(venv, vres, dptr) bind(C, name="xtb_getGradient")
      type(c_ptr), value :: venv, vres
      real(c_double), intent(inout) :: dptr(3, *)
      character(len=*), parameter :: source = "xtb_api_getGradient"
      type(environment_t), pointer :: env
      type(results_t), pointer :: res
      if (.not. res_checks(venv, vres, source, env, res)) return
      if (.not. res%has_gradient) then
         call env_error(env, "Gradient is not available in results", source)
         return
      end if
      dptr(1:3, 1:res%n) = res%gradient
   end subroutine getGradient_api

   subroutine getPCGradient_api(venv, vres, dptr) bind(C, name="xtb_getPCGradient")
      type(c_ptr), value :: venv, vres
      real(c_double), intent(inout) :: dptr(3, *)
      character(len=*), parameter :: source = "xtb_api_getPCGradient"
      type(environment_t), pointer :: env
      type(results_t), pointer :: res
      if (.not. res_checks(venv, vres, source, env, res)) return
      call env_error(env, "External charge gradients are not available in results", source)
   end subroutine getPCGradient_api

   subroutine getVirial_api(venv, vres, dptr) bind(C, name="xtb_getVirial")
      type(c_ptr), value :: venv, vres
      real(c_double), intent(inout) :: dptr(3, 3)
      character(len=*), parameter :: source = "xtb_api_getVirial"
      type(environment_t), pointer :: env
      type(results_t), pointer :: res
      if (.not. res_checks(venv, vres, source, env, res)) return
      if (.not. res%has_sigma) then
         call env_error(env, "Virial is not available in results", source)
         return
      end if
      dptr = res%sigma
   end subroutine getVirial_api

   subroutine getDipole_api(venv, vres, dptr) bind(C, name="xtb_getDipole")
      type(c_ptr), value :: venv, vres
      real(c_double), intent(inout) :: dptr(3)
      character(len=*), parameter :: source = "xtb_api_getDipole"
      type(environment_t), pointer :: env
      type(results_t), pointer :: res
      if (.not. res_checks(venv, vres, source, env, res)) return
      if (.not. res%has_dipole) then
         call env_error(env, "Dipole moment is not available in results", source)
         return
      end if
      dptr = res%dipole
   end subroutine getDipole_api

   subroutine getCharges_api(venv, vres, dptr) bind(C, name="xtb_getCharges")
      type(c_ptr), value :: venv, vres
      real(c_double), intent(inout) :: dptr(*)
      character(len=*), parameter :: source = "xtb_api_getCharges"
      type(environment_t), pointer :: env
      type(results_t), pointer :: res
      if (.not. res_checks(venv, vres, source, env, res)) return
      if (.not. res%has_wfn) then
         call env_error(env, "Partial charges are not available in results", source)
         return
      end if
      dptr(1:res%n) = res%q
   end subroutine getCharges_api

   subroutine getBondOrders_api(venv, vres, dptr) bind(C, name="xtb_getBondOrders")
      type(c_ptr), value :: venv, vres
      real(c_double), intent(inout) :: dptr(*)
      character(len=*), parameter :: source = "xtb_api_getBondOrders"
      type(environment_t), pointer :: env
      type(results_t), pointer :: res
      if (.not. res_checks(venv, vres, source, env, res)) return
      if (.not. res%has_wfn) then
         call env_error(env, "Bond orders are not available in results", source)
         return
      end if
      dptr(1:res%n * res%n) = reshape(res%wbo, [res%n * res%n])
   end subroutine getBondOrders_api

   subroutine getNao_api(venv, vres, iptr) bind(C, name="xtb_getNao")
      type(c_ptr), value :: venv, vres
      integer(c_int), intent(inout) :: iptr
      type(environment_t), pointer :: env
      type(results_t), pointer :: res
      if (.not. res_checks(venv, vres, "xtb_api_getNao", env, res)) return
      if (res%has_wfn) then
         iptr = res%nao
      else
         iptr = 0
      end if
   end subroutine getNao_api

   subroutine getOrbitalEigenvalues_api(venv, vres, dptr) bind(C, name="xtb_getOrbitalEigenvalues")
      type(c_ptr), value :: venv, vres
      real(c_double), intent(inout) :: dptr(*)
      character(len=*), parameter :: source = "xtb_api_getOrbitalEigenvalues"
      type(environment_t), pointer :: env
      type(results_t), pointer :: res
      if (.not. res_checks(venv, vres, source, env, res)) return
      if (.not. res%has_wfn) then
         call env_error(env, "Orbital eigenvalues are not available in results", source)
         return
      end if
      dptr(1:res%nao) = res%emo * evtoau
   end subroutine getOrbitalEigenvalues_api

   subroutine getOrbitalOccupations_api(venv, vres, dptr) bind(C, name="xtb_getOrbitalOccupations")
      type(c_ptr), value :: venv, vres
      real(c_double), intent(inout) :: dptr(*)
      character(len=*), parameter :: source = "xtb_api_getOrbitalOccupations"
      type(environment_t), pointer :: env
      type(results_t), pointer :: res
      if (.not. res_checks(venv, vres, source, env, res)) return
      if (.not. res%has_wfn) then
         call env_error(env, "Occupation numbers are not available in results", source)
         return
      end if
      dptr(1:res%nao) = res%focc
   end subroutine getOrbitalOccupations_api

   subroutine getOrbitalCoefficients_api(venv, vres, dptr) bind(C, name="xtb_getOrbitalCoefficients")
      type(c_ptr), value :: venv, vres
      real(c_double), intent(inout) :: dptr(*)
      character(len=*), parameter :: source = "xtb_api_getOrbitalCoefficients"
      type(environment_t), pointer :: env
      type(results_t), pointer :: res
      if (.not. res_checks(venv, vres, source, env, res)) return
      if (.not. res%has_wfn) then
         call env_error(env, "Orbital coefficients are not available in results", source)
         return
      end if
      dptr(1:res%nao * res%nao) = reshape(res%coeff, [res%nao * res%nao])
   end subroutine getOrbitalCoefficients_api

end module xtbamd_capi
