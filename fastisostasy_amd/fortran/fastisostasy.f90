! fastisostasy.f90 -- Fortran drop-in for the LV-ELVA path of FastIsostasy.
!
! Keeps the reference's public module name and procedures
! (src/fastisostasy.f90:34-60; bsl_* from src/barysealevel.f90:31-36):
!
!     call bsl_init(bsl, path_par, time)
!     call isos_init(isos, path_par, "isostasy", nx, ny, dx, dy)
!     call isos_init_state(isos, z_bed, H_ice, time, bsl)
!     call bsl_update(bsl, time)
!     call isos_update(isos, H_ice, time, bsl)
!     ... isos%out%z_bed ...
!     call isos_end(isos)
!
! Parameters are read from the same &isostasy / &barysealevel namelist files
! as the reference (fastisostasy.f90:925-1015, barysealevel.f90:40-193) with
! standard NAMELIST input, converted to the C ABI's fi_params (lengths in m)
! and handed to libfastiso.so through iso_c_binding.  A nonzero status of the
! library becomes `error stop` with the library's message, as the reference
! stops.  Arrays cross the boundary as real(c_double); with wp = sp the shim
! converts at the boundary.
module fastisostasy

    use, intrinsic :: iso_c_binding
    implicit none

    private

    integer, parameter, public :: sp = kind(1.0)
    integer, parameter, public :: dp = kind(1.d0)
    ! working precision of the caller's arrays, as the reference's precision
    ! module (precision.f90:10): sp when built with -DFI_WP_SP, else dp
#ifdef FI_WP_SP
    integer, parameter, public :: wp = sp
#else
    integer, parameter, public :: wp = dp
#endif

    integer, parameter :: FI_MAXL = 8

    type, bind(C) :: fi_params
        integer(c_int) :: heterogeneous_ssh, interactive_sealevel, include_elastic
        integer(c_int) :: correct_compression, correct_distortion
        integer(c_int) :: method
        real(c_double) :: dt_diagnostics, min_pad
        integer(c_int) :: mantle, lithosphere, layering, lumping, viscosity_scaling_method
        real(c_double) :: viscosity_scaling
        integer(c_int) :: nl
        real(c_double) :: zl(FI_MAXL), viscosities(FI_MAXL)
        real(c_double) :: dl, eta_ref
        real(c_double) :: rho_water, rho_ice, rho_seawater, rho_uppermantle, rho_litho
        real(c_double) :: E
        real(c_double) :: nu, g, r_earth, m_earth
        integer(c_int) :: dt_method
        real(c_double) :: atol, rtol, dt_min, dt_max, dt_init, q
        integer(c_int) :: precision
        real(c_double) :: tau
    end type

    type, public :: isos_param_class
        real(wp) :: rho_ice, rho_uppermantle, rho_seawater, rho_water, g, nu, E
        character(len=56) :: dt_method
        integer :: precision = 0
        character(len=256) :: restart = "None"
        logical :: use_restart = .false.
    end type

    type, public :: isos_domain_class
        integer :: nx_ice, ny_ice, nx, ny
        real(wp) :: dx, dy
        real(wp), allocatable :: D_lith(:, :)
    end type

    type, public :: isos_state_class
        real(wp) :: bsl, deltaV_bsl, V_af, V_den
    end type

    type, public :: ode_class
        real(wp) :: t, dt
    end type

    type, public :: isos_out_class
        real(wp), allocatable :: Hice(:, :), canom_full(:, :), dwdt(:, :), w(:, :), we(:, :)
        real(wp), allocatable :: z_ss(:, :), dz_ss(:, :), z_bed(:, :)
        logical, allocatable :: maskocean(:, :), maskgrounded(:, :), maskcontinent(:, :)
    end type

    type, public :: isos_class
        type(c_ptr) :: h = c_null_ptr
        type(isos_param_class) :: par
        type(isos_domain_class) :: domain
        type(isos_state_class) :: now
        type(isos_out_class) :: out
        type(ode_class) :: ode
    end type

    type, public :: bsl_class
        type(c_ptr) :: h = c_null_ptr
        character(len=56) :: method
        real(wp) :: bsl_init, bsl_now, A_ocean_pd, A_ocean_now, time
    end type

    public :: isos_init, isos_init_ref, isos_init_state, isos_update, isos_end
    public :: isos_restart_write, isos_restart_read
    public :: bsl_init, bsl_update, bsl_end, bsl_restart_write
    public :: isos_write_init, isos_write_step, bsl_write_init, bsl_write_step
    public :: fi_select_precision

    integer, parameter :: FI_Z_BED = 0, FI_Z_SS = 1, FI_DZ_SS = 2, FI_W = 3, FI_WE = 4, FI_DWDT = 5
    integer, parameter :: FI_CANOM_FULL = 6, FI_HICE = 7, FI_MASKOCEAN = 8, FI_MASKGROUNDED = 9
    integer, parameter :: FI_MASKCONTINENT = 10, FI_D_LITH = 14
    integer, parameter :: FI_S_N = 0, FI_S_T = 1, FI_S_DT = 2, FI_S_BSL = 3, FI_S_DELTAV = 4
    integer, parameter :: FI_S_V_AF = 5, FI_S_V_DEN = 6

    integer, save :: default_precision = 0

    interface
        function c_fi_last_error() bind(C, name="fi_last_error") result(p)
            import :: c_ptr
            type(c_ptr) :: p
        end function
        function c_fi_isos_create(p, nx, ny, dx, dy, K, rho, he, eta, out) &
                bind(C, name="fi_isos_create") result(rc)
            import :: c_int, c_double, c_ptr, fi_params
            type(fi_params), intent(in) :: p
            integer(c_int), value :: nx, ny
            real(c_double), value :: dx, dy
            type(c_ptr), value :: K, rho, he, eta
            type(c_ptr), intent(out) :: out
            integer(c_int) :: rc
        end function
        function c_fi_isos_init_ref(h, z_bed, H_ice, bsl, dz_ss) bind(C, name="fi_isos_init_ref") result(rc)
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: h
            real(c_double), intent(in) :: z_bed(*), H_ice(*)
            type(c_ptr), value :: bsl, dz_ss
            integer(c_int) :: rc
        end function
        function c_fi_isos_init_state(h, z_bed, H_ice, time, b) bind(C, name="fi_isos_init_state") result(rc)
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: h
            real(c_double), intent(in) :: z_bed(*), H_ice(*)
            real(c_double), value :: time
            type(c_ptr), value :: b
            integer(c_int) :: rc
        end function
        function c_fi_isos_update(h, H_ice, time, b) bind(C, name="fi_isos_update") result(rc)
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: h
            real(c_double), intent(in) :: H_ice(*)
            real(c_double), value :: time
            type(c_ptr), value :: b
            integer(c_int) :: rc
        end function
        function c_fi_isos_get(h, field, dst) bind(C, name="fi_isos_get") result(rc)
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: h
            integer(c_int), value :: field
            real(c_double), intent(out) :: dst(*)
            integer(c_int) :: rc
        end function
        function c_fi_isos_get_full(h, field, dst) bind(C, name="fi_isos_get_full") result(rc)
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: h
            integer(c_int), value :: field
            real(c_double), intent(out) :: dst(*)
            integer(c_int) :: rc
        end function
        function c_fi_isos_get_scalar(h, which, v) bind(C, name="fi_isos_get_scalar") result(rc)
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: h
            integer(c_int), value :: which
            real(c_double), intent(out) :: v
            integer(c_int) :: rc
        end function
        function c_fi_isos_destroy(h) bind(C, name="fi_isos_destroy") result(rc)
            import :: c_int, c_ptr
            type(c_ptr), value :: h
            integer(c_int) :: rc
        end function
        function c_fi_bsl_create(method, bsl_init, A_pd, nvec, bvec, avec, year_bp, out) &
                bind(C, name="fi_bsl_create") result(rc)
            import :: c_int, c_ptr, c_double
            integer(c_int), value :: method, nvec
            real(c_double), value :: bsl_init, A_pd, year_bp
            type(c_ptr), value :: bvec, avec
            type(c_ptr), intent(out) :: out
            integer(c_int) :: rc
        end function
        function c_fi_bsl_create_series(method, bsl_init, A_pd, nvec, bvec, avec, ns, st, sv, year_bp, out) &
                bind(C, name="fi_bsl_create_series") result(rc)
            import :: c_int, c_ptr, c_double
            integer(c_int), value :: method, nvec, ns
            real(c_double), value :: bsl_init, A_pd, year_bp
            type(c_ptr), value :: bvec, avec
            real(c_double), intent(in) :: st(*), sv(*)
            type(c_ptr), intent(out) :: out
            integer(c_int) :: rc
        end function
        function c_fi_bsl_read_series(path, max_n, t, v, n) bind(C, name="fi_bsl_read_series") result(rc)
            import :: c_int, c_double, c_char
            character(kind=c_char), intent(in) :: path(*)
            integer(c_int), value :: max_n
            real(c_double), intent(out) :: t(*), v(*)
            integer(c_int), intent(out) :: n
            integer(c_int) :: rc
        end function
        function c_fi_isos_restart_write(h, path, time) bind(C, name="fi_isos_restart_write") result(rc)
            import :: c_int, c_ptr, c_double, c_char
            type(c_ptr), value :: h
            character(kind=c_char), intent(in) :: path(*)
            real(c_double), value :: time
            integer(c_int) :: rc
        end function
        function c_fi_isos_restart_read(h, path, time) bind(C, name="fi_isos_restart_read") result(rc)
            import :: c_int, c_ptr, c_double, c_char
            type(c_ptr), value :: h
            character(kind=c_char), intent(in) :: path(*)
            real(c_double), value :: time
            integer(c_int) :: rc
        end function
        function c_fi_isos_write_init(h, xc, yc, path, time) bind(C, name="fi_isos_write_init") result(rc)
            import :: c_int, c_ptr, c_double, c_char
            type(c_ptr), value :: h
            real(c_double), intent(in) :: xc(*), yc(*)
            character(kind=c_char), intent(in) :: path(*)
            real(c_double), value :: time
            integer(c_int) :: rc
        end function
        function c_fi_isos_write_step(h, path, time, hice, zb) bind(C, name="fi_isos_write_step") result(rc)
            import :: c_int, c_ptr, c_double, c_char
            type(c_ptr), value :: h, hice, zb
            character(kind=c_char), intent(in) :: path(*)
            real(c_double), value :: time
            integer(c_int) :: rc
        end function
        function c_fi_bsl_write_init(b, path, time) bind(C, name="fi_bsl_write_init") result(rc)
            import :: c_int, c_ptr, c_double, c_char
            type(c_ptr), value :: b
            character(kind=c_char), intent(in) :: path(*)
            real(c_double), value :: time
            integer(c_int) :: rc
        end function
        function c_fi_bsl_write_step(b, path, time) bind(C, name="fi_bsl_write_step") result(rc)
            import :: c_int, c_ptr, c_double, c_char
            type(c_ptr), value :: b
            character(kind=c_char), intent(in) :: path(*)
            real(c_double), value :: time
            integer(c_int) :: rc
        end function
        function c_fi_bsl_restart_write(b, path, time) bind(C, name="fi_bsl_restart_write") result(rc)
            import :: c_int, c_ptr, c_double, c_char
            type(c_ptr), value :: b
            character(kind=c_char), intent(in) :: path(*)
            real(c_double), value :: time
            integer(c_int) :: rc
        end function
        function c_fi_bsl_restart_read(path, b, a) bind(C, name="fi_bsl_restart_read") result(rc)
            import :: c_int, c_double, c_char
            character(kind=c_char), intent(in) :: path(*)
            real(c_double), intent(out) :: b, a
            integer(c_int) :: rc
        end function
        function c_fi_bsl_update(b, year_bp) bind(C, name="fi_bsl_update") result(rc)
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: b
            real(c_double), value :: year_bp
            integer(c_int) :: rc
        end function
        function c_fi_bsl_get(b, bsl_now, A_now) bind(C, name="fi_bsl_get") result(rc)
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: b
            real(c_double), intent(out) :: bsl_now, A_now
            integer(c_int) :: rc
        end function
        function c_fi_bsl_destroy(b) bind(C, name="fi_bsl_destroy") result(rc)
            import :: c_int, c_ptr
            type(c_ptr), value :: b
            integer(c_int) :: rc
        end function
        function c_strlen(s) bind(C, name="strlen") result(n)
            import :: c_ptr, c_size_t
            type(c_ptr), value :: s
            integer(c_size_t) :: n
        end function
    end interface

contains

    ! Precision of the RHS for handles created afterwards: 0 FP64 (default), 1 mixed.
    subroutine fi_select_precision(p)
        integer, intent(IN) :: p
        default_precision = p
    end subroutine fi_select_precision

    subroutine check(rc, where)
        integer(c_int), intent(IN) :: rc
        character(len=*), intent(IN) :: where
        character(kind=c_char), pointer :: msg(:)
        character(len=512) :: text
        type(c_ptr) :: p
        integer :: i, n
        if (rc == 0) return
        p = c_fi_last_error()
        n = int(c_strlen(p))
        call c_f_pointer(p, msg, [n])
        text = ""
        do i = 1, min(n, 512)
            text(i:i) = msg(i)
        end do
        write(*, *) trim(where), ": ", trim(text)
        error stop "fastisostasy: library error"
    end subroutine check

    integer function pick(name, opts) result(k)
        character(len=*), intent(IN) :: name
        character(len=*), intent(IN) :: opts(:)
        integer :: i
        do i = 1, size(opts)
            if (trim(name) == trim(opts(i))) then
                k = i - 1
                return
            end if
        end do
        write(*, *) "fastisostasy: unknown option ", trim(name)
        error stop "fastisostasy: unsupported option"
    end function pick

    ! isos_par_load (src/fastisostasy.f90:925-1015) with NAMELIST input
    subroutine read_isostasy_nml(filename, group, p, restart_path)
        character(len=*), intent(IN) :: filename, group
        type(fi_params), intent(OUT) :: p
        character(len=256), intent(OUT) :: restart_path
        logical :: heterogeneous_ssh, interactive_sealevel, include_elastic
        logical :: correct_compression, correct_distortion
        integer :: method, nl, u, ios
        real(dp) :: dt_diagnostics, pad, viscosity_scaling, ref_viscosity, tau
        real(dp) :: rheo_smoothing_radius, litho_thickness_scaling, dl
        real(dp) :: zl(FI_MAXL), viscosities(FI_MAXL)
        real(dp) :: E, nu, rho_ice, rho_seawater, rho_water, rho_uppermantle, rho_litho, g, r_earth, m_earth
        real(dp) :: dt_init, dt_min, dt_max, rtol, atol, q
        character(len=256) :: mantle, lithosphere, layering, lumping, viscosity_scaling_method
        character(len=256) :: restart, mask_file, rheology_file, dt_method
        character(len=256) :: litho_thickness_varname, log10_viscosity_varname, stddev_log10_viscosity_varname
        namelist /isostasy/ heterogeneous_ssh, interactive_sealevel, include_elastic, correct_compression, &
            correct_distortion, method, dt_diagnostics, pad, mantle, lithosphere, layering, lumping, &
            viscosity_scaling_method, viscosity_scaling, nl, zl, viscosities, ref_viscosity, tau, &
            rheo_smoothing_radius, litho_thickness_scaling, dl, E, nu, rho_ice, rho_seawater, rho_water, &
            rho_uppermantle, rho_litho, g, r_earth, m_earth, restart, mask_file, rheology_file, &
            litho_thickness_varname, log10_viscosity_varname, stddev_log10_viscosity_varname, &
            dt_method, dt_init, dt_min, dt_max, rtol, atol, q

        if (trim(group) /= "isostasy") error stop "fastisostasy: only the &isostasy group is supported"
        heterogeneous_ssh = .false.; interactive_sealevel = .false.; include_elastic = .false.
        correct_compression = .false.; correct_distortion = .false.
        method = 3; nl = 1; dt_diagnostics = 10.0_dp; pad = 0.0_dp; viscosity_scaling = 1.0_dp
        ref_viscosity = 1.0e21_dp; tau = 3000.0_dp; rheo_smoothing_radius = 0.0_dp
        litho_thickness_scaling = 1.0_dp; dl = 0.0_dp; zl = 0.0_dp; viscosities = 0.0_dp
        mantle = "uniform"; lithosphere = "uniform"; layering = "equalized"; lumping = "freqdomain"
        viscosity_scaling_method = "scale"; restart = "None"; mask_file = "None"; rheology_file = "None"
        dt_method = "tsit54"
        open(newunit=u, file=filename, status="old", action="read", iostat=ios)
        if (ios /= 0) error stop "fastisostasy: cannot open parameter file"
        read(u, nml=isostasy, iostat=ios)
        close(u)
        if (ios /= 0) error stop "fastisostasy: cannot read the &isostasy namelist"
        restart_path = restart
        if (trim(mask_file) /= "None" .and. trim(mask_file) /= "none" .and. trim(mask_file) /= "no") &
            error stop "fastisostasy: mask files are not supported"
        if (nl < 1 .or. nl > FI_MAXL) error stop "fastisostasy: nl out of range"

        p%heterogeneous_ssh = merge(1, 0, heterogeneous_ssh)
        p%interactive_sealevel = merge(1, 0, interactive_sealevel)
        p%include_elastic = merge(1, 0, include_elastic)
        p%correct_compression = merge(1, 0, correct_compression)
        p%correct_distortion = merge(1, 0, correct_distortion)
        p%method = method
        p%dt_diagnostics = dt_diagnostics
        p%min_pad = pad * 1e3_dp
        p%mantle = pick(mantle, [character(len=16) :: "uniform", "gaussian_plus", "gaussian_minus"])
        p%lithosphere = pick(lithosphere, [character(len=16) :: "uniform", "gaussian_plus", "gaussian_minus"])
        p%layering = pick(layering, [character(len=16) :: "parallel", "equalized"])
        p%lumping = pick(lumping, [character(len=16) :: "bottom", "min", "max", "mean", "logmean", "freqdomain"])
        select case (trim(viscosity_scaling_method))
        case ("min"); p%viscosity_scaling_method = 1
        case ("scale"); p%viscosity_scaling_method = 2
        case ("max"); p%viscosity_scaling_method = 3
        case ("stddev"); p%viscosity_scaling_method = 4
        case default; p%viscosity_scaling_method = 0
        end select
        p%viscosity_scaling = viscosity_scaling
        p%nl = nl
        p%zl = zl * 1e3_dp
        p%viscosities = viscosities
        p%dl = dl * 1e3_dp
        p%eta_ref = ref_viscosity
        p%rho_water = rho_water; p%rho_ice = rho_ice; p%rho_seawater = rho_seawater
        p%rho_uppermantle = rho_uppermantle; p%rho_litho = rho_litho
        p%E = E; p%nu = nu; p%g = g; p%r_earth = r_earth * 1e3_dp; p%m_earth = m_earth
        p%dt_method = pick(dt_method, [character(len=16) :: "euler", "rk4", "bs32", "tsit54"])
        p%atol = atol; p%rtol = rtol; p%dt_min = dt_min; p%dt_max = dt_max; p%dt_init = dt_init; p%q = q
        p%precision = default_precision
        p%tau = tau
    end subroutine read_isostasy_nml

    ! isos_init (src/fastisostasy.f90:64-473)
    subroutine isos_init(isos, filename, group, nx_ice, ny_ice, dx, dy, rho_ice, K)
        type(isos_class), intent(INOUT) :: isos
        character(len=*), intent(IN) :: filename, group
        integer, intent(IN) :: nx_ice, ny_ice
        real(wp), intent(IN) :: dx, dy
        real(wp), intent(IN), optional :: rho_ice
        real(wp), intent(IN), optional, target :: K(:, :)
        type(fi_params), target :: p
        real(c_double), target :: rho
        real(c_double), allocatable, target :: Kd(:, :)
        type(c_ptr) :: prho, pK
        real(c_double) :: v
        integer :: n

        call read_isostasy_nml(filename, group, p, isos%par%restart)
        ! par%use_restart (fastisostasy.f90:997-1004)
        isos%par%use_restart = .not. (trim(isos%par%restart) == "None" .or. trim(isos%par%restart) == "none" &
            .or. trim(isos%par%restart) == "no")
        prho = c_null_ptr
        pK = c_null_ptr
        if (present(rho_ice)) then
            rho = real(rho_ice, c_double)
            prho = c_loc(rho)
        end if
        if (present(K)) then
            allocate(Kd(size(K, 1), size(K, 2)))
            Kd = real(K, c_double)
            pK = c_loc(Kd)
        end if
        call check(c_fi_isos_create(p, int(nx_ice, c_int), int(ny_ice, c_int), real(dx, c_double), &
            real(dy, c_double), pK, prho, c_null_ptr, c_null_ptr, isos%h), "isos_init")
        call check(c_fi_isos_get_scalar(isos%h, FI_S_N, v), "isos_init")
        n = nint(v)
        isos%domain%nx_ice = nx_ice; isos%domain%ny_ice = ny_ice
        isos%domain%nx = n; isos%domain%ny = n
        isos%domain%dx = dx; isos%domain%dy = dy
        allocate(isos%domain%D_lith(n, n))
        call get_full(isos, FI_D_LITH, isos%domain%D_lith)
        isos%par%rho_ice = p%rho_ice; isos%par%rho_uppermantle = p%rho_uppermantle
        isos%par%rho_seawater = p%rho_seawater; isos%par%rho_water = p%rho_water
        isos%par%g = p%g; isos%par%nu = p%nu; isos%par%E = p%E
        isos%par%precision = p%precision
        allocate(isos%out%Hice(nx_ice, ny_ice), isos%out%canom_full(nx_ice, ny_ice), &
            isos%out%dwdt(nx_ice, ny_ice), isos%out%w(nx_ice, ny_ice), isos%out%we(nx_ice, ny_ice), &
            isos%out%z_ss(nx_ice, ny_ice), isos%out%dz_ss(nx_ice, ny_ice), isos%out%z_bed(nx_ice, ny_ice), &
            isos%out%maskocean(nx_ice, ny_ice), isos%out%maskgrounded(nx_ice, ny_ice), &
            isos%out%maskcontinent(nx_ice, ny_ice))
    end subroutine isos_init

    subroutine get_full(isos, field, dst)
        type(isos_class), intent(IN) :: isos
        integer, intent(IN) :: field
        real(wp), intent(OUT) :: dst(:, :)
        real(c_double), allocatable :: tmp(:, :)
        allocate(tmp(size(dst, 1), size(dst, 2)))
        call check(c_fi_isos_get_full(isos%h, int(field, c_int), tmp), "get_full")
        dst = real(tmp, wp)
    end subroutine get_full

    subroutine get_out(isos, field, dst)
        type(isos_class), intent(IN) :: isos
        integer, intent(IN) :: field
        real(wp), intent(OUT) :: dst(:, :)
        real(c_double), allocatable :: tmp(:, :)
        allocate(tmp(size(dst, 1), size(dst, 2)))
        call check(c_fi_isos_get(isos%h, int(field, c_int), tmp), "get")
        dst = real(tmp, wp)
    end subroutine get_out

    subroutine get_mask(isos, field, dst)
        type(isos_class), intent(IN) :: isos
        integer, intent(IN) :: field
        logical, intent(OUT) :: dst(:, :)
        real(c_double), allocatable :: tmp(:, :)
        allocate(tmp(size(dst, 1), size(dst, 2)))
        call check(c_fi_isos_get(isos%h, int(field, c_int), tmp), "get")
        dst = tmp /= 0.0_c_double
    end subroutine get_mask

    ! cropstate2out (src/isos_utils.f90:386-405) and the scalars callers read
    subroutine refresh(isos)
        type(isos_class), intent(INOUT) :: isos
        real(c_double) :: v
        call get_out(isos, FI_HICE, isos%out%Hice)
        call get_out(isos, FI_Z_SS, isos%out%z_ss)
        call get_out(isos, FI_Z_BED, isos%out%z_bed)
        call get_out(isos, FI_DWDT, isos%out%dwdt)
        call get_out(isos, FI_W, isos%out%w)
        call get_out(isos, FI_WE, isos%out%we)
        call get_out(isos, FI_DZ_SS, isos%out%dz_ss)
        call get_out(isos, FI_CANOM_FULL, isos%out%canom_full)
        call get_mask(isos, FI_MASKOCEAN, isos%out%maskocean)
        call get_mask(isos, FI_MASKGROUNDED, isos%out%maskgrounded)
        call get_mask(isos, FI_MASKCONTINENT, isos%out%maskcontinent)
        call check(c_fi_isos_get_scalar(isos%h, FI_S_T, v), "scalar"); isos%ode%t = v
        call check(c_fi_isos_get_scalar(isos%h, FI_S_DT, v), "scalar"); isos%ode%dt = v
        call check(c_fi_isos_get_scalar(isos%h, FI_S_BSL, v), "scalar"); isos%now%bsl = v
        call check(c_fi_isos_get_scalar(isos%h, FI_S_DELTAV, v), "scalar"); isos%now%deltaV_bsl = v
        call check(c_fi_isos_get_scalar(isos%h, FI_S_V_AF, v), "scalar"); isos%now%V_af = v
        call check(c_fi_isos_get_scalar(isos%h, FI_S_V_DEN, v), "scalar"); isos%now%V_den = v
    end subroutine refresh

    subroutine bsl_refresh(bsl)
        type(bsl_class), intent(INOUT) :: bsl
        real(c_double) :: b, a
        call check(c_fi_bsl_get(bsl%h, b, a), "bsl")
        bsl%bsl_now = b
        bsl%A_ocean_now = a
    end subroutine bsl_refresh

    ! isos_init_ref (src/fastisostasy.f90:550-575)
    subroutine isos_init_ref(isos, z_bed, H_ice, bsl, dz_ss)
        type(isos_class), intent(INOUT) :: isos
        real(wp), intent(IN) :: z_bed(:, :), H_ice(:, :)
        real(wp), intent(IN), optional :: bsl
        real(wp), intent(IN), optional :: dz_ss(:, :)
        real(c_double), target :: b
        real(c_double), allocatable, target :: dz(:, :)
        type(c_ptr) :: pb, pdz
        pb = c_null_ptr
        pdz = c_null_ptr
        if (present(bsl)) then
            b = real(bsl, c_double)
            pb = c_loc(b)
        end if
        if (present(dz_ss)) then
            allocate(dz(size(dz_ss, 1), size(dz_ss, 2)))
            dz = real(dz_ss, c_double)
            pdz = c_loc(dz)
        end if
        call check(c_fi_isos_init_ref(isos%h, real(z_bed, c_double), real(H_ice, c_double), pb, pdz), &
            "isos_init_ref")
    end subroutine isos_init_ref

    ! isos_init_state (src/fastisostasy.f90:596-660)
    subroutine isos_init_state(isos, z_bed, H_ice, time, bsl)
        type(isos_class), intent(INOUT) :: isos
        real(wp), intent(IN) :: z_bed(:, :), H_ice(:, :)
        real(wp), intent(IN) :: time
        type(bsl_class), intent(INOUT) :: bsl
        if (isos%par%use_restart) call isos_restart_read(isos, trim(isos%par%restart), time)
        call check(c_fi_isos_init_state(isos%h, real(z_bed, c_double), real(H_ice, c_double), &
            real(time, c_double), bsl%h), "isos_init_state")
        call refresh(isos)
        call bsl_refresh(bsl)
    end subroutine isos_init_state

    ! isos_restart_write / isos_restart_read (src/isostasy_io.f90:56-194): binary
    ! container of the same variables (the library's restart format)
    subroutine isos_restart_write(isos, filename, time, init)
        type(isos_class), intent(IN) :: isos
        character(len=*), intent(IN) :: filename
        real(wp), intent(IN) :: time
        logical, optional, intent(IN) :: init
        call check(c_fi_isos_restart_write(isos%h, trim(filename)//c_null_char, real(time, c_double)), &
            "isos_restart_write")
    end subroutine isos_restart_write

    subroutine isos_restart_read(isos, filename, time)
        type(isos_class), intent(INOUT) :: isos
        character(len=*), intent(IN) :: filename
        real(wp), intent(IN) :: time
        call check(c_fi_isos_restart_read(isos%h, trim(filename)//c_null_char, real(time, c_double)), &
            "isos_restart_read")
    end subroutine isos_restart_read

    ! bsl_restart_write (src/barysealevel.f90:426-450)
    subroutine bsl_restart_write(bsl, filename, time, init)
        type(bsl_class), intent(IN) :: bsl
        character(len=*), intent(IN) :: filename
        real(wp), intent(IN) :: time
        logical, optional, intent(IN) :: init
        call check(c_fi_bsl_restart_write(bsl%h, trim(filename)//c_null_char, real(time, c_double)), &
            "bsl_restart_write")
    end subroutine bsl_restart_write

    ! isos_write_init / isos_write_step (src/isostasy_io.f90:197-405) and
    ! bsl_write_init / bsl_write_step (src/barysealevel.f90:453-494): the
    ! library's output container (the reference's variables, not NetCDF)
    subroutine isos_write_init(isos, xc, yc, filename, time_init)
        type(isos_class), intent(IN) :: isos
        real(wp), intent(IN) :: xc(:), yc(:)
        character(len=*), intent(IN) :: filename
        real(wp), intent(IN) :: time_init
        if (size(xc) /= isos%domain%nx_ice .or. size(yc) /= isos%domain%ny_ice) &
            error stop "isos_write_init:: xc / yc do not match the ice grid"
        call check(c_fi_isos_write_init(isos%h, real(xc, c_double), real(yc, c_double), &
            trim(filename)//c_null_char, real(time_init, c_double)), "isos_write_init")
    end subroutine isos_write_init

    subroutine isos_write_step(isos, filename, time, H_ice, z_bed_bench)
        type(isos_class), intent(IN) :: isos
        character(len=*), intent(IN) :: filename
        real(wp), intent(IN) :: time
        real(wp), intent(IN) :: H_ice(:, :)
        real(wp), intent(IN), optional :: z_bed_bench(:, :)
        real(c_double), allocatable, target :: zb(:, :)
        type(c_ptr) :: pzb
        pzb = c_null_ptr
        if (present(z_bed_bench)) then
            allocate(zb(size(z_bed_bench, 1), size(z_bed_bench, 2)))
            zb = real(z_bed_bench, c_double)
            pzb = c_loc(zb)
        end if
        call check(c_fi_isos_write_step(isos%h, trim(filename)//c_null_char, real(time, c_double), &
            c_null_ptr, pzb), "isos_write_step")
    end subroutine isos_write_step

    subroutine bsl_write_init(bsl, filename, time_init)
        type(bsl_class), intent(IN) :: bsl
        character(len=*), intent(IN) :: filename
        real(wp), intent(IN) :: time_init
        call check(c_fi_bsl_write_init(bsl%h, trim(filename)//c_null_char, real(time_init, c_double)), &
            "bsl_write_init")
    end subroutine bsl_write_init

    subroutine bsl_write_step(bsl, filename, time)
        type(bsl_class), intent(IN) :: bsl
        character(len=*), intent(IN) :: filename
        real(wp), intent(IN) :: time
        call check(c_fi_bsl_write_step(bsl%h, trim(filename)//c_null_char, real(time, c_double)), &
            "bsl_write_step")
    end subroutine bsl_write_step

    ! isos_update (src/fastisostasy.f90:731-872); dwdt_corr is unused, as in the reference
    subroutine isos_update(isos, H_ice, time, bsl, dwdt_corr)
        type(isos_class), intent(INOUT) :: isos
        real(wp), intent(IN) :: H_ice(:, :)
        real(wp), intent(IN) :: time
        type(bsl_class), intent(INOUT) :: bsl
        real(wp), intent(IN), optional :: dwdt_corr(:, :)
        call check(c_fi_isos_update(isos%h, real(H_ice, c_double), real(time, c_double), bsl%h), &
            "isos_update")
        call refresh(isos)
        call bsl_refresh(bsl)
    end subroutine isos_update

    ! isos_end (src/fastisostasy.f90:911-923)
    subroutine isos_end(isos)
        type(isos_class), intent(INOUT) :: isos
        integer(c_int) :: rc
        if (c_associated(isos%h)) rc = c_fi_isos_destroy(isos%h)
        isos%h = c_null_ptr
        if (allocated(isos%domain%D_lith)) deallocate(isos%domain%D_lith)
    end subroutine isos_end

    ! bsl_init (src/barysealevel.f90:40-153): methods const, fastiso, file, mixed
    subroutine read_barysealevel_nml(filename, method, bsl0, A_pd, series_path)
        character(len=*), intent(IN) :: filename
        character(len=256), intent(OUT) :: method, series_path
        real(dp), intent(OUT) :: bsl0, A_pd
        character(len=256) :: sl_path, sl_name, A_ocean_path, restart
        real(dp) :: bsl_init, A_ocean_pd
        integer :: u, ios
        real(c_double) :: rb, ra
        namelist /barysealevel/ method, bsl_init, sl_path, sl_name, A_ocean_pd, A_ocean_path, restart
        method = "const"; bsl_init = 0.0_dp; A_ocean_pd = 3.625e14_dp
        sl_path = "None"; sl_name = "None"; A_ocean_path = "None"; restart = "None"
        open(newunit=u, file=filename, status="old", action="read", iostat=ios)
        if (ios /= 0) error stop "fastisostasy: cannot open parameter file"
        read(u, nml=barysealevel, iostat=ios)
        close(u)
        if (ios /= 0) error stop "fastisostasy: cannot read the &barysealevel namelist"
        if (trim(A_ocean_path) /= "None" .and. trim(A_ocean_path) /= "none" .and. trim(A_ocean_path) /= "no") &
            error stop "fastisostasy: A_ocean_path tables are not supported by the namelist shim"
        ! const with a restart file takes bsl_init from it (barysealevel.f90:57-73)
        if (trim(method) == "const" .and. .not. (trim(restart) == "None" .or. trim(restart) == "none" .or. &
                trim(restart) == "no")) then
            call check(c_fi_bsl_restart_read(trim(restart)//c_null_char, rb, ra), "bsl_init")
            bsl_init = rb
        end if
        bsl0 = bsl_init
        A_pd = A_ocean_pd
        series_path = sl_path
    end subroutine read_barysealevel_nml

    subroutine bsl_init(bsl, filename, year_bp)
        type(bsl_class), intent(INOUT) :: bsl
        character(len=*), intent(IN) :: filename
        real(wp), intent(IN) :: year_bp
        character(len=256) :: method, series_path
        real(dp) :: bsl0, A_pd
        integer, parameter :: nmax = 10000
        real(c_double), allocatable :: st(:), sv(:)
        integer(c_int) :: ns
        integer :: m, k
        call read_barysealevel_nml(filename, method, bsl0, A_pd, series_path)
        select case (trim(method))
        case ("const"); m = 0
        case ("fastiso"); m = 1
        case ("file"); m = 2
        case ("mixed"); m = 3
        case default; error stop "bsl_init:: Error: method not recognized."
        end select
        if (m <= 1) then
            call check(c_fi_bsl_create(int(m, c_int), real(bsl0, c_double), real(A_pd, c_double), 0_c_int, &
                c_null_ptr, c_null_ptr, real(year_bp, c_double), bsl%h), "bsl_init")
        else
            ! read_series (barysealevel.f90:254-315); NetCDF series need the netCDF library
            k = len_trim(series_path)
            if (k >= 2) then
                if (series_path(k-1:k) == "nc") error stop "fastisostasy: NetCDF sea-level series are not supported"
            end if
            allocate(st(nmax), sv(nmax))
            call check(c_fi_bsl_read_series(trim(series_path)//c_null_char, int(nmax, c_int), st, sv, ns), &
                "read_series")
            call check(c_fi_bsl_create_series(int(m, c_int), real(bsl0, c_double), real(A_pd, c_double), 0_c_int, &
                c_null_ptr, c_null_ptr, ns, st, sv, real(year_bp, c_double), bsl%h), "bsl_init")
        end if
        bsl%method = method
        bsl%A_ocean_pd = A_pd
        bsl%time = year_bp
        call bsl_refresh(bsl)
        bsl%bsl_init = bsl%bsl_now   ! bsl_now = bsl_init after bsl_init (barysealevel.f90:100)
    end subroutine bsl_init

    ! bsl_update (src/barysealevel.f90:195-228)
    subroutine bsl_update(bsl, year_bp)
        type(bsl_class), intent(INOUT) :: bsl
        real(wp), intent(IN) :: year_bp
        call check(c_fi_bsl_update(bsl%h, real(year_bp, c_double)), "bsl_update")
        bsl%time = year_bp
        call bsl_refresh(bsl)
    end subroutine bsl_update

    subroutine bsl_end(bsl)
        type(bsl_class), intent(INOUT) :: bsl
        integer(c_int) :: rc
        if (c_associated(bsl%h)) rc = c_fi_bsl_destroy(bsl%h)
        bsl%h = c_null_ptr
    end subroutine bsl_end

end module fastisostasy
