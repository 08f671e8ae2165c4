! drive_disc.f90 -- Fortran caller of the drop-in module, shaped like the
! reference driver (tests/test_isostasy.f90:350-402): disc load switched on
! at t = 1e-9 yr, coupling step 10 yr, bsl_update then isos_update.
!
!   drive_disc <par.nml> <nx> <dx_m> <years> <out.bin> [mixed]
!
! Writes the final out%w and out%z_bed (nx*nx doubles each, Fortran order)
! and prints one line per 1000 yr.
program drive_disc
    use fastisostasy
    implicit none
    type(isos_class) :: isos
    type(bsl_class) :: bsl
    character(len=512) :: par, outf, arg
    integer :: nx, n, nt, i, j, u
    real(wp) :: dx, years, time, xc, r2
    real(wp), allocatable :: z_bed(:, :), H0(:, :), H(:, :)

    call get_command_argument(1, par)
    call get_command_argument(2, arg); read(arg, *) nx
    call get_command_argument(3, arg); read(arg, *) dx
    call get_command_argument(4, arg); read(arg, *) years
    call get_command_argument(5, outf)
    call get_command_argument(6, arg)
    if (trim(arg) == "mixed") call fi_select_precision(1)

    allocate(z_bed(nx, nx), H0(nx, nx), H(nx, nx))
    z_bed = 1e6_wp
    H0 = 0.0_wp
    xc = (nx - 1) / 2.0_wp
    do j = 1, nx
        do i = 1, nx
            r2 = ((i - 1 - xc) * dx)**2 + ((j - 1 - xc) * dx)**2
            if (r2 <= (1000e3_wp)**2) H0(i, j) = 1000.0_wp
        end do
    end do

    time = 0.0_wp
    call bsl_init(bsl, trim(par), time)
    call isos_init(isos, trim(par), "isostasy", nx, nx, dx, dx)
    H = 0.0_wp
    call isos_init_state(isos, z_bed, H, time, bsl)
    nt = nint(years / 10.0_wp)
    do n = 1, nt
        time = n * 10.0_wp
        H = H0                      ! interp_2d of (0, 1e-9) -> (0, H0) for t > 1e-9
        call bsl_update(bsl, time)
        call isos_update(isos, H, time, bsl)
        if (mod(n, 100) == 0) write(*, '(a,f10.1,a,f12.5,a,f8.4)') "time ", time, "  min(w) ", &
            minval(isos%out%w), "  dt ", isos%ode%dt
    end do
    open(newunit=u, file=trim(outf), access="stream", form="unformatted", status="replace")
    write(u) isos%out%w
    write(u) isos%out%z_bed
    close(u)
    call isos_end(isos)
    call bsl_end(bsl)
end program drive_disc
