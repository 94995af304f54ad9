!*************************************************************************
! Host driver for the MI355X HERCULES time step.
!
! Reads parameters.input with the reference's four namelists
! (src/module_parameters.f90:145-157) and the initial fields
! results/init_{u,v,w,p,t}.dat (raw FP64, global nx*ny*nz, Fortran order,
! the 2DECOMP decomp_2d_write_one format read at src/module_tools.f90:273-286),
! runs imax steps of the drive loop (src/module_navier_stokes.f90:139-210)
! on the GPU and writes results/bak_{u,v,w,p,t}.dat every ibackup steps
! (src/module_io.f90:311-360; with AB2, dts = 1, also the right-hand-side
! history bak_{ff,gg,hh,ss}.dat, which a restart reads back from
! init_{ff,gg,hh,ss}.dat, src/module_tools.f90:292-307), the grid
! results/geo.info (src/module_parameters.f90:495-504) and the mean fields results/mean_*.dat every
! imeanfl steps (calculate_mean_fields summed on the GPU from the first
! step, written as output_mean_fields does, src/module_io.f90:546-598),
! the time series results/tmsr_{u,v,w,t,p}.dat every step when istmsr = 1
! (:211-305), the instantaneous fields results/inst_{u,v,w,p,t}NNN.dat every
! iinstfl steps (:366-403) and the x-y, x-z, y-z slices
! results/slices/{xy,xz,yz}2d_* every intv_2d steps (:408-542).
!*************************************************************************
program hercules_hip

    use, intrinsic :: iso_c_binding
    use module_hercules_hip
    implicit none

    ! namelist variables (same names, types and groups as the reference)
    integer :: p_row, p_col, nx, ny, nz, ochannel
    real(c_double) :: lx, ly, lz, zstretch
    integer :: dp_opt, cds, issk, dts, isnoise, isdamp, nzdamp
    real(c_double) :: noise_mag, cdamp
    real(c_double) :: nu, alpha, got, fc, ug, vg, dt, u_mrf, t_ref, tbot, ttop, ri_str, ri_end
    integer :: imax, isscalar, is_ri_var
    integer :: restart, istmsr, ibackup, iinstfl, imeanfl, isxy2d, isxz2d, isyz2d, intv_2d
    integer, dimension(50) :: xy2d_id, xz2d_id, yz2d_id

    namelist /domain/ p_row, p_col, lx, ly, lz, nx, ny, nz, zstretch, ochannel
    namelist /modeling/ dp_opt, cds, dts, isnoise, noise_mag, isdamp, nzdamp, cdamp, issk
    namelist /constants/ nu, alpha, got, fc, ug, vg, dt, imax, u_mrf, isscalar, t_ref, &
                         tbot, ttop, is_ri_var, ri_str, ri_end
    namelist /io/ restart, istmsr, ibackup, iinstfl, imeanfl, &
                  isxy2d, xy2d_id, isxz2d, xz2d_id, isyz2d, yz2d_id, intv_2d

    type(hc_params) :: prm
    real(c_double), allocatable, target :: f(:,:,:), g(:,:,:)
    real(c_double), allocatable, target :: zc(:), dzk(:), l_t(:), l_b(:)
    real(c_double) :: cfl_max
    real :: cpu_t
    integer :: stps, indx_bk, indx_m, indx_ins, indx_slc, nsteps_env, q
    real(c_double), allocatable, target :: col(:)
    character(len=2), dimension(11), parameter :: mean_names = &
        (/ 'u ', 'v ', 'w ', 't ', 'uu', 'vv', 'ww', 'uw', 'vw', 'tt', 'tw' /)
    character(len=32) :: arg

    ! slice lists end at their first zero entry
    xy2d_id(:) = 0; xz2d_id(:) = 0; yz2d_id(:) = 0
    open(10, file='parameters.input')
    read(10, nml=domain)
    read(10, nml=modeling)
    read(10, nml=constants)
    read(10, nml=io)
    close(10)

    if (mod(nx, 2) .ne. 0 .or. mod(ny, 2) .ne. 0) then
        print *, 'Please use even numbers for nx and ny! Quit!'
        stop
    endif
    if (isnoise .ne. 0) then
        print *, 'isnoise=1 is not supported: prepare seeded initial fields instead'
        stop
    endif

    ! optional first argument: number of steps (defaults to imax)
    nsteps_env = imax
    if (command_argument_count() >= 1) then
        call get_command_argument(1, arg)
        read(arg, *) nsteps_env
    endif

    prm%nx = nx; prm%ny = ny; prm%nz = nz; prm%ochannel = ochannel; prm%dp_opt = dp_opt
    prm%cds = cds; prm%dts = dts; prm%issk = issk; prm%isdamp = isdamp; prm%nzdamp = nzdamp
    prm%isscalar = isscalar; prm%is_ri_var = is_ri_var; prm%imax = imax
    prm%lx = lx; prm%ly = ly; prm%lz = lz; prm%zstretch = zstretch; prm%cdamp = cdamp
    prm%nu = nu; prm%alpha = alpha; prm%got = got; prm%fc = fc; prm%ug = ug; prm%vg = vg
    prm%dt = dt; prm%u_mrf = u_mrf; prm%t_ref = t_ref; prm%tbot = tbot; prm%ttop = ttop
    prm%ri_str = ri_str; prm%ri_end = ri_end; prm%wall_rows = 0; prm%precision = 0
    prm%tbc_top = 0; prm%dtdz_top = 0.0d0; prm%zorder = 2
    call hip_create(prm)

    allocate(f(nx + 4, ny + 4, nz + 4), g(nx, ny, nz))
    f(:,:,:) = 0.d0
    if (restart .eq. 1) then
        print *, 'Reading Initial Fields.'
        call read_one('results/init_u.dat', -u_mrf, HC_U)
        call read_one('results/init_v.dat', 0.d0, HC_V)
        call read_one('results/init_w.dat', 0.d0, HC_W)
        call read_one('results/init_p.dat', 0.d0, HC_P)
        call read_one('results/init_t.dat', 0.d0, HC_T)
        ! for RK3, the f, g, h, s of previous steps are not needed
        if (dts .ne. 2) then
            call read_one('results/init_ff.dat', 0.d0, HC_FF)
            call read_one('results/init_gg.dat', 0.d0, HC_GG)
            call read_one('results/init_hh.dat', 0.d0, HC_HH)
            call read_one('results/init_ss.dat', 0.d0, HC_SS)
        endif
    endif
    call hip_init()

    ! the grid, as get_mesh_param writes it
    allocate(zc(nz), dzk(nz), l_t(nz), l_b(nz))
    call hip_check(hc_get_mesh(hip_ctx, c_loc(zc), c_loc(dzk), c_loc(l_t), c_loc(l_b)), 'hc_get_mesh')
    write(*,*) 'Writing Geo Info to Files.'
    open(11, file='results/geo.info')
    do q = 1, nz
        write(11, '(4e14.6)') zc(q), dzk(q), l_t(q), l_b(q)
    enddo
    close(11)
    call hip_check(hc_mean_fields(hip_ctx, 1_c_int), 'hc_mean_fields')

    if (istmsr .eq. 1) then  ! the time-series files start empty (src/module_navier_stokes.f90:82-93)
        open(99, file='results/tmsr_u.dat', status='replace'); close(99)
        open(98, file='results/tmsr_v.dat', status='replace'); close(98)
        open(97, file='results/tmsr_w.dat', status='replace'); close(97)
        open(96, file='results/tmsr_t.dat', status='replace'); close(96)
        open(95, file='results/tmsr_p.dat', status='replace'); close(95)
    endif
    allocate(col(nz))

    indx_bk = 1
    indx_m = 1
    indx_ins = 1
    indx_slc = 1
    do stps = 1, nsteps_env
        ! screen_cpu_time (src/module_io.f90:32-80): CFL of the fields entering the step
        call hip_check(hc_screen_cfl(hip_ctx, cfl_max), 'screen_cfl')
        call cpu_time(cpu_t)
        print *, ' '
        write (*, 184) stps * dt, cpu_t, cfl_max
184     format(' Time = ', es12.5, '; CPU Time = ', es11.4, '; CFL max = ', f5.2)
        ! the drive loop body (src/module_navier_stokes.f90:143-165) through the
        ! drop-in module: the step prints the divergence lines of every substep
        ! and stops on a NaN, as screen_div_error does
        if (is_ri_var .eq. 1) call update_ri(stps)
        if (cds .eq. 3) then
            call time_advancement_sp()
        else
            call time_advancement_fd()
        endif
        if (dp_opt .eq. 2) call hip_calculate_ubar()
        ! time series of the column at the first interior point (ghst+1, ghst+1)
        if (istmsr .eq. 1) then
            call write_series('results/tmsr_u.dat', HC_U, u_mrf, stps * dt)
            call write_series('results/tmsr_v.dat', HC_V, 0.d0, stps * dt)
            call write_series('results/tmsr_w.dat', HC_W, 0.d0, stps * dt)
            call write_series('results/tmsr_t.dat', HC_T, 0.d0, stps * dt)
            call write_series('results/tmsr_p.dat', HC_P, 0.d0, stps * dt)
        endif
        if (indx_bk / ibackup .eq. 1 .or. stps .eq. nsteps_env) then
            call write_one('results/bak_u.dat', u_mrf, HC_U)
            call write_one('results/bak_v.dat', 0.d0, HC_V)
            call write_one('results/bak_w.dat', 0.d0, HC_W)
            call write_one('results/bak_p.dat', 0.d0, HC_P)
            call write_one('results/bak_t.dat', 0.d0, HC_T)
            if (dts .ne. 2) then
                call write_one('results/bak_ff.dat', 0.d0, HC_FF)
                call write_one('results/bak_gg.dat', 0.d0, HC_GG)
                call write_one('results/bak_hh.dat', 0.d0, HC_HH)
                call write_one('results/bak_ss.dat', 0.d0, HC_SS)
            endif
            indx_bk = 0
        endif
        indx_bk = indx_bk + 1
        if (indx_m / imeanfl .eq. 1) then
            print *, 'Writing Mean Fields...'
            do q = 1, HC_MEAN_COUNT
                call write_mean('results/mean_' // trim(mean_names(q)) // '.dat', q - 1)
            enddo
            indx_m = 0
        endif
        indx_m = indx_m + 1
        if (indx_ins / iinstfl .eq. 1) then
            print *, 'Writing Instantaneous Fields...'
            write (arg, '(i3.3)') stps / iinstfl
            call write_one('results/inst_u' // trim(arg) // '.dat', u_mrf, HC_U)
            call write_one('results/inst_v' // trim(arg) // '.dat', 0.d0, HC_V)
            call write_one('results/inst_w' // trim(arg) // '.dat', 0.d0, HC_W)
            call write_one('results/inst_p' // trim(arg) // '.dat', 0.d0, HC_P)
            call write_one('results/inst_t' // trim(arg) // '.dat', 0.d0, HC_T)
            indx_ins = 0
        endif
        indx_ins = indx_ins + 1
        if (isxy2d .eq. 1 .or. isxz2d .eq. 1 .or. isyz2d .eq. 1) then
            if (indx_slc / intv_2d .eq. 1) then
                call write_slices(stps / intv_2d)
                indx_slc = 0
            endif
        endif
        indx_slc = indx_slc + 1
    enddo

    call hip_destroy()

contains

    subroutine read_one(fname, shift, field)
        character(len=*), intent(in) :: fname
        real(c_double), intent(in) :: shift
        integer(c_int), intent(in) :: field
        integer :: u_
        open(newunit=u_, file=fname, access='stream', form='unformatted', status='old')
        read(u_) g
        close(u_)
        f(:,:,:) = 0.d0
        f(3:nx + 2, 3:ny + 2, 3:nz + 2) = g + shift
        call hip_upload(field, f)
    end subroutine

    subroutine write_one(fname, shift, field)
        character(len=*), intent(in) :: fname
        real(c_double), intent(in) :: shift
        integer(c_int), intent(in) :: field
        integer :: u_
        call hip_download(field, f)
        g = f(3:nx + 2, 3:ny + 2, 3:nz + 2) + shift
        open(newunit=u_, file=fname, access='stream', form='unformatted', status='replace')
        write(u_) g
        close(u_)
    end subroutine

    ! one line: time, then the column's nz values (formats of output_time_series)
    subroutine write_series(fname, field, shift, time)
        character(len=*), intent(in) :: fname
        integer(c_int), intent(in) :: field
        real(c_double), intent(in) :: shift, time
        integer :: u_, k
        call hip_check(hc_get_column(hip_ctx, field, 0_c_int, 0_c_int, c_loc(col)), 'hc_get_column')
        open(newunit=u_, file=fname, status='old', position='append', action='write')
        do k = 1, nz
            if (k .eq. 1) then
                write(u_, '(2f14.8)', advance='no') time, col(k) + shift
            elseif (k .eq. nz) then
                write(u_, '(1f14.8)') col(k) + shift
            else
                write(u_, '(1f14.8)', advance='no') col(k) + shift
            endif
        enddo
        close(u_)
    end subroutine

    ! output_2d_slices: planes of u + u_mrf, v, w, p, t at the listed indices
    subroutine write_slices(nout)
        integer, intent(in) :: nout
        character(len=6) :: ff1
        character(len=4) :: ii1
        character(len=1), dimension(5), parameter :: fn = (/ 'u', 'v', 'w', 'p', 't' /)
        integer(c_int), dimension(5), parameter :: fid = (/ HC_U, HC_V, HC_W, HC_P, HC_T /)
        integer :: n, m, nxy, nxz, nyz, u_
        real(c_double) :: shift
        write (ff1, '(i6.6)') nout
        nxy = count_ids(xy2d_id); nxz = count_ids(xz2d_id); nyz = count_ids(yz2d_id)
        if (isxy2d .eq. 1) print *, 'Writing X-Y Slices...'
        if (isxz2d .eq. 1) print *, 'Writing X-Z Slices...'
        if (isyz2d .eq. 1) print *, 'Writing Y-Z Slices...'
        do m = 1, 5
            shift = 0.d0
            if (m .eq. 1) shift = u_mrf
            call hip_download(fid(m), f)
            g = f(3:nx + 2, 3:ny + 2, 3:nz + 2) + shift
            if (isxy2d .eq. 1) then
                do n = 1, nxy
                    write (ii1, '(i4.4)') xy2d_id(n)
                    open(newunit=u_, file='results/slices/xy2d_' // fn(m) // '_k' // ii1 // '_t' // ff1, &
                         access='stream', form='unformatted', status='replace')
                    write(u_) g(:, :, xy2d_id(n))
                    close(u_)
                enddo
            endif
            if (isxz2d .eq. 1) then
                do n = 1, nxz
                    write (ii1, '(i4.4)') xz2d_id(n)
                    open(newunit=u_, file='results/slices/xz2d_' // fn(m) // '_j' // ii1 // '_t' // ff1, &
                         access='stream', form='unformatted', status='replace')
                    write(u_) g(:, xz2d_id(n), :)
                    close(u_)
                enddo
            endif
            if (isyz2d .eq. 1) then
                do n = 1, nyz
                    write (ii1, '(i4.4)') yz2d_id(n)
                    open(newunit=u_, file='results/slices/yz2d_' // fn(m) // '_i' // ii1 // '_t' // ff1, &
                         access='stream', form='unformatted', status='replace')
                    write(u_) g(yz2d_id(n), :, :)
                    close(u_)
                enddo
            endif
        enddo
    end subroutine

    ! entries before the first zero (output_2d_slices counts the list this way)
    integer function count_ids(ids)
        integer, dimension(:), intent(in) :: ids
        integer :: n
        count_ids = 0
        do n = 1, size(ids)
            if (ids(n) .eq. 0) exit
            count_ids = n
        enddo
    end function

    subroutine write_mean(fname, stat)
        character(len=*), intent(in) :: fname
        integer, intent(in) :: stat
        integer :: u_
        call hip_check(hc_get_mean_field(hip_ctx, int(stat, c_int), c_loc(g), HC_LAYOUT_INTERIOR), &
                       'hc_get_mean_field')
        open(newunit=u_, file=fname, access='stream', form='unformatted', status='replace')
        write(u_) g
        close(u_)
    end subroutine

end program
