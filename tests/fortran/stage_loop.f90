!*************************************************************************
! A caller that keeps the reference's own substep loop on the host
! (time_advancement_fd, src/module_navier_stokes.f90:247-355) and calls the
! per-stage subroutines of module_hercules_hip with the reference's argument
! lists: get_momentum_cnvdiff, poisson_solver_fft, get_p_eqn_src,
! calculate_uvw, get_temperature_cnvdiff, screen_div_error, calculate_new_p,
! calculate_ubar.  Reads parameters.input and results/init_*.dat like the
! driver and writes results/stage_{u,v,w,p,t}.dat after the steps.  Used by
! the tests to check the per-stage path against the fused device step.
!*************************************************************************
program stage_loop

    use, intrinsic :: iso_c_binding
    use module_hercules_hip
    implicit none

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

    ! RK3 coefficients of Orlandi (src/module_parameters.f90:210-212)
    real(c_double), parameter :: rka(3) = (/ 8.d0 / 15.d0, 5.d0 / 12.d0, 3.d0 / 4.d0 /)
    real(c_double), parameter :: rkb(3) = (/ 0.d0, -17.d0 / 60.d0, -5.d0 / 12.d0 /)
    real(c_double), parameter :: rkc(3) = (/ 8.d0 / 15.d0, 2.d0 / 15.d0, 1.d0 / 3.d0 /)

    type(hc_params) :: prm
    real(c_double), allocatable, target, dimension(:,:,:) :: u, v, w, p, t, phi, ff, gg, hh, ss, &
        ff0, gg0, hh0, ss0, qu, qv, qw, qp, qt, dum
    real(c_double), allocatable :: g(:,:,:), aa(:,:,:), cc(:,:,:)
    complex(c_double_complex), allocatable :: bb(:,:,:)
    integer :: stps, t_its, nsub

    open(10, file='parameters.input')
    read(10, nml=domain)
    read(10, nml=modeling)
    read(10, nml=constants)
    read(10, nml=io)
    close(10)

    prm%nx = nx; prm%ny = ny; prm%nz = nz; prm%ochannel = ochannel; prm%dp_opt = dp_opt
    prm%cds = cds; prm%dts = dts; prm%issk = issk; prm%isdamp = isdamp; prm%nzdamp = nzdamp
    prm%isscalar = isscalar; prm%is_ri_var = is_ri_var; prm%imax = imax
    prm%lx = lx; prm%ly = ly; prm%lz = lz; prm%zstretch = zstretch; prm%cdamp = cdamp
    prm%nu = nu; prm%alpha = alpha; prm%got = got; prm%fc = fc; prm%ug = ug; prm%vg = vg
    prm%dt = dt; prm%u_mrf = u_mrf; prm%t_ref = t_ref; prm%tbot = tbot; prm%ttop = ttop
    prm%ri_str = ri_str; prm%ri_end = ri_end; prm%wall_rows = 0; prm%precision = 0
    prm%tbc_top = 0; prm%dtdz_top = 0.0d0; prm%zorder = 2
    call hip_create(prm)
    ! the loop below screens after each substep itself, as the reference does
    call hip_screen(.false.)

    allocate(u(nx + 4, ny + 4, nz + 4))
    allocate(v, w, p, t, phi, ff, gg, hh, ss, ff0, gg0, hh0, ss0, qu, qv, qw, qp, qt, dum, mold=u)
    allocate(g(nx, ny, nz), aa(1, 1, 1), cc(1, 1, 1), bb(1, 1, 1))
    u = 0.d0; v = 0.d0; w = 0.d0; p = 0.d0; t = 0.d0; phi = 0.d0; dum = 0.d0
    ff = 0.d0; gg = 0.d0; hh = 0.d0; ss = 0.d0
    call read_one('results/init_u.dat', -u_mrf, u)
    call read_one('results/init_v.dat', 0.d0, v)
    call read_one('results/init_w.dat', 0.d0, w)
    call read_one('results/init_p.dat', 0.d0, p)
    call read_one('results/init_t.dat', 0.d0, t)
    call hip_upload(HC_U, u); call hip_upload(HC_V, v); call hip_upload(HC_W, w)
    call hip_upload(HC_P, p); call hip_upload(HC_T, t)
    call hip_init()
    if (dp_opt .eq. 2) call calculate_ubar(u)

    nsub = 3
    if (dts .ne. 2) nsub = 1
    do stps = 1, imax
        if (is_ri_var .eq. 1) call update_ri(stps)
        do t_its = 1, nsub
            if (dts .eq. 2) then
                call assign_rk_coeff(rka(t_its), rkb(t_its), rkc(t_its))
            else
                call assign_rk_coeff(1.5d0, -0.5d0, 1.d0)
            endif
            ff0 = ff; gg0 = gg; hh0 = hh
            call get_momentum_cnvdiff(u, v, w, p, t, dum, dum, dum, dum, dum, dum, dum, dum, &
                                      ff, ff0, gg, gg0, hh, hh0, qu, qv, qw)
            call poisson_solver_fft(aa, bb, cc, qu, u, 0, 0, 0, 1)
            call poisson_solver_fft(aa, bb, cc, qv, v, 0, 0, 0, 0)
            call poisson_solver_fft(aa, bb, cc, qw, w, 1, 0, 0, 0)
            call get_p_eqn_src(u, v, w, qp)
            call poisson_solver_fft(aa, bb, cc, qp, phi, 0, 1, 0, 0)
            call calculate_uvw(u, v, w, phi)
            if (isscalar .eq. 1) then
                ss0 = ss
                call get_temperature_cnvdiff(u, v, w, t, dum, dum, dum, ss, ss0, qt)
                call poisson_solver_fft(aa, bb, cc, qt, t, 0, 0, 1, 0)
            endif
            if (dts .ne. 2 .or. t_its .eq. 3) then
                call screen_div_error(u, v, w, 1, 1)
            else
                call screen_div_error(u, v, w, 1, 0)
            endif
            call calculate_new_p(phi, p)
        enddo
        if (dp_opt .eq. 2) call calculate_ubar(u)
    enddo

    call write_one('results/stage_u.dat', u_mrf, u)
    call write_one('results/stage_v.dat', 0.d0, v)
    call write_one('results/stage_w.dat', 0.d0, w)
    call write_one('results/stage_p.dat', 0.d0, p)
    call write_one('results/stage_t.dat', 0.d0, t)
    call hip_destroy()

contains

    subroutine read_one(fname, shift, a)
        character(len=*), intent(in) :: fname
        real(c_double), intent(in) :: shift
        real(c_double), intent(inout) :: a(:,:,:)
        integer :: u_
        open(newunit=u_, file=fname, access='stream', form='unformatted', status='old')
        read(u_) g
        close(u_)
        a(3:nx + 2, 3:ny + 2, 3:nz + 2) = g + shift
    end subroutine

    subroutine write_one(fname, shift, a)
        character(len=*), intent(in) :: fname
        real(c_double), intent(in) :: shift
        real(c_double), intent(in) :: a(:,:,:)
        integer :: u_
        g = a(3:nx + 2, 3:ny + 2, 3:nz + 2) + shift
        open(newunit=u_, file=fname, access='stream', form='unformatted', status='replace')
        write(u_) g
        close(u_)
    end subroutine

end program
