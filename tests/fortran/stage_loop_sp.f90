!*************************************************************************
! A caller that keeps the reference's own spectral substep loop on the host
! (time_advancement_sp, src/module_navier_stokes.f90:393-509, and the
! spectral start-up of drive, :96-123) and calls the spectral per-stage
! subroutines of module_hercules_hip with the reference's argument lists and
! complex work arrays: rayleigh_damping_spec, get_momentum_cnvdiff_spec,
! poisson_solver_fft_spec, cmplx_2_real, get_div_spec, calculate_new_p_spec,
! calculate_uvw_spec, spectral_truncation, get_temperature_cnvdiff_spec,
! real_2_cmplx.  Reads parameters.input and results/init_*.dat like the
! driver and writes results/stage_{u,v,w,p,t}.dat after the steps.  Used by
! the tests to check the per-stage path against the fused device step.
!*************************************************************************
program stage_loop_sp

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
    real(c_double), allocatable, target, dimension(:,:,:) :: u, v, w, p, t, phi, ff, gg, hh, ss
    complex(c_double_complex), allocatable, target, dimension(:,:,:) :: c_u, c_v, c_w, c_p, c_t, c_phi, &
        c_ff, c_gg, c_hh, c_ss, c_qu, c_qv, c_qw
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
    call hip_screen(.false.)

    allocate(u(nx + 4, ny + 4, nz + 4))
    allocate(v, w, p, t, phi, ff, gg, hh, ss, mold=u)
    allocate(c_u(nx, ny, nz))
    allocate(c_v, c_w, c_p, c_t, c_phi, c_ff, c_gg, c_hh, c_ss, c_qu, c_qv, c_qw, mold=c_u)
    allocate(g(nx, ny, nz), aa(1, 1, 1), cc(1, 1, 1), bb(1, 1, 1))
    u = 0.d0; v = 0.d0; w = 0.d0; p = 0.d0; t = 0.d0; phi = 0.d0
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

    ! the spectral start-up of drive (src/module_navier_stokes.f90:96-123)
    call real_2_cmplx(u, c_u); call real_2_cmplx(v, c_v); call real_2_cmplx(w, c_w)
    call real_2_cmplx(p, c_p); call real_2_cmplx(t, c_t); call real_2_cmplx(ff, c_ff)
    call real_2_cmplx(gg, c_gg); call real_2_cmplx(hh, c_hh); call real_2_cmplx(ss, c_ss)
    call spectral_truncation(c_u); call spectral_truncation(c_v); call spectral_truncation(c_w)
    call spectral_truncation(c_p); call spectral_truncation(c_t)
    call cmplx_2_real(c_u, u); call cmplx_2_real(c_v, v); call cmplx_2_real(c_w, w)
    call cmplx_2_real(c_p, p); call cmplx_2_real(c_t, t)

    nsub = 3
    if (dts .ne. 2) nsub = 1
    do stps = 1, imax
        if (is_ri_var .eq. 1) call update_ri(stps)
        ! time_advancement_sp (src/module_navier_stokes.f90:393-509) on the host
        do t_its = 1, nsub
            if (isdamp .eq. 1) then
                call rayleigh_damping_spec(c_u); call rayleigh_damping_spec(c_v)
                call rayleigh_damping_spec(c_w); call rayleigh_damping_spec(c_p)
                call rayleigh_damping_spec(c_t)
            endif
            if (dts .eq. 2) then
                call assign_rk_coeff(rka(t_its), rkb(t_its), rkc(t_its))
            else
                call assign_rk_coeff(1.5d0, -0.5d0, 1.d0)
            endif
            call get_momentum_cnvdiff_spec(u, v, w, p, t, c_u, c_v, c_w, c_p, c_ff, c_gg, c_hh, c_qu, c_qv, c_qw)
            call poisson_solver_fft_spec(aa, bb, cc, c_qu, c_u, 0, 0, 0, 1)
            call poisson_solver_fft_spec(aa, bb, cc, c_qv, c_v, 0, 0, 0, 0)
            call poisson_solver_fft_spec(aa, bb, cc, c_qw, c_w, 1, 0, 0, 0)
            call cmplx_2_real(c_w, w)
            call get_div_spec(c_u, c_v, w, c_qu)
            c_qu(:,:,:) = c_qu(:,:,:) / dt
            call poisson_solver_fft_spec(aa, bb, cc, c_qu, c_phi, 0, 1, 0, 0)
            call cmplx_2_real(c_phi, phi)
            call calculate_new_p_spec(phi, c_phi, c_p)
            call calculate_uvw_spec(phi, c_phi, c_u, c_v, c_w)
            call spectral_truncation(c_u); call spectral_truncation(c_v)
            call spectral_truncation(c_w); call spectral_truncation(c_p)
            call cmplx_2_real(c_u, u); call cmplx_2_real(c_v, v)
            call cmplx_2_real(c_w, w); call cmplx_2_real(c_p, p)
            if (isscalar .eq. 1) then
                call get_temperature_cnvdiff_spec(u, v, w, t, c_t, c_ss, c_qu)
                call poisson_solver_fft_spec(aa, bb, cc, c_qu, c_t, 0, 0, 1, 0)
                call spectral_truncation(c_t)
                call cmplx_2_real(c_t, t)
            endif
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
