!*************************************************************************
! ISO_C_BINDING layer over the MI355X HERCULES library (include/hercules_amd.h).
!
! The interfaces bind the C ABI one to one.  The module procedures below
! keep the names and argument lists of the reference subroutines they
! replace, so a caller written against the reference's modules links against
! these instead:
!   time_advancement_fd     src/module_navier_stokes.f90:223-365
!   time_advancement_sp     src/module_navier_stokes.f90:371-517
!   poisson_solver_fft      src/module_poisson_solver.f90:483-658
!   get_momentum_cnvdiff    src/module_cnvdiff.f90:35-374
!   get_temperature_cnvdiff src/module_cnvdiff.f90:381-473
!   get_p_eqn_src           src/module_poisson_solver.f90:127-176
!   calculate_uvw           src/module_navier_stokes.f90:659-734
!   calculate_new_p         src/module_navier_stokes.f90:739-808
!   screen_div_error        src/module_io.f90:85-206
!   calculate_ubar          src/module_parameters.f90:513-546
!   update_ri               src/module_parameters.f90:580-590
!   assign_rk_coeff         src/module_parameters.f90:618-641
! time_advancement_fd / _sp run on the device-resident fields, which
! hip_upload / hip_download move across as the reference's ghosted
! (nx+4, ny+4, nz+4) arrays; like the reference they screen the divergence
! after every substep, print its lines and stop on a NaN.  The per-stage
! routines take their array arguments as data (copied in and out on each
! call), for callers that keep the reference's own substep loop.  A nonzero
! status prints the library's message and stops, as the reference does on
! bad flags.
!*************************************************************************
module module_hercules_hip

    use, intrinsic :: iso_c_binding
    implicit none

    type, bind(C) :: hc_params
        integer(c_int) :: nx, ny, nz, ochannel, dp_opt, cds, dts, issk, isdamp, nzdamp, &
                          isscalar, is_ri_var, imax
        real(c_double) :: lx, ly, lz, zstretch, cdamp, nu, alpha, got, fc, ug, vg, dt, &
                          u_mrf, t_ref, tbot, ttop, ri_str, ri_end
        integer(c_int) :: wall_rows
        integer(c_int) :: precision   ! 0 FP64, 1 mixed FP32 + FP64 low wavenumbers
        integer(c_int) :: tbc_top     ! 0 t = ttop at the top wall, 1 dT/dz = dtdz_top
        real(c_double) :: dtdz_top
        integer(c_int) :: zorder      ! 2 (or 0) as the reference, 4 or 6: 4th- or 6th-order z operators
    end type

    integer(c_int), parameter :: HC_U = 0, HC_V = 1, HC_W = 2, HC_P = 3, HC_T = 4
    integer(c_int), parameter :: HC_FF = 5, HC_GG = 6, HC_HH = 7, HC_SS = 8
    ! statistics of calculate_mean_fields, in the order of output_mean_fields
    integer(c_int), parameter :: HC_MEAN_COUNT = 11
    integer(c_int), parameter :: HC_LAYOUT_INTERIOR = 0, HC_LAYOUT_GHOSTED = 1
    integer(c_int), parameter :: HC_EQ_U = 0, HC_EQ_V = 1, HC_EQ_W = 2, HC_EQ_P = 3, HC_EQ_T = 4

    interface
        function hc_last_error() bind(C, name="hc_last_error") result(msg)
            import :: c_ptr
            type(c_ptr) :: msg
        end function
        function hc_create(prm, ctx) bind(C, name="hc_create") result(rc)
            import :: hc_params, c_ptr, c_int
            type(hc_params), intent(in) :: prm
            type(c_ptr), intent(out) :: ctx
            integer(c_int) :: rc
        end function
        function hc_destroy(ctx) bind(C, name="hc_destroy") result(rc)
            import :: c_ptr, c_int
            type(c_ptr), value :: ctx
            integer(c_int) :: rc
        end function
        function hc_set_field(ctx, field, src, layout) bind(C, name="hc_set_field") result(rc)
            import :: c_ptr, c_int
            type(c_ptr), value :: ctx, src
            integer(c_int), value :: field, layout
            integer(c_int) :: rc
        end function
        function hc_get_field(ctx, field, dst, layout) bind(C, name="hc_get_field") result(rc)
            import :: c_ptr, c_int
            type(c_ptr), value :: ctx, dst
            integer(c_int), value :: field, layout
            integer(c_int) :: rc
        end function
        function hc_init(ctx) bind(C, name="hc_init") result(rc)
            import :: c_ptr, c_int
            type(c_ptr), value :: ctx
            integer(c_int) :: rc
        end function
        function hc_time_advancement_fd(ctx) bind(C, name="hc_time_advancement_fd") result(rc)
            import :: c_ptr, c_int
            type(c_ptr), value :: ctx
            integer(c_int) :: rc
        end function
        function hc_time_advancement_sp(ctx) bind(C, name="hc_time_advancement_sp") result(rc)
            import :: c_ptr, c_int
            type(c_ptr), value :: ctx
            integer(c_int) :: rc
        end function
        function hc_drive_steps(ctx, nsteps) bind(C, name="hc_drive_steps") result(rc)
            import :: c_ptr, c_int
            type(c_ptr), value :: ctx
            integer(c_int), value :: nsteps
            integer(c_int) :: rc
        end function
        function hc_poisson_solver_fft(ctx, equation, rk_c, r_src, r_var, layout) &
                 bind(C, name="hc_poisson_solver_fft") result(rc)
            import :: c_ptr, c_int, c_double
            type(c_ptr), value :: ctx, r_src, r_var
            integer(c_int), value :: equation, layout
            real(c_double), value :: rk_c
            integer(c_int) :: rc
        end function
        function hc_calculate_ubar(ctx, ubar) bind(C, name="hc_calculate_ubar") result(rc)
            import :: c_ptr, c_int, c_double
            type(c_ptr), value :: ctx
            real(c_double), intent(out) :: ubar
            integer(c_int) :: rc
        end function
        function hc_get_column(ctx, field, i, j, dst) bind(C, name="hc_get_column") result(rc)
            import :: c_ptr, c_int
            type(c_ptr), value :: ctx, dst
            integer(c_int), value :: field, i, j
            integer(c_int) :: rc
        end function
        function hc_screen_cfl(ctx, cfl_max) bind(C, name="hc_screen_cfl") result(rc)
            import :: c_ptr, c_int, c_double
            type(c_ptr), value :: ctx
            real(c_double), intent(out) :: cfl_max
            integer(c_int) :: rc
        end function
        function hc_get_mesh(ctx, zc, dz, l_t, l_b) bind(C, name="hc_get_mesh") result(rc)
            import :: c_ptr, c_int
            type(c_ptr), value :: ctx, zc, dz, l_t, l_b
            integer(c_int) :: rc
        end function
        function hc_mean_fields(ctx, on) bind(C, name="hc_mean_fields") result(rc)
            import :: c_ptr, c_int
            type(c_ptr), value :: ctx
            integer(c_int), value :: on
            integer(c_int) :: rc
        end function
        function hc_get_mean_field(ctx, stat, dst, layout) bind(C, name="hc_get_mean_field") result(rc)
            import :: c_ptr, c_int
            type(c_ptr), value :: ctx, dst
            integer(c_int), value :: stat, layout
            integer(c_int) :: rc
        end function
        function hc_screen_div_error(ctx, div_avg, div_max) bind(C, name="hc_screen_div_error") result(rc)
            import :: c_ptr, c_int, c_double
            type(c_ptr), value :: ctx
            real(c_double), intent(out) :: div_avg, div_max
            integer(c_int) :: rc
        end function
        function hc_get_scalar(ctx, id, value) bind(C, name="hc_get_scalar") result(rc)
            import :: c_ptr, c_int, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: id
            real(c_double), intent(out) :: value
            integer(c_int) :: rc
        end function
        function hc_project(ctx) bind(C, name="hc_project") result(rc)
            import :: c_ptr, c_int
            type(c_ptr), value :: ctx
            integer(c_int) :: rc
        end function
        function hc_calculate_ubar_of(ctx, u, layout, ubar) bind(C, name="hc_calculate_ubar_of") result(rc)
            import :: c_ptr, c_int, c_double
            type(c_ptr), value :: ctx, u
            integer(c_int), value :: layout
            real(c_double), intent(out) :: ubar
            integer(c_int) :: rc
        end function
        function hc_update_ri(ctx, stps) bind(C, name="hc_update_ri") result(rc)
            import :: c_ptr, c_int
            type(c_ptr), value :: ctx
            integer(c_int), value :: stps
            integer(c_int) :: rc
        end function
        function hc_screen_substeps(ctx, on) bind(C, name="hc_screen_substeps") result(rc)
            import :: c_ptr, c_int
            type(c_ptr), value :: ctx
            integer(c_int), value :: on
            integer(c_int) :: rc
        end function
        function hc_get_screen(ctx, substep, div_avg, div_max, div_hist_max) bind(C, name="hc_get_screen") result(rc)
            import :: c_ptr, c_int, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: substep
            real(c_double), intent(out) :: div_avg, div_max, div_hist_max
            integer(c_int) :: rc
        end function
        function hc_screen_div_error_ref(ctx, isfinal, div_avg, div_max, div_hist_max) &
                 bind(C, name="hc_screen_div_error_ref") result(rc)
            import :: c_ptr, c_int, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: isfinal
            real(c_double), intent(out) :: div_avg, div_max, div_hist_max
            integer(c_int) :: rc
        end function
        function hc_get_momentum_cnvdiff(ctx, rk_a, rk_b, rk_c, u, v, w, p, t, ff, ff0, gg, gg0, hh, hh0, &
                                         qu, qv, qw, layout) bind(C, name="hc_get_momentum_cnvdiff") result(rc)
            import :: c_ptr, c_int, c_double
            type(c_ptr), value :: ctx, u, v, w, p, t, ff, ff0, gg, gg0, hh, hh0, qu, qv, qw
            real(c_double), value :: rk_a, rk_b, rk_c
            integer(c_int), value :: layout
            integer(c_int) :: rc
        end function
        function hc_get_temperature_cnvdiff(ctx, rk_a, rk_b, rk_c, u, v, w, t, ss, ss0, qt, layout) &
                 bind(C, name="hc_get_temperature_cnvdiff") result(rc)
            import :: c_ptr, c_int, c_double
            type(c_ptr), value :: ctx, u, v, w, t, ss, ss0, qt
            real(c_double), value :: rk_a, rk_b, rk_c
            integer(c_int), value :: layout
            integer(c_int) :: rc
        end function
        function hc_get_p_eqn_src(ctx, u, v, w, qp, layout) bind(C, name="hc_get_p_eqn_src") result(rc)
            import :: c_ptr, c_int
            type(c_ptr), value :: ctx, u, v, w, qp
            integer(c_int), value :: layout
            integer(c_int) :: rc
        end function
        function hc_calculate_uvw(ctx, rk_c, u, v, w, phi, layout) bind(C, name="hc_calculate_uvw") result(rc)
            import :: c_ptr, c_int, c_double
            type(c_ptr), value :: ctx, u, v, w, phi
            real(c_double), value :: rk_c
            integer(c_int), value :: layout
            integer(c_int) :: rc
        end function
        function hc_calculate_new_p(ctx, rk_c, phi, p, layout) bind(C, name="hc_calculate_new_p") result(rc)
            import :: c_ptr, c_int, c_double
            type(c_ptr), value :: ctx, phi, p
            real(c_double), value :: rk_c
            integer(c_int), value :: layout
            integer(c_int) :: rc
        end function
        function hc_get_momentum_cnvdiff_spec(ctx, rk_a, rk_b, rk_c, u, v, w, p, t, c_u, c_v, c_w, c_p, &
                                              c_ff, c_gg, c_hh, c_qu, c_qv, c_qw, layout) &
                 bind(C, name="hc_get_momentum_cnvdiff_spec") result(rc)
            import :: c_ptr, c_int, c_double
            type(c_ptr), value :: ctx, u, v, w, p, t, c_u, c_v, c_w, c_p, c_ff, c_gg, c_hh, c_qu, c_qv, c_qw
            real(c_double), value :: rk_a, rk_b, rk_c
            integer(c_int), value :: layout
            integer(c_int) :: rc
        end function
        function hc_get_temperature_cnvdiff_spec(ctx, rk_a, rk_b, rk_c, u, v, w, t, c_t, c_ss, c_qt, layout) &
                 bind(C, name="hc_get_temperature_cnvdiff_spec") result(rc)
            import :: c_ptr, c_int, c_double
            type(c_ptr), value :: ctx, u, v, w, t, c_t, c_ss, c_qt
            real(c_double), value :: rk_a, rk_b, rk_c
            integer(c_int), value :: layout
            integer(c_int) :: rc
        end function
        function hc_poisson_solver_fft_spec(ctx, equation, rk_c, c_src, c_var) &
                 bind(C, name="hc_poisson_solver_fft_spec") result(rc)
            import :: c_ptr, c_int, c_double
            type(c_ptr), value :: ctx, c_src, c_var
            integer(c_int), value :: equation
            real(c_double), value :: rk_c
            integer(c_int) :: rc
        end function
        function hc_get_div_spec(ctx, c_u, c_v, w, c_div, layout) bind(C, name="hc_get_div_spec") result(rc)
            import :: c_ptr, c_int
            type(c_ptr), value :: ctx, c_u, c_v, w, c_div
            integer(c_int), value :: layout
            integer(c_int) :: rc
        end function
        function hc_calculate_uvw_spec(ctx, rk_c, phi, c_phi, c_u, c_v, c_w, layout) &
                 bind(C, name="hc_calculate_uvw_spec") result(rc)
            import :: c_ptr, c_int, c_double
            type(c_ptr), value :: ctx, phi, c_phi, c_u, c_v, c_w
            real(c_double), value :: rk_c
            integer(c_int), value :: layout
            integer(c_int) :: rc
        end function
        function hc_calculate_new_p_spec(ctx, rk_c, phi, c_phi, c_p, layout) &
                 bind(C, name="hc_calculate_new_p_spec") result(rc)
            import :: c_ptr, c_int, c_double
            type(c_ptr), value :: ctx, phi, c_phi, c_p
            real(c_double), value :: rk_c
            integer(c_int), value :: layout
            integer(c_int) :: rc
        end function
        function hc_real_2_cmplx(ctx, r_in, c_out, layout) bind(C, name="hc_real_2_cmplx") result(rc)
            import :: c_ptr, c_int
            type(c_ptr), value :: ctx, r_in, c_out
            integer(c_int), value :: layout
            integer(c_int) :: rc
        end function
        function hc_cmplx_2_real(ctx, c_in, r_out, layout) bind(C, name="hc_cmplx_2_real") result(rc)
            import :: c_ptr, c_int
            type(c_ptr), value :: ctx, c_in, r_out
            integer(c_int), value :: layout
            integer(c_int) :: rc
        end function
        function hc_spectral_truncation(ctx, c_inout) bind(C, name="hc_spectral_truncation") result(rc)
            import :: c_ptr, c_int
            type(c_ptr), value :: ctx, c_inout
            integer(c_int) :: rc
        end function
        function hc_rayleigh_damping_spec(ctx, c_inout) bind(C, name="hc_rayleigh_damping_spec") result(rc)
            import :: c_ptr, c_int
            type(c_ptr), value :: ctx, c_inout
            integer(c_int) :: rc
        end function
        function c_strlen(s) bind(C, name="strlen") result(n)
            import :: c_ptr, c_size_t
            type(c_ptr), value :: s
            integer(c_size_t) :: n
        end function
    end interface

    ! the device-resident channel and the RK coefficient of the current substep
    type(c_ptr), save :: hip_ctx = c_null_ptr
    real(c_double), save :: hip_rk_a = 0.d0, hip_rk_b = 0.d0, hip_rk_c = 1.d0
    ! the parameters the channel was created with, and the module's bulk
    ! velocity (src/module_parameters.f90:205) as calculate_ubar leaves it
    type(hc_params), save :: hip_prm
    real(c_double), save :: ubar = 0.d0
    ! divergence lines after every substep of time_advancement_fd / _sp
    logical, save :: hip_screen_on = .true.

contains

    subroutine hip_check(rc, where)
        integer(c_int), intent(in) :: rc
        character(len=*), intent(in) :: where
        character(kind=c_char), pointer :: msg(:)
        type(c_ptr) :: p
        integer :: n
        if (rc .eq. 0) return
        p = hc_last_error()
        n = int(c_strlen(p))
        call c_f_pointer(p, msg, [n])
        print *, where, ' failed (status ', rc, '): ', msg(1:n)
        stop
    end subroutine

    ! create the device channel (initiate_parameters .. get_*_eqn_coeff)
    subroutine hip_create(prm)
        type(hc_params), intent(in) :: prm
        call hip_check(hc_create(prm, hip_ctx), 'hc_create')
        hip_prm = prm
        call hip_screen(hip_screen_on)
    end subroutine

    ! screen_div_error after every substep of the device step (on by default)
    subroutine hip_screen(on)
        logical, intent(in) :: on
        integer(c_int) :: flag
        flag = 0
        if (on) flag = 1
        call hip_check(hc_screen_substeps(hip_ctx, flag), 'hc_screen_substeps')
        hip_screen_on = on
    end subroutine

    ! the lines screen_div_error prints (src/module_io.f90:160-199)
    subroutine hip_print_div(div_avg, div_max, div_hist_max, isfinal, ishistout)
        real(c_double), intent(in) :: div_avg, div_max, div_hist_max
        integer, intent(in) :: isfinal, ishistout
        real(c_double) :: v
        if (isfinal .eq. 0) then
            write (*, 104) div_avg, div_max
104         format(' DIV-ERR INIT : DIV_AVG = ', es11.4, ';  DIV_MAX = ', es11.4)
        elseif (ishistout .eq. 0) then
            write (*, 105) div_avg, div_max
105         format(' DIV-ERR FINAL: DIV_AVG = ', es11.4, ';  DIV_MAX = ', es11.4)
        else
            write (*, 115) div_avg, div_max
115         format(' DIV-ERR FINAL: DIV_AVG = ', es11.4, ';  DIV_MAX = ', es11.4)
            write (*, 106) div_hist_max
106         format(' DIV-ERR HIST_MAX = ', es11.4)
            if (hip_prm%is_ri_var .eq. 1) then
                call hip_check(hc_get_scalar(hip_ctx, 3_c_int, v), 'hc_get_scalar')
                write (*, 258) v
258             format(' Ri_b = ', es12.5)
            endif
            if (hip_prm%dp_opt .eq. 2) then
                call hip_check(hc_get_scalar(hip_ctx, 1_c_int, v), 'hc_get_scalar')
                write (*, 259) v + hip_prm%u_mrf
259             format(' Ubar = ', es12.5)
            endif
        endif
    end subroutine

    ! the screens recorded after each substep of the last device step
    subroutine hip_print_substeps()
        integer :: s, nsub
        real(c_double) :: a, m, h
        if (.not. hip_screen_on) return
        nsub = 3
        if (hip_prm%dts .ne. 2) nsub = 1
        do s = 1, nsub
            call hip_check(hc_get_screen(hip_ctx, int(s, c_int), a, m, h), 'screen_div_error')
            if (hip_prm%dts .ne. 2 .or. s .eq. 3) then
                call hip_print_div(a, m, h, 1, 1)
            else
                call hip_print_div(a, m, h, 1, 0)
            endif
        enddo
    end subroutine

    subroutine hip_destroy()
        integer(c_int) :: rc
        if (c_associated(hip_ctx)) rc = hc_destroy(hip_ctx)
        hip_ctx = c_null_ptr
    end subroutine

    ! move one ghosted module field to / from the device
    subroutine hip_upload(field, var)
        integer(c_int), intent(in) :: field
        real(c_double), dimension(:,:,:), intent(in), target, contiguous :: var
        call hip_check(hc_set_field(hip_ctx, field, c_loc(var), HC_LAYOUT_GHOSTED), 'hc_set_field')
    end subroutine

    subroutine hip_download(field, var)
        integer(c_int), intent(in) :: field
        real(c_double), dimension(:,:,:), intent(inout), target, contiguous :: var
        call hip_check(hc_get_field(hip_ctx, field, c_loc(var), HC_LAYOUT_GHOSTED), 'hc_get_field')
    end subroutine

    subroutine hip_init()
        call hip_check(hc_init(hip_ctx), 'hc_init')
    end subroutine

    ! src/module_parameters.f90:618-641 (the library applies the same
    ! coefficients internally; kept so poisson_solver_fft sees rk_c)
    subroutine assign_rk_coeff(a_in, b_in, c_in)
        real(c_double), intent(in) :: a_in, b_in, c_in
        hip_rk_a = a_in
        hip_rk_b = b_in
        hip_rk_c = c_in
    end subroutine

    ! src/module_navier_stokes.f90:223-365, on the device-resident state
    subroutine time_advancement_fd()
        call hip_check(hc_time_advancement_fd(hip_ctx), 'time_advancement_fd')
        call hip_print_substeps()
    end subroutine

    ! src/module_navier_stokes.f90:371-517 (cds=3), on the device-resident state
    subroutine time_advancement_sp()
        call hip_check(hc_time_advancement_sp(hip_ctx), 'time_advancement_sp')
        call hip_print_substeps()
    end subroutine

    ! src/module_parameters.f90:580-590
    subroutine update_ri(stps)
        integer, intent(in) :: stps
        call hip_check(hc_update_ri(hip_ctx, int(stps, c_int)), 'update_ri')
    end subroutine

    ! src/module_io.f90:85-206 on the given velocity (which becomes the
    ! channel's u, v, w); prints the reference's lines and stops on a NaN
    subroutine screen_div_error(u, v, w, isfinal, ishistout)
        real(c_double), dimension(:,:,:), intent(in), target, contiguous :: u, v, w
        integer, intent(in) :: isfinal, ishistout
        real(c_double) :: a, m, h
        call hip_upload(HC_U, u)
        call hip_upload(HC_V, v)
        call hip_upload(HC_W, w)
        call hip_check(hc_screen_div_error_ref(hip_ctx, int(isfinal, c_int), a, m, h), 'screen_div_error')
        call hip_print_div(a, m, h, isfinal, ishistout)
    end subroutine

    ! src/module_cnvdiff.f90:35-374.  The interpolants uhx .. why are formed
    ! by the library from u, v, w (the caller's copies are not read).
    subroutine get_momentum_cnvdiff(u, v, w, p, t, uhx, uhy, uhz, vhx, vhy, vhz, whx, why, &
                                    ff, ff0, gg, gg0, hh, hh0, qu, qv, qw)
        real(c_double), dimension(:,:,:), intent(in), target, contiguous :: u, v, w, p, t, ff0, gg0, hh0
        real(c_double), dimension(:,:,:), intent(in) :: uhx, uhy, uhz, vhx, vhy, vhz, whx, why
        real(c_double), dimension(:,:,:), intent(inout), target, contiguous :: ff, gg, hh, qu, qv, qw
        call hip_check(hc_get_momentum_cnvdiff(hip_ctx, hip_rk_a, hip_rk_b, hip_rk_c, c_loc(u), c_loc(v), &
                                               c_loc(w), c_loc(p), c_loc(t), c_loc(ff), c_loc(ff0), c_loc(gg), &
                                               c_loc(gg0), c_loc(hh), c_loc(hh0), c_loc(qu), c_loc(qv), &
                                               c_loc(qw), HC_LAYOUT_GHOSTED), 'get_momentum_cnvdiff')
        if (size(uhx) + size(uhy) + size(uhz) + size(vhx) + size(vhy) + size(vhz) + size(whx) + size(why) < 0) stop
    end subroutine

    ! src/module_cnvdiff.f90:381-473 (thx, thy, thz formed by the library)
    subroutine get_temperature_cnvdiff(u, v, w, t, thx, thy, thz, ss, ss0, qt)
        real(c_double), dimension(:,:,:), intent(in), target, contiguous :: u, v, w, t, ss0
        real(c_double), dimension(:,:,:), intent(in) :: thx, thy, thz
        real(c_double), dimension(:,:,:), intent(inout), target, contiguous :: ss, qt
        call hip_check(hc_get_temperature_cnvdiff(hip_ctx, hip_rk_a, hip_rk_b, hip_rk_c, c_loc(u), c_loc(v), &
                                                  c_loc(w), c_loc(t), c_loc(ss), c_loc(ss0), c_loc(qt), &
                                                  HC_LAYOUT_GHOSTED), 'get_temperature_cnvdiff')
        if (size(thx) + size(thy) + size(thz) < 0) stop
    end subroutine

    ! src/module_poisson_solver.f90:127-176
    subroutine get_p_eqn_src(u, v, w, qp)
        real(c_double), dimension(:,:,:), intent(in), target, contiguous :: u, v, w
        real(c_double), dimension(:,:,:), intent(inout), target, contiguous :: qp
        call hip_check(hc_get_p_eqn_src(hip_ctx, c_loc(u), c_loc(v), c_loc(w), c_loc(qp), HC_LAYOUT_GHOSTED), &
                       'get_p_eqn_src')
    end subroutine

    ! src/module_navier_stokes.f90:659-734
    subroutine calculate_uvw(u, v, w, phi)
        real(c_double), dimension(:,:,:), intent(inout), target, contiguous :: u, v, w
        real(c_double), dimension(:,:,:), intent(in), target, contiguous :: phi
        call hip_check(hc_calculate_uvw(hip_ctx, hip_rk_c, c_loc(u), c_loc(v), c_loc(w), c_loc(phi), &
                                        HC_LAYOUT_GHOSTED), 'calculate_uvw')
    end subroutine

    ! src/module_navier_stokes.f90:739-808
    subroutine calculate_new_p(phi, p)
        real(c_double), dimension(:,:,:), intent(in), target, contiguous :: phi
        real(c_double), dimension(:,:,:), intent(inout), target, contiguous :: p
        call hip_check(hc_calculate_new_p(hip_ctx, hip_rk_c, c_loc(phi), c_loc(p), HC_LAYOUT_GHOSTED), &
                       'calculate_new_p')
    end subroutine

    ! src/module_poisson_solver.f90:483-658.  aa, bb, cc are accepted for
    ! interface compatibility and ignored: the flags select the equation and
    ! its coefficients are held by the library (assign_abc at rk_c).
    subroutine poisson_solver_fft(aa, bb, cc, r_src, r_var, zstagger, delnull, tbnd, ubnd)
        real(c_double), dimension(:,:,:), intent(inout) :: aa, cc
        complex(c_double_complex), dimension(:,:,:), intent(inout) :: bb
        real(c_double), dimension(:,:,:), intent(in), target, contiguous :: r_src
        real(c_double), dimension(:,:,:), intent(out), target, contiguous :: r_var
        integer, intent(in) :: zstagger, delnull, tbnd, ubnd
        integer(c_int) :: eq
        eq = flags_equation(zstagger, delnull, tbnd, ubnd)
        ! the reference zeroes r_var first (:527)
        r_var(:,:,:) = 0.d0
        call hip_check(hc_poisson_solver_fft(hip_ctx, eq, hip_rk_c, c_loc(r_src), c_loc(r_var), &
                                             HC_LAYOUT_GHOSTED), 'poisson_solver_fft')
        if (size(aa) < 0 .or. size(bb) < 0 .or. size(cc) < 0) stop
    end subroutine

    ! the equation a flag tuple of poisson_solver_fft(_spec) selects
    function flags_equation(zstagger, delnull, tbnd, ubnd) result(eq)
        integer, intent(in) :: zstagger, delnull, tbnd, ubnd
        integer(c_int) :: eq
        if (zstagger .eq. 1 .and. delnull .eq. 0 .and. tbnd .eq. 0 .and. ubnd .eq. 0) then
            eq = HC_EQ_W
        elseif (zstagger .eq. 0 .and. delnull .eq. 1 .and. tbnd .eq. 0 .and. ubnd .eq. 0) then
            eq = HC_EQ_P
        elseif (zstagger .eq. 0 .and. delnull .eq. 0 .and. tbnd .eq. 1 .and. ubnd .eq. 0) then
            eq = HC_EQ_T
        elseif (zstagger .eq. 0 .and. delnull .eq. 0 .and. tbnd .eq. 0 .and. ubnd .eq. 1) then
            eq = HC_EQ_U
        elseif (zstagger .eq. 0 .and. delnull .eq. 0 .and. tbnd .eq. 0 .and. ubnd .eq. 0) then
            eq = HC_EQ_V
        else
            print *, 'Invalid zstagger/delnull/tbnd/ubnd flags!'
            stop
        endif
    end function

    ! ---- the spectral substep's stages (src/module_spectral.f90); the complex
    ! work arrays are the reference's c(nx, ny, nz), the real ones its ghosted
    ! module arrays

    ! get_momentum_cnvdiff_spec (:35-504)
    subroutine get_momentum_cnvdiff_spec(u, v, w, p, t, c_u, c_v, c_w, c_p, c_ff, c_gg, c_hh, c_qu, c_qv, c_qw)
        real(c_double), dimension(:,:,:), intent(in), target, contiguous :: u, v, w, p, t
        complex(c_double_complex), dimension(:,:,:), intent(in), target, contiguous :: c_u, c_v, c_w, c_p
        complex(c_double_complex), dimension(:,:,:), intent(inout), target, contiguous :: c_ff, c_gg, c_hh
        complex(c_double_complex), dimension(:,:,:), intent(out), target, contiguous :: c_qu, c_qv, c_qw
        call hip_check(hc_get_momentum_cnvdiff_spec(hip_ctx, hip_rk_a, hip_rk_b, hip_rk_c, c_loc(u), c_loc(v), &
                                                    c_loc(w), c_loc(p), c_loc(t), c_loc(c_u), c_loc(c_v), &
                                                    c_loc(c_w), c_loc(c_p), c_loc(c_ff), c_loc(c_gg), c_loc(c_hh), &
                                                    c_loc(c_qu), c_loc(c_qv), c_loc(c_qw), HC_LAYOUT_GHOSTED), &
                       'get_momentum_cnvdiff_spec')
    end subroutine

    ! get_temperature_cnvdiff_spec (:509-595)
    subroutine get_temperature_cnvdiff_spec(u, v, w, t, c_t, c_ss, c_qt)
        real(c_double), dimension(:,:,:), intent(in), target, contiguous :: u, v, w, t
        complex(c_double_complex), dimension(:,:,:), intent(in), target, contiguous :: c_t
        complex(c_double_complex), dimension(:,:,:), intent(inout), target, contiguous :: c_ss
        complex(c_double_complex), dimension(:,:,:), intent(out), target, contiguous :: c_qt
        call hip_check(hc_get_temperature_cnvdiff_spec(hip_ctx, hip_rk_a, hip_rk_b, hip_rk_c, c_loc(u), c_loc(v), &
                                                       c_loc(w), c_loc(t), c_loc(c_t), c_loc(c_ss), c_loc(c_qt), &
                                                       HC_LAYOUT_GHOSTED), 'get_temperature_cnvdiff_spec')
    end subroutine

    ! poisson_solver_fft_spec (:602-746); the coefficient arrays are held by the library
    subroutine poisson_solver_fft_spec(aa, bb, cc, c_src_in, c_var_out, zstagger, delnull, tbnd, ubnd)
        real(c_double), dimension(:,:,:), intent(inout) :: aa, cc
        complex(c_double_complex), dimension(:,:,:), intent(inout) :: bb
        complex(c_double_complex), dimension(:,:,:), intent(in), target, contiguous :: c_src_in
        complex(c_double_complex), dimension(:,:,:), intent(out), target, contiguous :: c_var_out
        integer, intent(in) :: zstagger, delnull, tbnd, ubnd
        call hip_check(hc_poisson_solver_fft_spec(hip_ctx, flags_equation(zstagger, delnull, tbnd, ubnd), hip_rk_c, &
                                                  c_loc(c_src_in), c_loc(c_var_out)), 'poisson_solver_fft_spec')
        if (size(aa) < 0 .or. size(bb) < 0 .or. size(cc) < 0) stop
    end subroutine

    ! get_div_spec (:752-795)
    subroutine get_div_spec(c_u, c_v, w, c_div)
        complex(c_double_complex), dimension(:,:,:), intent(in), target, contiguous :: c_u, c_v
        real(c_double), dimension(:,:,:), intent(in), target, contiguous :: w
        complex(c_double_complex), dimension(:,:,:), intent(out), target, contiguous :: c_div
        call hip_check(hc_get_div_spec(hip_ctx, c_loc(c_u), c_loc(c_v), c_loc(w), c_loc(c_div), HC_LAYOUT_GHOSTED), &
                       'get_div_spec')
    end subroutine

    ! calculate_uvw_spec (:801-849)
    subroutine calculate_uvw_spec(phi, c_phi, c_u, c_v, c_w)
        real(c_double), dimension(:,:,:), intent(in), target, contiguous :: phi
        complex(c_double_complex), dimension(:,:,:), intent(in), target, contiguous :: c_phi
        complex(c_double_complex), dimension(:,:,:), intent(inout), target, contiguous :: c_u, c_v, c_w
        call hip_check(hc_calculate_uvw_spec(hip_ctx, hip_rk_c, c_loc(phi), c_loc(c_phi), c_loc(c_u), c_loc(c_v), &
                                             c_loc(c_w), HC_LAYOUT_GHOSTED), 'calculate_uvw_spec')
    end subroutine

    ! calculate_new_p_spec (:854-894)
    subroutine calculate_new_p_spec(phi, c_phi, c_p)
        real(c_double), dimension(:,:,:), intent(in), target, contiguous :: phi
        complex(c_double_complex), dimension(:,:,:), intent(in), target, contiguous :: c_phi
        complex(c_double_complex), dimension(:,:,:), intent(inout), target, contiguous :: c_p
        call hip_check(hc_calculate_new_p_spec(hip_ctx, hip_rk_c, c_loc(phi), c_loc(c_phi), c_loc(c_p), &
                                               HC_LAYOUT_GHOSTED), 'calculate_new_p_spec')
    end subroutine

    ! real_2_cmplx / cmplx_2_real (:1157-1217)
    subroutine real_2_cmplx(r_in, c_out)
        real(c_double), dimension(:,:,:), intent(in), target, contiguous :: r_in
        complex(c_double_complex), dimension(:,:,:), intent(out), target, contiguous :: c_out
        call hip_check(hc_real_2_cmplx(hip_ctx, c_loc(r_in), c_loc(c_out), HC_LAYOUT_GHOSTED), 'real_2_cmplx')
    end subroutine
    subroutine cmplx_2_real(c_in, r_out)
        complex(c_double_complex), dimension(:,:,:), intent(in), target, contiguous :: c_in
        real(c_double), dimension(:,:,:), intent(inout), target, contiguous :: r_out
        call hip_check(hc_cmplx_2_real(hip_ctx, c_loc(c_in), c_loc(r_out), HC_LAYOUT_GHOSTED), 'cmplx_2_real')
    end subroutine

    ! spectral_truncation (:1222-1268) and rayleigh_damping_spec (:1345-1391)
    subroutine spectral_truncation(c_inout3)
        complex(c_double_complex), dimension(:,:,:), intent(inout), target, contiguous :: c_inout3
        call hip_check(hc_spectral_truncation(hip_ctx, c_loc(c_inout3)), 'spectral_truncation')
    end subroutine
    subroutine rayleigh_damping_spec(c_varin)
        complex(c_double_complex), dimension(:,:,:), intent(inout), target, contiguous :: c_varin
        call hip_check(hc_rayleigh_damping_spec(hip_ctx, c_loc(c_varin)), 'rayleigh_damping_spec')
    end subroutine

    ! src/module_parameters.f90:513-546: the module's ubar from the given u
    ! (which the next dp_opt=2 update of dpx_drive reads)
    subroutine calculate_ubar(u)
        real(c_double), dimension(:,:,:), intent(in), target, contiguous :: u
        call hip_check(hc_calculate_ubar_of(hip_ctx, c_loc(u), HC_LAYOUT_GHOSTED, ubar), 'calculate_ubar')
    end subroutine

    ! the module's ubar from the device-resident u (the coarse path)
    subroutine hip_calculate_ubar()
        call hip_check(hc_calculate_ubar(hip_ctx, ubar), 'calculate_ubar')
    end subroutine

end module
