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
!   calculate_ubar          src/module_parameters.f90:513-546
!   assign_rk_coeff         src/module_parameters.f90:618-641
! Fields live on the GPU between calls; hip_upload / hip_download move the
! reference's ghosted (nx+4, ny+4, nz+4) arrays across.  A nonzero status
! prints the library's message and stops, as the reference does on bad flags.
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
        integer(c_int) :: zorder      ! 2 (or 0) as the reference, 6 sixth-order z operators
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
        function c_strlen(s) bind(C, name="strlen") result(n)
            import :: c_ptr, c_size_t
            type(c_ptr), value :: s
            integer(c_size_t) :: n
        end function
    end interface

    ! the device-resident channel and the RK coefficient of the current substep
    type(c_ptr), save :: hip_ctx = c_null_ptr
    real(c_double), save :: hip_rk_a = 0.d0, hip_rk_b = 0.d0, hip_rk_c = 1.d0

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
    end subroutine

    ! src/module_navier_stokes.f90:371-517 (cds=3), on the device-resident state
    subroutine time_advancement_sp()
        call hip_check(hc_time_advancement_sp(hip_ctx), 'time_advancement_sp')
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
        ! the reference zeroes r_var first (:527)
        r_var(:,:,:) = 0.d0
        call hip_check(hc_poisson_solver_fft(hip_ctx, eq, hip_rk_c, c_loc(r_src), c_loc(r_var), &
                                             HC_LAYOUT_GHOSTED), 'poisson_solver_fft')
        if (size(aa) < 0 .or. size(bb) < 0 .or. size(cc) < 0) stop
    end subroutine

    ! src/module_parameters.f90:513-546
    subroutine calculate_ubar(ubar)
        real(c_double), intent(out) :: ubar
        call hip_check(hc_calculate_ubar(hip_ctx, ubar), 'calculate_ubar')
    end subroutine

end module
