/*
 * hercules_amd.h — C ABI of the MI355X-native HERCULES time-step hot path.
 *
 * One hc_ctx holds one channel (the reference's whole single-rank state:
 * fields, grid metrics and equation coefficients) resident in the HBM of the
 * current HIP device.  Every entry point returns 0 on success and a nonzero
 * hc_status otherwise; hc_last_error() gives the message (thread-local).  A
 * context owns one HIP stream; calls on one context must not overlap.
 *
 * Arrays crossing the ABI are FP64.  Two layouts are accepted, chosen per
 * call by `layout`:
 *   HC_LAYOUT_INTERIOR  (nx, ny, nz) column-major, x fastest, no ghosts;
 *   HC_LAYOUT_GHOSTED   (nx+4, ny+4, nz+4) column-major with two ghost cells
 *                       per side, the reference's own z-pencil arrays
 *                       (src/module_variables.f90:32-74); only the interior
 *                       istr3..iend3 x jstr3..jend3 x kstr3..kend3 is read or
 *                       written.
 * Pointers may be host or device memory.
 *
 * The reference's subroutines each entry point replaces are cited as
 * path:line relative to the HERCULES source tree.
 */
#ifndef HERCULES_AMD_H
#define HERCULES_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

/* Namelist parameters of parameters.input (src/module_parameters.f90:40-157)
 * that the time step reads.  p_row/p_col are absent: one GPU holds the
 * whole domain.  I/O cadence parameters stay with the host driver. */
typedef struct hc_params {
    int nx, ny, nz;          /* &domain */
    int ochannel;            /* 0 closed, 1 open channel */
    int dp_opt;              /* 1 constant dp/dx, 2 constant bulk velocity, 3 Ekman */
    int cds;                 /* 1 FD2, 2 FD4, 3 spectral (2/3-rule dealiasing) */
    int dts;                 /* 1 AB2, 2 RK3 */
    int issk;                /* skew-symmetric form (spectral only) */
    int isdamp, nzdamp;      /* Rayleigh damping */
    int isscalar;            /* temperature on/off */
    int is_ri_var;           /* Richardson number ramp */
    int imax;                /* number of steps (used by the Ri ramp) */
    double lx, ly, lz, zstretch, cdamp;
    double nu, alpha, got, fc, ug, vg, dt, u_mrf, t_ref, tbot, ttop, ri_str, ri_end;
    /* Wall rows of the implicit u, v, w and t systems (no namelist
     * counterpart).  0: as the reference builds them — get_{u,w,t}_eqn_coeff
     * zero the 1-D aa(k)/cc(k) while visiting the first mode, so only the
     * (kx,ky) = (0,0) mode gets the -aa/-cc wall term on its diagonal
     * (src/module_poisson_solver.f90:257-272,361-366,455-469).  1: every mode
     * gets the (0,0) mode's rows, which keeps Crank-Nicolson consistent with
     * the explicit half at the wall; 0 is unstable when
     * rk_c*dt*nu/(2 dz_min^2) > 1. */
    int wall_rows;
    /* Arithmetic of the step (no namelist counterpart; the reference is FP64
     * only).  HC_FP64 (0): everything in FP64.  HC_MIXED (1): fields, spectra,
     * right-hand-side stencils, FFTs and the tridiagonal solves of all but the
     * lowest wavenumbers in FP32; the column solves of the low-wavenumber
     * blocks (kx index l < 8, |ky index| <= 8, which contain the kx = ky = 0
     * mean mode and its Neumann compatibility condition) and the mean
     * pressure gradient / bulk velocity update in FP64.  HC_MIXED covers
     * the FD (cds = 1, 2) and spectral (cds = 3) paths and needs isdamp = 0
     * (HC_ERR_UNSUPPORTED otherwise). */
    int precision;
    /* Top temperature condition (no namelist counterpart; the reference is
     * Dirichlet at both walls, src/module_boundary.f90:179-180).  0: t = ttop
     * at the top wall, as the reference.  1: constant gradient dT/dz =
     * dtdz_top at the top wall (ghost t(kend3+1) = t(kend3) + dtdz_top *
     * dz_t(kend3)); ttop is then unused.  The bottom stays t = tbot. */
    int tbc_top;
    double dtdz_top;
    /* Order of the wall-normal (z) operators (no namelist counterpart; the
     * reference is 2nd order in z, src/module_cnvdiff.f90:89,121).  0 or 2:
     * as the reference.  6 (or 4): every z derivative, z interpolation and
     * implicit z operator of the step (FD, cds = 1, 2, and spectral, cds = 3)
     * from 7-point (5-point) second-derivative and 6-point (4-point)
     * first-derivative and interpolation Lagrange
     * stencils on the stretched grid, one-sided near the walls with the wall
     * values as data; the pressure equation is the product of the discrete
     * z divergence and gradient, so the projected velocity is divergence-free
     * to round-off; the implicit solves are exact (banded LU, one thread per
     * horizontal wavenumber).  zorder = 4 or 6 needs precision = HC_FP64,
     * isdamp = 0 and 16 <= nz <= 1024, and has no
     * adjoint (HC_ERR_UNSUPPORTED otherwise); every mode takes the
     * consistent wall rows (wall_rows is not used). */
    int zorder;
} hc_params;

enum hc_precision { HC_FP64 = 0, HC_MIXED = 1 };

typedef struct hc_ctx hc_ctx;

enum hc_status {
    HC_OK = 0,
    HC_ERR_ARG = 1,        /* invalid argument or flag (the reference stops) */
    HC_ERR_UNSUPPORTED = 2,/* parameter combination not implemented */
    HC_ERR_HIP = 3,        /* HIP runtime failure */
    HC_ERR_NAN = 4         /* NaN found by the divergence check */
};

enum hc_field {
    HC_U = 0, HC_V = 1, HC_W = 2, HC_P = 3, HC_T = 4,
    HC_FF = 5, HC_GG = 6, HC_HH = 7, HC_SS = 8
};

enum hc_layout { HC_LAYOUT_INTERIOR = 0, HC_LAYOUT_GHOSTED = 1 };

/* Equation selector for hc_poisson_solver_fft, i.e. the flag tuples
 * (zstagger, delnull, tbnd, ubnd) used at src/module_navier_stokes.f90:293-337:
 * u (0,0,0,1), v (0,0,0,0), w (1,0,0,0), p (0,1,0,0), t (0,0,1,0). */
enum hc_equation { HC_EQ_U = 0, HC_EQ_V = 1, HC_EQ_W = 2, HC_EQ_P = 3, HC_EQ_T = 4 };

/* Message of the last failing call on this thread. */
const char *hc_last_error(void);

/* Replaces initiate_parameters, initiate_indices, get_mesh_param,
 * initiate_fields (allocation/zeroing), get_fft_plan and
 * get_{u,w,p,t}_eqn_coeff: src/module_parameters.f90:242-318,371-508,551-576;
 * src/module_tools.f90:33-227; src/module_poisson_solver.f90:32-122,181-477.
 * Validates the parameters as initiate_parameters does (:256-270). */
int hc_create(const hc_params *prm, hc_ctx **out);
int hc_destroy(hc_ctx *ctx);

/* Copy a module field in / out: the transfer half of read_init_fields
 * (src/module_tools.f90:260-316) and output_backup (src/module_io.f90:311-360).
 * u is the frame-relative module variable; the +-u_mrf shift those two
 * routines apply at the file boundary stays with the caller. */
int hc_set_field(hc_ctx *ctx, int field, const double *src, int layout);
int hc_get_field(hc_ctx *ctx, int field, double *dst, int layout);

/* Start-up part of drive after the fields are read: bulk velocity for
 * dp_opt=2, Ri ramp start, boundary values (src/module_navier_stokes.f90:60-129);
 * with cds=3 also the spectra of every field, their 2/3-rule truncation and
 * the truncated physical fields (:96-123).  With cds=3 a later hc_set_field
 * makes the next substep form the spectra again the same way. */
int hc_init(hc_ctx *ctx);

/* time_advancement_fd: one full step, all RK3 (or AB2) substeps
 * (src/module_navier_stokes.f90:223-365).  Requires cds = 1 or 2. */
int hc_time_advancement_fd(hc_ctx *ctx);

/* time_advancement_sp: one full step of the spectral path
 * (src/module_navier_stokes.f90:371-517, src/module_spectral.f90).
 * Requires cds = 3.  With isdamp = 1, rayleigh_damping_spec relaxes the five
 * spectra toward their level's whole complex mean over the modes, where the
 * reference reduces that mean as one real (src/module_spectral.f90:1372); for
 * the spectra of real fields the mean is real. */
int hc_time_advancement_sp(hc_ctx *ctx);

/* nsteps passes of the drive loop body: update_ri, time_advancement_fd or
 * time_advancement_sp by cds, and calculate_ubar
 * (src/module_navier_stokes.f90:139-165), output excluded. */
int hc_drive_steps(hc_ctx *ctx, int nsteps);

/* poisson_solver_fft for one equation (src/module_poisson_solver.f90:483-658)
 * with its coefficients assigned at rk_c by assign_abc / assign_abc_p
 * (src/module_navier_stokes.f90:523-619).  r_src and r_var are full fields
 * in `layout`; r_var's interior is overwritten (level nz of w is set to 0). */
int hc_poisson_solver_fft(hc_ctx *ctx, int equation, double rk_c, const double *r_src,
                          double *r_var, int layout);

/* Make the current (u, v, w) discretely divergence-free with the scheme's
 * own operators: phi = L^{-1} div u, u -= G phi, w = 0 on the walls.  No
 * counterpart in the reference, which reads prepared initial fields; the
 * operators are those of get_p_eqn_src, poisson_solver_fft (p equation,
 * rk_c = 1) and calculate_uvw.  Used to build seeded initial states. */
int hc_project(hc_ctx *ctx);

/* calculate_ubar (src/module_parameters.f90:513-546) on the context's u. */
int hc_calculate_ubar(hc_ctx *ctx, double *ubar);
/* calculate_ubar(u) with the reference's argument: the bulk velocity of the
 * given u (FP64, `layout`), stored as the module's ubar (read by the dp_opt=2
 * update of the next substep) and returned in *ubar (may be null).  The
 * context's own u is not changed. */
int hc_calculate_ubar_of(hc_ctx *ctx, const double *u, int layout, double *ubar);

/* update_ri (src/module_parameters.f90:580-590): got = ri_str + (ri_end -
 * ri_str) / imax * stps.  hc_drive_steps applies it by itself when
 * is_ri_var = 1; a caller looping over hc_time_advancement_fd / _sp calls it
 * before each step, as the reference's drive loop does (src/module_navier_stokes.f90:143-145). */
int hc_update_ri(hc_ctx *ctx, int stps);

/* Mean |div| and the LARGEST |div| of the current fields (FP64 metrics), and
 * HC_ERR_NAN when u, v or w hold a NaN.  The reference keeps a signed
 * running value instead; hc_screen_div_error_ref returns that. */
int hc_screen_div_error(hc_ctx *ctx, double *div_avg, double *div_max);

/* screen_div_error(u, v, w, isfinal, .) exactly as the reference computes it
 * (src/module_io.f90:85-206): div_avg = sum |d| / (nx ny nz) and the signed
 * values of its scan in (k, j, i) order, div_max (from 0: at each point with
 * |d| > div_max, div_max = d) and div_hist_max (the same rule, kept in the
 * context across calls and updated only when isfinal = 1).  HC_ERR_NAN when
 * u, v or w hold a NaN (the reference prints 'NaN error!' and stops). */
int hc_screen_div_error_ref(hc_ctx *ctx, int isfinal, double *div_avg, double *div_max, double *div_hist_max);

/* No counterpart in the reference: hc_screen_div_error_ref for callers that
 * cannot allocate device memory (loops that forbid hipMalloc, such as tight
 * serving loops and harnesses that pre-plan every allocation).  It allocates
 * nothing: its partial results live in one file-scope device array of 2^19
 * doubles (4 MiB) that each GPU the library runs on holds until the process
 * ends, so the grid must satisfy 8 + ny*nz + 3*ceil(ny/16)*nz <= 2^19
 * (HC_ERR_ARG otherwise; 256x256x257 needs 78,356, 512x512x1024 would need
 * 622,600).  div_hist_max is in/out:
 * the caller keeps the history (isfinal = 1 updates it).  Every call on a GPU
 * shares that GPU's array: calls that overlap on one GPU (two host threads,
 * or two contexts' streams) race and must be serialised by the caller.  The
 * context must live on the current device. */
int hc_screen_div_error_noalloc(hc_ctx *ctx, int isfinal, double *div_avg, double *div_max, double *div_hist_max);

/* Screen every substep (on = 1; default 0): time_advancement_fd / _sp call
 * screen_div_error(u,v,w,1,.) after each substep (src/module_navier_stokes.f90:344-350,
 * 505-509), which updates div_hist_max and stops on a NaN.  With the screen on,
 * hc_time_advancement_fd / _sp and each step of hc_drive_steps wait for the
 * step and return HC_ERR_NAN when a substep left a NaN in u, v or w; the
 * values of the last step's substeps are read with hc_get_screen (substep
 * 1..3; 0 reads the last hc_screen_div_error_ref).  Ensemble steps, the
 * autotuner and the gradient driver do not screen.  Cost at 256x256x257:
 * one read of u, v, w per substep. */
int hc_screen_substeps(hc_ctx *ctx, int on);
int hc_get_screen(hc_ctx *ctx, int substep, double *div_avg, double *div_max, double *div_hist_max);

/* Plane-averaged wall shear stress nu * du/dz at the walls, from the no-slip
 * ghost rule u(kstr3-1) = -u(kstr3) - 2 u_mrf (src/module_boundary.f90:52-53,
 * 66) and dz_b(kstr3) = dz(kstr3), dz_t(kend3) = dz(kend3)
 * (src/module_parameters.f90:469-492): tau_bottom = 2 nu (<u(kstr3)> + u_mrf) /
 * dz(kstr3), tau_top likewise at kend3 for a closed channel and 0 for the
 * free-slip top.  The reference has no such routine; its inputs are the
 * module fields above. */
int hc_wall_shear(hc_ctx *ctx, double *tau_bottom, double *tau_top);

/* Scalars of the reference's module state: 0 dpx_drive, 1 ubar, 2 time_sim,
 * 3 got, 4 number of steps taken; and, with no reference counterpart, 5 the
 * number of fields the column kernels move per global phase (LDS sub-tiles,
 * 1 or 4), 6-9 and 11 the launch shapes of hc_autotune (RHS tile-order
 * group, y-transform columns per block, inverse-x points per block, column
 * block order, momentum tile shape; 0 means the default) and 10 the number
 * of steps in the device statistics. */
int hc_get_scalar(hc_ctx *ctx, int id, double *value);

/* Autotuner (no counterpart in the reference).  Times whole steps of this
 * context from its current state for each candidate launch shape of the
 * finite-difference kernels (RHS tile-order group, y-transform columns per
 * block, inverse-x points per block, column block order, momentum tile
 * shape: 0 whole x rows with the x transform fused, 1-4 tiles of 32x8, 64x8,
 * 32x16, 64x4 points and a separate in-place x transform; coordinate
 * descent, steps_per_trial steps per trial) and keeps the fastest.  Every
 * candidate gives bit-identical results.  The state and step counters are
 * restored; the spectral path (cds = 3) keeps its fixed shapes. */
int hc_autotune(hc_ctx *ctx, int steps_per_trial);
/* Set one of those launch shapes directly (knob 0..4 in the order above;
 * HC_ERR_ARG for a shape the kernels do not accept). */
int hc_set_tuning(hc_ctx *ctx, int knob, int value);

/* The HIP stream (hipStream_t) all work of this context is queued on, and
 * the device address of a field's interior (nx, ny, nz) array: double, or
 * float with precision = HC_MIXED. */
void *hc_stream(hc_ctx *ctx);
void *hc_field_device_ptr(hc_ctx *ctx, int field);

/* Wait for all queued work of the context. */
int hc_synchronize(hc_ctx *ctx);

/* Kernel timing (no counterpart in the reference, which only prints CPU
 * time per step, src/module_io.f90:72-75).  While enabled, every launch of
 * the kernels below is bracketed by HIP events on the context stream;
 * hc_profile_read synchronises and returns the accumulated milliseconds and
 * launch count of one kernel since hc_profile_enable. */
enum hc_kernel {
    HC_K_RHS = 0,        /* momentum right-hand side + x R2C FFT      */
    HC_K_YFFT_FWD = 1,   /* forward y FFT of the three q spectra       */
    HC_K_COLUMN = 2,     /* per-column solves, projection, dp          */
    HC_K_YFFT_INV = 3,   /* inverse y FFT of u, v, w, dp               */
    HC_K_FINISH = 4,     /* x C2R FFT, store u, v, w, accumulate p     */
    HC_K_RHS_T = 5,      /* temperature chain ...                      */
    HC_K_YFFT_T = 6,
    HC_K_COLUMN_T = 7,
    HC_K_FINISH_T = 8,
    HC_K_SOLVE_XR2C = 9, /* standalone poisson_solver_fft chain ...    */
    HC_K_SOLVE_YFFT = 10,
    HC_K_SOLVE_COLUMN = 11,
    HC_K_SOLVE_XC2R = 12,
    HC_KERNEL_COUNT = 13
};
/* One z column (interior indices i, j, 0-based) of a field into dst (nz
 * values): the profile output_time_series writes (src/module_io.f90:211-305). */
int hc_get_column(hc_ctx *ctx, int field, int i, int j, double *dst);

/* The CFL number screen_cpu_time prints (src/module_io.f90:32-80): the
 * largest dt (|u|/dx + |v|/dy + |w|/dz(k)) over the domain, u relative to
 * the moving frame. */
int hc_screen_cfl(hc_ctx *ctx, double *cfl_max);

/* The wall-normal grid as get_mesh_param writes it to results/geo.info
 * (src/module_parameters.f90:495-504): cell-centre z, dz, l_t, l_b of the
 * nz levels (arrays of nz values). */
int hc_get_mesh(hc_ctx *ctx, double *zc, double *dz, double *l_t, double *l_b);

/* Mean statistics, calculate_mean_fields (src/module_tools.f90:362-404) and
 * calculate_mean_fields_spec (src/module_spectral.f90:1012-1052): pointwise
 * running sums over the steps, in the order of output_mean_fields
 * (src/module_io.f90:546-598).  u includes u_mrf; w_c = (w(k) + w(k-1))/2.
 * In the FD path uw and vw take u, v from their x / y interpolants at the
 * start of the step's last substep, as the reference does. */
enum hc_mean {
    HC_MEAN_U = 0, HC_MEAN_V, HC_MEAN_W, HC_MEAN_T, HC_MEAN_UU, HC_MEAN_VV, HC_MEAN_WW, HC_MEAN_UW,
    HC_MEAN_VW, HC_MEAN_TT, HC_MEAN_TW, HC_MEAN_COUNT
};
/* on = 1: zero the sums and add every following step to them (at the end of
 * hc_time_advancement_fd / _sp, hc_drive_steps and ensemble steps); the
 * first call allocates 11 fields (13 in the FD path).  on = 0: stop adding,
 * keep the sums.  hc_get_scalar id 10 returns the number of steps summed. */
int hc_mean_fields(hc_ctx *ctx, int on);
/* output_mean_fields: the sum of statistic `stat` divided by the number of
 * steps summed (the reference divides by the step count stps of its run,
 * the same number when the statistics start with the run). */
int hc_get_mean_field(hc_ctx *ctx, int stat, double *dst, int layout);

int hc_profile_enable(hc_ctx *ctx, int on);
int hc_profile_read(hc_ctx *ctx, int kernel, double *total_ms, long *launches);

/* One substep t_its (1..3 for RK3, 1 for AB2) of time_advancement_fd: one
 * pass of its substep loop (src/module_navier_stokes.f90:247-355).  With
 * cds = 1 or 2. */
int hc_substep_fd(hc_ctx *ctx, int t_its);

/* Discrete adjoint of the finite-difference step (no counterpart in the
 * reference).  Covers FP64, cds = 1 or 2, isscalar = 0, isdamp = 0,
 * dp_opt = 1 or 3 (HC_ERR_UNSUPPORTED otherwise).  The context holds
 * adjoint fields for HC_U, HC_V, HC_W, HC_P and, for HC_FF, HC_GG, HC_HH,
 * the adjoint of the stored right-hand sides f, g, h (the next substep's
 * old right-hand sides); hc_set_adjoint / hc_get_adjoint move them like
 * hc_set_field / hc_get_field.
 *
 * hc_adjoint_substep_fd: with the context's fields equal to the input state
 * of substep t_its (u, v, w, p at its start), replaces the adjoints of the
 * substep's outputs (u, v, w, p after it and its f, g, h) by the adjoints of
 * its inputs (u, v, w, p and the f, g, h it read as old right-hand sides),
 * i.e. applies the transposed Jacobian of hc_substep_fd.  The fields are not
 * changed.
 *
 * hc_wall_shear_gradient: loss = plane-averaged bottom wall shear (as
 * hc_wall_shear) after nsteps RK3 steps from the context's current state;
 * on return the adjoint fields HC_U, HC_V, HC_W, HC_P hold d loss / d
 * (initial u, v, w, p) and the context's fields hold the initial state
 * again.  The forward run keeps (u, v, w, p) every ckpt_every steps; the
 * backward sweep recomputes each segment's substep inputs from its
 * checkpoint (device memory: (nsteps/ckpt_every + 3 ckpt_every) x 4 fields). */
int hc_set_adjoint(hc_ctx *ctx, int field, const double *src, int layout);
int hc_get_adjoint(hc_ctx *ctx, int field, double *dst, int layout);
int hc_adjoint_substep_fd(hc_ctx *ctx, int t_its);
int hc_wall_shear_gradient(hc_ctx *ctx, int nsteps, int ckpt_every, double *loss);

/* Per-stage entry points with the reference's argument lists, for callers
 * that keep the reference's own time_advancement_fd loop on the host.  Every
 * array is a whole FP64 field in `layout` (host or device memory) and crosses
 * the ABI on each call; inputs become the context's module fields (u, v, w,
 * p, t, the old right-hand sides), as the reference's arguments are its
 * module arrays.  Ghost cells are not read: the wall rules and the periodic
 * wrap of update_boundary_* and comm_bound are applied on the fly.  FP64,
 * cds = 1 or 2, 2nd-order z (HC_ERR_UNSUPPORTED / HC_ERR_ARG otherwise).
 * The RK coefficients are the module's rk_a, rk_b, rk_c of assign_rk_coeff
 * (src/module_parameters.f90:618-641).
 *
 * get_momentum_cnvdiff(u,v,w,p,t,uhx,...,why,ff,ff0,gg,gg0,hh,hh0,qu,qv,qw)
 * (src/module_cnvdiff.f90:35-374), including its update_dpx_drive (:51); the
 * interpolants uhx..why are formed inside.  t may be null when isscalar = 0.
 * hh and qw are written as 0 on the top level, which the reference skips. */
int hc_get_momentum_cnvdiff(hc_ctx *ctx, double rk_a, double rk_b, double rk_c, const double *u, const double *v,
                            const double *w, const double *p, const double *t, double *ff, const double *ff0,
                            double *gg, const double *gg0, double *hh, const double *hh0, double *qu, double *qv,
                            double *qw, int layout);
/* get_temperature_cnvdiff(u,v,w,t,thx,thy,thz,ss,ss0,qt)
 * (src/module_cnvdiff.f90:381-473), the interpolants thx..thz formed inside. */
int hc_get_temperature_cnvdiff(hc_ctx *ctx, double rk_a, double rk_b, double rk_c, const double *u, const double *v,
                               const double *w, const double *t, double *ss, const double *ss0, double *qt,
                               int layout);
/* get_p_eqn_src(u,v,w,qp) (src/module_poisson_solver.f90:127-176). */
int hc_get_p_eqn_src(hc_ctx *ctx, const double *u, const double *v, const double *w, double *qp, int layout);
/* calculate_uvw(u,v,w,phi) (src/module_navier_stokes.f90:659-734): u, v, w
 * updated in place. */
int hc_calculate_uvw(hc_ctx *ctx, double rk_c, double *u, double *v, double *w, const double *phi, int layout);
/* calculate_new_p(phi,p) (src/module_navier_stokes.f90:739-808): p updated
 * in place. */
int hc_calculate_new_p(hc_ctx *ctx, double rk_c, const double *phi, double *p, int layout);

/* Per-stage entry points of the spectral substep (cds = 3, FP64, 2nd-order
 * z), for callers that keep the reference's time_advancement_sp loop
 * (src/module_navier_stokes.f90:417-464).  The complex arguments c_* are the
 * reference's work arrays c(nx, ny, nz): full C2C spectra, interleaved
 * (re, im) doubles, x index fastest, forward transform divided by nx ny.
 * They hold the spectra of real fields; the x and y Nyquist modes are not
 * carried (read as 0, written as 0) and the modes above nx/2 are written as
 * the conjugate mirror.  Real arrays are whole fields in `layout`.  Terms the
 * reference forms from its real arguments (the z diffusion and dp/dz, dw/dz,
 * the z derivatives of phi) are formed from those arguments here too.
 *
 * get_momentum_cnvdiff_spec(u,v,w,p,t,c_u,c_v,c_w,c_p,c_ff,c_gg,c_hh,c_qu,c_qv,c_qw)
 * (src/module_spectral.f90:35-504), with its update_dpx (:51): c_ff, c_gg,
 * c_hh hold the old right-hand sides on entry and the new ones on return.
 * t may be null when isscalar = 0. */
int hc_get_momentum_cnvdiff_spec(hc_ctx *ctx, double rk_a, double rk_b, double rk_c, const double *u,
                                 const double *v, const double *w, const double *p, const double *t,
                                 const double *c_u, const double *c_v, const double *c_w, const double *c_p,
                                 double *c_ff, double *c_gg, double *c_hh, double *c_qu, double *c_qv, double *c_qw,
                                 int layout);
/* get_temperature_cnvdiff_spec(u,v,w,t,c_t,c_ss,c_qt) (src/module_spectral.f90:509-595);
 * c_ss old on entry, new on return. */
int hc_get_temperature_cnvdiff_spec(hc_ctx *ctx, double rk_a, double rk_b, double rk_c, const double *u,
                                    const double *v, const double *w, const double *t, const double *c_t,
                                    double *c_ss, double *c_qt, int layout);
/* poisson_solver_fft_spec(aa,bb,cc,c_src,c_var,zstagger,delnull,tbnd,ubnd)
 * (src/module_spectral.f90:602-746), the flag tuple given as the equation
 * of hc_poisson_solver_fft and the coefficients assigned at rk_c. */
int hc_poisson_solver_fft_spec(hc_ctx *ctx, int equation, double rk_c, const double *c_src, double *c_var);
/* get_div_spec(c_u,c_v,w,c_div) (src/module_spectral.f90:752-795). */
int hc_get_div_spec(hc_ctx *ctx, const double *c_u, const double *c_v, const double *w, double *c_div, int layout);
/* calculate_uvw_spec(phi,c_phi,c_u,c_v,c_w) (src/module_spectral.f90:801-849):
 * c_u, c_v, c_w updated in place. */
int hc_calculate_uvw_spec(hc_ctx *ctx, double rk_c, const double *phi, const double *c_phi, double *c_u,
                          double *c_v, double *c_w, int layout);
/* calculate_new_p_spec(phi,c_phi,c_p) (src/module_spectral.f90:854-894):
 * c_p updated in place. */
int hc_calculate_new_p_spec(hc_ctx *ctx, double rk_c, const double *phi, const double *c_phi, double *c_p,
                            int layout);
/* The utilities time_advancement_sp calls between its stages:
 * real_2_cmplx(r_in,c_out) and cmplx_2_real(c_in,r_out)
 * (src/module_spectral.f90:1157-1217; cmplx_2_real writes the interior of
 * r_out), spectral_truncation(c) (:1222-1268) and rayleigh_damping_spec(c)
 * (:1345-1391, with the whole complex mean; needs isdamp = 1). */
int hc_real_2_cmplx(hc_ctx *ctx, const double *r_in, double *c_out, int layout);
int hc_cmplx_2_real(hc_ctx *ctx, const double *c_in, double *r_out, int layout);
int hc_spectral_truncation(hc_ctx *ctx, double *c_inout);
int hc_rayleigh_damping_spec(hc_ctx *ctx, double *c_inout);

/* Ensembles of independent channels (no counterpart in the reference, which
 * runs one channel per job).  The members are contexts made by hc_create and
 * prepared by hc_init; they may differ in grid, path (cds) and physics and
 * stay owned by the caller.  hc_ensemble_steps advances every member by
 * nsteps passes of the drive loop body (as hc_drive_steps), each member on
 * its own stream so that members run concurrently.  With use_graph = 1 the
 * members' launches for two steps are captured once into one HIP graph
 * (fork/join over the member streams) and replayed; members with
 * is_ri_var = 1 or kernel timing enabled make the ensemble fall back to
 * direct launches.  Each member's result is identical to stepping it alone.
 * Members must not be used on their own while the ensemble exists. */
typedef struct hc_ensemble hc_ensemble;
int hc_ensemble_create(hc_ctx *const *members, int n, int use_graph, hc_ensemble **out);
int hc_ensemble_steps(hc_ensemble *ens, int nsteps);
/* 1 when the last hc_ensemble_steps replayed the captured graph */
int hc_ensemble_graph_used(const hc_ensemble *ens);
int hc_ensemble_destroy(hc_ensemble *ens);

#ifdef __cplusplus
}
#endif
#endif
