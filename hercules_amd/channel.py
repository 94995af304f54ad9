"""Host-side mirror of HERCULES's time-stepping driver over the C ABI.

``Channel`` owns one device-resident channel (an ``hc_ctx``) and exposes the
reference's module subroutines by name: ``time_advancement_fd``,
``poisson_solver_fft``, ``calculate_ubar``, ``screen_div_error`` and the
drive loop body (src/module_navier_stokes.f90:31-365).  Fields cross as
interior arrays of shape (nz, ny, nx) — the reference's Fortran (i,j,k)
interior order — or as the reference's ghosted (nx+4, ny+4, nz+4) arrays.

Every call goes to the HIP library; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .params import Params

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libhercules_amd.so")

FIELDS = {"u": 0, "v": 1, "w": 2, "p": 3, "t": 4, "ff": 5, "gg": 6, "hh": 7, "ss": 8}
EQUATIONS = {"u": 0, "v": 1, "w": 2, "p": 3, "t": 4}
LAYOUT_INTERIOR, LAYOUT_GHOSTED = 0, 1
ABI_SYMBOLS = ("hc_last_error", "hc_create", "hc_destroy", "hc_set_field", "hc_get_field",
               "hc_init", "hc_project", "hc_time_advancement_fd", "hc_time_advancement_sp", "hc_drive_steps",
               "hc_poisson_solver_fft",
               "hc_calculate_ubar", "hc_screen_div_error", "hc_wall_shear", "hc_get_scalar", "hc_stream",
               "hc_field_device_ptr", "hc_synchronize", "hc_profile_enable", "hc_profile_read",
               "hc_ensemble_create", "hc_ensemble_steps", "hc_ensemble_graph_used", "hc_ensemble_destroy",
               "hc_substep_fd", "hc_set_adjoint", "hc_get_adjoint", "hc_adjoint_substep_fd",
               "hc_wall_shear_gradient", "hc_autotune", "hc_set_tuning", "hc_mean_fields", "hc_get_mean_field",
               "hc_get_mesh", "hc_screen_cfl", "hc_get_column", "hc_calculate_ubar_of", "hc_update_ri",
               "hc_screen_div_error_ref", "hc_screen_substeps", "hc_get_screen", "hc_get_momentum_cnvdiff",
               "hc_get_temperature_cnvdiff", "hc_get_p_eqn_src", "hc_calculate_uvw", "hc_calculate_new_p",
               "hc_get_momentum_cnvdiff_spec", "hc_get_temperature_cnvdiff_spec", "hc_poisson_solver_fft_spec",
               "hc_get_div_spec", "hc_calculate_uvw_spec", "hc_calculate_new_p_spec", "hc_real_2_cmplx",
               "hc_cmplx_2_real", "hc_spectral_truncation", "hc_rayleigh_damping_spec")
# statistics of calculate_mean_fields, in the order of output_mean_fields
MEANS = ("u", "v", "w", "t", "uu", "vv", "ww", "uw", "vw", "tt", "tw")
KERNELS = {"rhs": 0, "yfft_fwd": 1, "column": 2, "yfft_inv": 3, "finish": 4, "rhs_t": 5,
           "yfft_t": 6, "column_t": 7, "finish_t": 8, "solve_xr2c": 9, "solve_yfft": 10,
           "solve_column": 11, "solve_xc2r": 12}


class HerculesError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[hc status {code}] {msg}")
        self.code = code


_lib = None


def load_library(path: str = LIB_PATH):
    """Load the HIP library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"HIP library missing: {path}; run __graft_entry__.build()")
    L = C.CDLL(path)
    vp, dp, ip = C.c_void_p, C.POINTER(C.c_double), C.c_int
    L.hc_last_error.restype = C.c_char_p
    L.hc_create.argtypes = [vp, C.POINTER(vp)]
    L.hc_destroy.argtypes = [vp]
    L.hc_set_field.argtypes = [vp, ip, vp, ip]
    L.hc_get_field.argtypes = [vp, ip, vp, ip]
    L.hc_init.argtypes = [vp]
    L.hc_project.argtypes = [vp]
    L.hc_time_advancement_fd.argtypes = [vp]
    L.hc_time_advancement_sp.argtypes = [vp]
    L.hc_drive_steps.argtypes = [vp, ip]
    L.hc_poisson_solver_fft.argtypes = [vp, ip, C.c_double, vp, vp, ip]
    L.hc_calculate_ubar.argtypes = [vp, dp]
    L.hc_screen_div_error.argtypes = [vp, dp, dp]
    L.hc_wall_shear.argtypes = [vp, dp, dp]
    L.hc_ensemble_create.argtypes = [C.POINTER(vp), ip, ip, C.POINTER(vp)]
    L.hc_ensemble_steps.argtypes = [vp, ip]
    L.hc_ensemble_graph_used.argtypes = [vp]
    L.hc_ensemble_destroy.argtypes = [vp]
    L.hc_get_scalar.argtypes = [vp, ip, dp]
    L.hc_stream.restype = vp
    L.hc_stream.argtypes = [vp]
    L.hc_field_device_ptr.restype = vp
    L.hc_field_device_ptr.argtypes = [vp, ip]
    L.hc_synchronize.argtypes = [vp]
    L.hc_substep_fd.argtypes = [vp, ip]
    L.hc_autotune.argtypes = [vp, ip]
    L.hc_set_tuning.argtypes = [vp, ip, ip]
    L.hc_mean_fields.argtypes = [vp, ip]
    L.hc_get_mean_field.argtypes = [vp, ip, vp, ip]
    L.hc_get_mesh.argtypes = [vp, vp, vp, vp, vp]
    L.hc_screen_cfl.argtypes = [vp, dp]
    L.hc_get_column.argtypes = [vp, ip, ip, ip, vp]
    L.hc_set_adjoint.argtypes = [vp, ip, vp, ip]
    L.hc_get_adjoint.argtypes = [vp, ip, vp, ip]
    L.hc_adjoint_substep_fd.argtypes = [vp, ip]
    L.hc_wall_shear_gradient.argtypes = [vp, ip, ip, dp]
    L.hc_profile_enable.argtypes = [vp, ip]
    L.hc_profile_read.argtypes = [vp, ip, dp, C.POINTER(C.c_long)]
    L.hc_calculate_ubar_of.argtypes = [vp, vp, ip, dp]
    L.hc_update_ri.argtypes = [vp, ip]
    L.hc_screen_div_error_ref.argtypes = [vp, ip, dp, dp, dp]
    L.hc_screen_substeps.argtypes = [vp, ip]
    L.hc_get_screen.argtypes = [vp, ip, dp, dp, dp]
    d = C.c_double
    L.hc_get_momentum_cnvdiff.argtypes = [vp, d, d, d] + [vp] * 14 + [ip]
    L.hc_get_temperature_cnvdiff.argtypes = [vp, d, d, d] + [vp] * 7 + [ip]
    L.hc_get_p_eqn_src.argtypes = [vp] + [vp] * 4 + [ip]
    L.hc_calculate_uvw.argtypes = [vp, d] + [vp] * 4 + [ip]
    L.hc_calculate_new_p.argtypes = [vp, d, vp, vp, ip]
    L.hc_get_momentum_cnvdiff_spec.argtypes = [vp, d, d, d] + [vp] * 15 + [ip]
    L.hc_get_temperature_cnvdiff_spec.argtypes = [vp, d, d, d] + [vp] * 7 + [ip]
    L.hc_poisson_solver_fft_spec.argtypes = [vp, ip, d, vp, vp]
    L.hc_get_div_spec.argtypes = [vp] + [vp] * 4 + [ip]
    L.hc_calculate_uvw_spec.argtypes = [vp, d] + [vp] * 5 + [ip]
    L.hc_calculate_new_p_spec.argtypes = [vp, d, vp, vp, vp, ip]
    L.hc_real_2_cmplx.argtypes = [vp, vp, vp, ip]
    L.hc_cmplx_2_real.argtypes = [vp, vp, vp, ip]
    L.hc_spectral_truncation.argtypes = [vp, vp]
    L.hc_rayleigh_damping_spec.argtypes = [vp, vp]
    _lib = L
    return L


def _ptr(a: np.ndarray):
    return C.c_void_p(a.ctypes.data)


class Channel:
    """One channel resident on the current HIP device."""

    def __init__(self, prm: Params):
        self.prm = prm
        self.L = load_library()
        self._cp = prm.to_c()
        h = C.c_void_p()
        self._check(self.L.hc_create(C.byref(self._cp), C.byref(h)))
        self.h = h
        self.shape = (prm.nz, prm.ny, prm.nx)
        self.ghost_shape = (prm.nz + 4, prm.ny + 4, prm.nx + 4)

    def _check(self, rc):
        if rc:
            raise HerculesError(rc, self.L.hc_last_error().decode())

    def close(self):
        if getattr(self, "h", None):
            self.L.hc_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- fields
    def set_field(self, name, arr, layout=LAYOUT_INTERIOR):
        shape = self.shape if layout == LAYOUT_INTERIOR else self.ghost_shape
        a = np.ascontiguousarray(arr, dtype=np.float64)
        if a.shape != shape:
            raise ValueError(f"{name}: shape {a.shape} != {shape}")
        self._check(self.L.hc_set_field(self.h, FIELDS[name], _ptr(a), layout))

    def get_field(self, name, layout=LAYOUT_INTERIOR, out=None):
        shape = self.shape if layout == LAYOUT_INTERIOR else self.ghost_shape
        a = np.zeros(shape) if out is None else out
        self._check(self.L.hc_get_field(self.h, FIELDS[name], _ptr(a), layout))
        return a

    def get_column(self, name, i, j):
        """The z column (i, j) of a field (interior indices)."""
        a = np.zeros(self.prm.nz)
        self._check(self.L.hc_get_column(self.h, FIELDS[name], int(i), int(j), _ptr(a)))
        return a

    def screen_cfl(self):
        """The CFL number of screen_cpu_time (largest dt(|u|/dx + |v|/dy + |w|/dz))."""
        v = C.c_double()
        self._check(self.L.hc_screen_cfl(self.h, C.byref(v)))
        return v.value

    def mesh(self):
        """(zc, dz, l_t, l_b) of the nz levels, the columns of results/geo.info."""
        out = [np.zeros(self.prm.nz) for _ in range(4)]
        self._check(self.L.hc_get_mesh(self.h, *(_ptr(a) for a in out)))
        return tuple(out)

    # ---- mean statistics (calculate_mean_fields / output_mean_fields)
    def mean_fields(self, on=True):
        """Zero the running sums and add every following step (on), or stop adding."""
        self._check(self.L.hc_mean_fields(self.h, 1 if on else 0))

    def get_mean(self, name, layout=LAYOUT_INTERIOR):
        shape = self.shape if layout == LAYOUT_INTERIOR else self.ghost_shape
        a = np.zeros(shape)
        self._check(self.L.hc_get_mean_field(self.h, MEANS.index(name), _ptr(a), layout))
        return a

    def set_state(self, **fields):
        for k, v in fields.items():
            self.set_field(k, v)

    # ---- reference subroutines
    def init(self):
        self._check(self.L.hc_init(self.h))

    def project(self):
        """Make (u, v, w) discretely divergence-free (initial-state preparation)."""
        self._check(self.L.hc_project(self.h))

    def time_advancement_fd(self):
        self._check(self.L.hc_time_advancement_fd(self.h))

    def time_advancement_sp(self):
        self._check(self.L.hc_time_advancement_sp(self.h))

    def steps(self, n):
        self._check(self.L.hc_drive_steps(self.h, int(n)))

    def poisson_solver_fft(self, equation, rk_c, src, layout=LAYOUT_INTERIOR):
        shape = self.shape if layout == LAYOUT_INTERIOR else self.ghost_shape
        s = np.ascontiguousarray(src, dtype=np.float64)
        if s.shape != shape:
            raise ValueError(f"source shape {s.shape} != {shape}")
        out = np.zeros(shape)
        self._check(self.L.hc_poisson_solver_fft(self.h, EQUATIONS[equation], float(rk_c),
                                                 _ptr(s), _ptr(out), layout))
        return out

    def poisson_solver_fft_device(self, equation, rk_c, src_ptr: int, dst_ptr: int):
        """Device-resident variant: src/dst are device addresses of (nz,ny,nx) arrays."""
        self._check(self.L.hc_poisson_solver_fft(self.h, EQUATIONS[equation], float(rk_c),
                                                 C.c_void_p(src_ptr), C.c_void_p(dst_ptr),
                                                 LAYOUT_INTERIOR))

    def autotune(self, steps_per_trial=2):
        """Pick the fastest bit-identical launch shapes for this channel
        (hc_autotune); returns the choices."""
        self._check(self.L.hc_autotune(self.h, int(steps_per_trial)))
        return {k: int(self.scalar("tune_" + k)) for k in ("rgroup", "ytile", "c2r_pts", "col_remap", "rhs_xt")}

    def set_tuning(self, rgroup=None, ytile=None, c2r_pts=None, col_remap=None, rhs_xt=None):
        for knob, v in enumerate((rgroup, ytile, c2r_pts, col_remap, rhs_xt)):
            if v is not None:
                self._check(self.L.hc_set_tuning(self.h, knob, int(v)))

    def substep(self, t_its):
        """One substep of time_advancement_fd (t_its = 1..3 for RK3)."""
        self._check(self.L.hc_substep_fd(self.h, int(t_its)))

    # ---- discrete adjoint
    def set_adjoint(self, name, arr, layout=LAYOUT_INTERIOR):
        shape = self.shape if layout == LAYOUT_INTERIOR else self.ghost_shape
        a = np.ascontiguousarray(arr, dtype=np.float64)
        if a.shape != shape:
            raise ValueError(f"{name}: shape {a.shape} != {shape}")
        self._check(self.L.hc_set_adjoint(self.h, FIELDS[name], _ptr(a), layout))

    def set_field_from(self, name, ptr, layout=LAYOUT_INTERIOR):
        """hc_set_field from raw memory (host or device address of a float64
        array of the layout's shape)."""
        self._check(self.L.hc_set_field(self.h, FIELDS[name], C.c_void_p(int(ptr)), layout))

    def get_adjoint_into(self, name, ptr, layout=LAYOUT_INTERIOR):
        """hc_get_adjoint into raw memory (host or device address)."""
        self._check(self.L.hc_get_adjoint(self.h, FIELDS[name], C.c_void_p(int(ptr)), layout))

    def get_adjoint(self, name, layout=LAYOUT_INTERIOR):
        shape = self.shape if layout == LAYOUT_INTERIOR else self.ghost_shape
        a = np.zeros(shape)
        self._check(self.L.hc_get_adjoint(self.h, FIELDS[name], _ptr(a), layout))
        return a

    def adjoint_substep(self, t_its):
        """Transposed Jacobian of substep t_its at the current (input) state."""
        self._check(self.L.hc_adjoint_substep_fd(self.h, int(t_its)))

    def wall_shear_gradient(self, nsteps, ckpt_every=20):
        """Bottom wall shear after nsteps steps and, in the adjoint fields
        u, v, w, p, its gradient with respect to the current state."""
        v = C.c_double()
        self._check(self.L.hc_wall_shear_gradient(self.h, int(nsteps), int(ckpt_every), C.byref(v)))
        return v.value

    def calculate_ubar(self):
        v = C.c_double()
        self._check(self.L.hc_calculate_ubar(self.h, C.byref(v)))
        return v.value

    def calculate_ubar_of(self, u, layout=LAYOUT_INTERIOR):
        """calculate_ubar(u): bulk velocity of the given u, kept as the module's ubar."""
        a = self._arr(u, layout)
        v = C.c_double()
        self._check(self.L.hc_calculate_ubar_of(self.h, _ptr(a), layout, C.byref(v)))
        return v.value

    def update_ri(self, stps):
        """update_ri(stps): got = ri_str + (ri_end - ri_str) / imax * stps."""
        self._check(self.L.hc_update_ri(self.h, int(stps)))

    def screen_div_error(self):
        """(mean |div|, largest |div|) of the current fields."""
        a, m = C.c_double(), C.c_double()
        self._check(self.L.hc_screen_div_error(self.h, C.byref(a), C.byref(m)))
        return a.value, m.value

    def screen_div_error_ref(self, isfinal=1):
        """The reference's (div_avg, signed div_max, div_hist_max)."""
        a, m, h = C.c_double(), C.c_double(), C.c_double()
        self._check(self.L.hc_screen_div_error_ref(self.h, int(isfinal), C.byref(a), C.byref(m), C.byref(h)))
        return a.value, m.value, h.value

    def screen_substeps(self, on=True):
        """screen_div_error after every substep; NaN stops the step (HerculesError code 4)."""
        self._check(self.L.hc_screen_substeps(self.h, 1 if on else 0))

    def get_screen(self, substep):
        """(div_avg, div_max, div_hist_max) recorded after `substep` of the last step."""
        a, m, h = C.c_double(), C.c_double(), C.c_double()
        self._check(self.L.hc_get_screen(self.h, int(substep), C.byref(a), C.byref(m), C.byref(h)))
        return a.value, m.value, h.value

    # ---- per-stage subroutines with the reference's argument lists
    def _arr(self, x, layout):
        shape = self.shape if layout == LAYOUT_INTERIOR else self.ghost_shape
        a = np.ascontiguousarray(x, dtype=np.float64)
        if a.shape != shape:
            raise ValueError(f"shape {a.shape} != {shape}")
        return a

    def _out(self, layout):
        return np.zeros(self.shape if layout == LAYOUT_INTERIOR else self.ghost_shape)

    def get_momentum_cnvdiff(self, rk, u, v, w, p, t, ff0, gg0, hh0, layout=LAYOUT_INTERIOR):
        """get_momentum_cnvdiff with rk = (rk_a, rk_b, rk_c); returns ff, gg, hh, qu, qv, qw."""
        ins = [self._arr(x, layout) for x in (u, v, w, p)]
        tt = None if t is None else self._arr(t, layout)
        olds = [self._arr(x, layout) for x in (ff0, gg0, hh0)]
        outs = [self._out(layout) for _ in range(6)]
        ff, gg, hh, qu, qv, qw = outs
        self._check(self.L.hc_get_momentum_cnvdiff(
            self.h, float(rk[0]), float(rk[1]), float(rk[2]), *(_ptr(a) for a in ins),
            None if tt is None else _ptr(tt), _ptr(ff), _ptr(olds[0]), _ptr(gg), _ptr(olds[1]), _ptr(hh),
            _ptr(olds[2]), _ptr(qu), _ptr(qv), _ptr(qw), layout))
        return ff, gg, hh, qu, qv, qw

    def get_temperature_cnvdiff(self, rk, u, v, w, t, ss0, layout=LAYOUT_INTERIOR):
        """get_temperature_cnvdiff; returns ss, qt."""
        ins = [self._arr(x, layout) for x in (u, v, w, t)]
        s0 = self._arr(ss0, layout)
        ss, qt = self._out(layout), self._out(layout)
        self._check(self.L.hc_get_temperature_cnvdiff(self.h, float(rk[0]), float(rk[1]), float(rk[2]),
                                                      *(_ptr(a) for a in ins), _ptr(ss), _ptr(s0), _ptr(qt), layout))
        return ss, qt

    def get_p_eqn_src(self, u, v, w, layout=LAYOUT_INTERIOR):
        ins = [self._arr(x, layout) for x in (u, v, w)]
        qp = self._out(layout)
        self._check(self.L.hc_get_p_eqn_src(self.h, *(_ptr(a) for a in ins), _ptr(qp), layout))
        return qp

    def calculate_uvw(self, rk_c, u, v, w, phi, layout=LAYOUT_INTERIOR):
        """calculate_uvw: returns the projected (u, v, w) (copies of the inputs updated)."""
        uvw = [self._arr(x, layout).copy() for x in (u, v, w)]
        ph = self._arr(phi, layout)
        self._check(self.L.hc_calculate_uvw(self.h, float(rk_c), *(_ptr(a) for a in uvw), _ptr(ph), layout))
        return tuple(uvw)

    def calculate_new_p(self, rk_c, phi, p, layout=LAYOUT_INTERIOR):
        ph = self._arr(phi, layout)
        pp = self._arr(p, layout).copy()
        self._check(self.L.hc_calculate_new_p(self.h, float(rk_c), _ptr(ph), _ptr(pp), layout))
        return pp

    # ---- spectral per-stage subroutines: complex arrays (nz, ny, nx), i.e.
    # the reference's c(nx, ny, nz) work arrays
    def _spec(self, c):
        a = np.ascontiguousarray(c, dtype=np.complex128)
        if a.shape != self.shape:
            raise ValueError(f"spectrum shape {a.shape} != {self.shape}")
        return a

    def get_momentum_cnvdiff_spec(self, rk, u, v, w, p, t, c_u, c_v, c_w, c_p, c_ff, c_gg, c_hh,
                                  layout=LAYOUT_INTERIOR):
        """get_momentum_cnvdiff_spec; returns c_ff, c_gg, c_hh (new), c_qu, c_qv, c_qw."""
        ins = [self._arr(x, layout) for x in (u, v, w, p)]
        tt = None if t is None else self._arr(t, layout)
        cs = [self._spec(x) for x in (c_u, c_v, c_w, c_p)]
        fs = [self._spec(x).copy() for x in (c_ff, c_gg, c_hh)]
        qs = [np.zeros(self.shape, dtype=np.complex128) for _ in range(3)]
        self._check(self.L.hc_get_momentum_cnvdiff_spec(
            self.h, float(rk[0]), float(rk[1]), float(rk[2]), *(_ptr(a) for a in ins),
            None if tt is None else _ptr(tt), *(_ptr(a) for a in cs), *(_ptr(a) for a in fs),
            *(_ptr(a) for a in qs), layout))
        return (*fs, *qs)

    def get_temperature_cnvdiff_spec(self, rk, u, v, w, t, c_t, c_ss, layout=LAYOUT_INTERIOR):
        """get_temperature_cnvdiff_spec; returns c_ss (new), c_qt."""
        ins = [self._arr(x, layout) for x in (u, v, w, t)]
        ct = self._spec(c_t)
        ss = self._spec(c_ss).copy()
        qt = np.zeros(self.shape, dtype=np.complex128)
        self._check(self.L.hc_get_temperature_cnvdiff_spec(self.h, float(rk[0]), float(rk[1]), float(rk[2]),
                                                           *(_ptr(a) for a in ins), _ptr(ct), _ptr(ss), _ptr(qt),
                                                           layout))
        return ss, qt

    def poisson_solver_fft_spec(self, equation, rk_c, c_src):
        src = self._spec(c_src)
        out = np.zeros(self.shape, dtype=np.complex128)
        self._check(self.L.hc_poisson_solver_fft_spec(self.h, EQUATIONS[equation], float(rk_c), _ptr(src), _ptr(out)))
        return out

    def get_div_spec(self, c_u, c_v, w, layout=LAYOUT_INTERIOR):
        cu, cv = self._spec(c_u), self._spec(c_v)
        ww = self._arr(w, layout)
        out = np.zeros(self.shape, dtype=np.complex128)
        self._check(self.L.hc_get_div_spec(self.h, _ptr(cu), _ptr(cv), _ptr(ww), _ptr(out), layout))
        return out

    def calculate_uvw_spec(self, rk_c, phi, c_phi, c_u, c_v, c_w, layout=LAYOUT_INTERIOR):
        """calculate_uvw_spec: returns updated copies of c_u, c_v, c_w."""
        ph, cph = self._arr(phi, layout), self._spec(c_phi)
        uvw = [self._spec(x).copy() for x in (c_u, c_v, c_w)]
        self._check(self.L.hc_calculate_uvw_spec(self.h, float(rk_c), _ptr(ph), _ptr(cph), *(_ptr(a) for a in uvw),
                                                 layout))
        return tuple(uvw)

    def calculate_new_p_spec(self, rk_c, phi, c_phi, c_p, layout=LAYOUT_INTERIOR):
        ph, cph = self._arr(phi, layout), self._spec(c_phi)
        cp = self._spec(c_p).copy()
        self._check(self.L.hc_calculate_new_p_spec(self.h, float(rk_c), _ptr(ph), _ptr(cph), _ptr(cp), layout))
        return cp

    def real_2_cmplx(self, r, layout=LAYOUT_INTERIOR):
        rr = self._arr(r, layout)
        out = np.zeros(self.shape, dtype=np.complex128)
        self._check(self.L.hc_real_2_cmplx(self.h, _ptr(rr), _ptr(out), layout))
        return out

    def cmplx_2_real(self, c):
        cc = self._spec(c)
        out = np.zeros(self.shape)
        self._check(self.L.hc_cmplx_2_real(self.h, _ptr(cc), _ptr(out), LAYOUT_INTERIOR))
        return out

    def spectral_truncation(self, c):
        cc = self._spec(c).copy()
        self._check(self.L.hc_spectral_truncation(self.h, _ptr(cc)))
        return cc

    def rayleigh_damping_spec(self, c):
        cc = self._spec(c).copy()
        self._check(self.L.hc_rayleigh_damping_spec(self.h, _ptr(cc)))
        return cc

    def wall_shear(self):
        """(tau_bottom, tau_top): plane-averaged nu du/dz at the walls."""
        b, t = C.c_double(), C.c_double()
        self._check(self.L.hc_wall_shear(self.h, C.byref(b), C.byref(t)))
        return b.value, t.value

    def scalar(self, name):
        v = C.c_double()
        ids = {"dpx_drive": 0, "ubar": 1, "time_sim": 2, "got": 3, "stps": 4, "col_tiles": 5,
               "tune_rgroup": 6, "tune_ytile": 7, "tune_c2r_pts": 8, "tune_col_remap": 9,
               "mean_steps": 10, "tune_rhs_xt": 11}
        self._check(self.L.hc_get_scalar(self.h, ids[name], C.byref(v)))
        return v.value

    def stream(self) -> int:
        return self.L.hc_stream(self.h) or 0

    def device_ptr(self, name) -> int:
        return self.L.hc_field_device_ptr(self.h, FIELDS[name]) or 0

    def synchronize(self):
        self._check(self.L.hc_synchronize(self.h))

    # ---- kernel timing
    def profile(self, on=True):
        self._check(self.L.hc_profile_enable(self.h, 1 if on else 0))

    def profile_read(self):
        """{kernel: (total_ms, launches)} accumulated since profile(True)."""
        out = {}
        for name, kid in KERNELS.items():
            ms, n = C.c_double(), C.c_long()
            self._check(self.L.hc_profile_read(self.h, kid, C.byref(ms), C.byref(n)))
            if n.value:
                out[name] = (ms.value, n.value)
        return out


class Ensemble:
    """Independent channels advanced together (hc_ensemble_*): each member on
    its own stream, optionally replayed from one captured HIP graph."""

    def __init__(self, members, use_graph=True):
        self.L = load_library()
        self.members = list(members)
        arr = (C.c_void_p * len(self.members))(*[m.h.value for m in self.members])
        h = C.c_void_p()
        rc = self.L.hc_ensemble_create(arr, len(self.members), int(bool(use_graph)), C.byref(h))
        if rc:
            raise HerculesError(rc, self.L.hc_last_error().decode())
        self.h = h

    def steps(self, n):
        rc = self.L.hc_ensemble_steps(self.h, int(n))
        if rc:
            raise HerculesError(rc, self.L.hc_last_error().decode())

    @property
    def graph_used(self):
        return bool(self.L.hc_ensemble_graph_used(self.h))

    def close(self):
        if getattr(self, "h", None):
            self.L.hc_ensemble_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
