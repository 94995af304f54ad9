"""ORACLE — test infrastructure only.

ctypes wrapper around ``oracle/libhc_oracle.so``, the plain-C restatement of
HERCULES's finite-difference time step (``oracle/hc_oracle.c``).  Only the
tests, ``__graft_entry__.smoke()`` and the cpu_baseline leg of ``bench.py``
may import this module; it is the checker, never the product path.

Arrays crossing this wrapper are interior fields in C order, shape
``(nz, ny, nx)`` (x fastest), which is the reference's Fortran ``(i,j,k)``
interior order (src/module_tools.f90:273-286 reads them the same way).
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field, fields

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libhc_oracle.so")

_INT_FIELDS = ("nx", "ny", "nz", "ochannel", "dp_opt", "cds", "dts", "issk", "isdamp",
               "nzdamp", "isscalar", "is_ri_var", "imax")
_DBL_FIELDS = ("lx", "ly", "lz", "zstretch", "cdamp", "nu", "alpha", "got", "fc", "ug", "vg",
               "dt", "u_mrf", "t_ref", "tbot", "ttop", "ri_str", "ri_end")


class OParams(C.Structure):
    _fields_ = ([(n, C.c_int) for n in _INT_FIELDS] + [(n, C.c_double) for n in _DBL_FIELDS] +
                [("wall_rows", C.c_int), ("precision", C.c_int), ("tbc_top", C.c_int),
                 ("dtdz_top", C.c_double)])


FIELD_IDS = {"u": 0, "v": 1, "w": 2, "p": 3, "t": 4, "phi": 5, "ff": 6, "gg": 7, "hh": 8,
             "ss": 9, "qu": 10, "qv": 11, "qw": 12, "uhx": 13, "uhy": 14, "uhz": 15,
             "vhx": 16, "vhy": 17, "vhz": 18, "whx": 19, "why": 20, "thx": 21, "thy": 22,
             "thz": 23, "ff0": 24, "gg0": 25, "hh0": 26, "ss0": 27}
EQ_IDS = {"u": 0, "v": 1, "w": 2, "p": 3, "t": 4}
SPEC_IDS = {"u": 0, "v": 1, "w": 2, "p": 3, "t": 4, "phi": 5, "ff": 6, "gg": 7, "hh": 8, "ss": 9,
            "qu": 10, "qv": 11, "qw": 12}

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"oracle library not built: {_LIB_PATH} (run make -C oracle)")
        L = C.CDLL(_LIB_PATH)
        vp, dp, ip = C.c_void_p, C.POINTER(C.c_double), C.c_int
        L.ho_create.restype = vp
        L.ho_create.argtypes = [C.POINTER(OParams)]
        L.ho_destroy.argtypes = [vp]
        L.ho_init.argtypes = [vp]
        L.ho_steps.argtypes = [vp, ip]
        L.ho_error.restype = C.c_char_p
        L.ho_error.argtypes = [vp]
        L.ho_set_field.argtypes = [vp, ip, dp]
        L.ho_get_field.argtypes = [vp, ip, dp]
        L.ho_get_ghosted.argtypes = [vp, ip, dp]
        L.ho_get_mesh.argtypes = [vp, dp, dp, dp, dp, dp]
        L.ho_get_coeff.argtypes = [vp, ip, dp, dp, dp]
        L.ho_solve.argtypes = [vp, ip, C.c_double, dp, dp]
        L.ho_div_stats.argtypes = [vp, dp, dp]
        L.ho_screen.argtypes = [vp, ip, dp, dp, dp]
        L.ho_last_screen.argtypes = [vp, dp, dp, dp]
        L.ho_fft1d.argtypes = [dp, ip, ip]
        L.ho_mesh.argtypes = [vp]
        L.ho_coefficients.argtypes = [vp]
        L.ho_calc_ubar.argtypes = [vp]
        L.ho_stage_interp.argtypes = [vp]
        L.ho_stage_momentum.argtypes = [vp, ip]
        L.ho_stage_p_src.argtypes = [vp]
        L.ho_stage_bc_uvw.argtypes = [vp]
        L.ho_stage_temperature.argtypes = [vp, ip]
        L.ho_stage_damping.argtypes = [vp, ip]
        L.ho_stage_new_p.argtypes = [vp, ip]
        L.ho_stage_project.argtypes = [vp, ip]
        L.ho_get_spectrum.argtypes = [vp, ip, dp]
        L.ho_set_spectrum.argtypes = [vp, ip, dp]
        L.ho_stage_momentum_spec.argtypes = [vp, ip]
        L.ho_stage_damping_spec.argtypes = [vp, ip]
        L.ho_stage_temperature_spec.argtypes = [vp, ip]
        L.ho_stage_poisson_spec.argtypes = [vp, ip, C.c_double]
        L.ho_stage_div_spec.argtypes = [vp]
        L.ho_stage_uvw_spec.argtypes = [vp, ip]
        L.ho_stage_new_p_spec.argtypes = [vp, ip]
        L.ho_spec_divergence.argtypes = [vp, dp]
        L.ho_get_scalar.restype = C.c_double
        L.ho_mean_enable.argtypes = [vp]
        L.ho_get_mean.argtypes = [vp, ip, dp]
        L.ho_get_scalar.argtypes = [vp, ip]
        L.ho_set_scalar.argtypes = [vp, ip, C.c_double]
        _lib = L
    return _lib


def _dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def make_params(**kw) -> OParams:
    p = OParams()
    for n in _INT_FIELDS:
        setattr(p, n, int(kw.get(n, 0)))
    for n in _DBL_FIELDS:
        setattr(p, n, float(kw.get(n, 0.0)))
    p.wall_rows = int(kw.get("wall_rows", 0))
    p.tbc_top = int(kw.get("tbc_top", 0))
    p.dtdz_top = float(kw.get("dtdz_top", 0.0))
    return p


class Oracle:
    """One HERCULES state on the CPU (single rank)."""

    def __init__(self, **params):
        self.params = dict(params)
        self._p = make_params(**params)
        self.h = lib().ho_create(C.byref(self._p))
        if not self.h:
            raise ValueError("oracle rejected the parameters")
        self.nx, self.ny, self.nz = self._p.nx, self._p.ny, self._p.nz

    def close(self):
        if self.h:
            lib().ho_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc:
            raise RuntimeError(lib().ho_error(self.h).decode())

    @property
    def shape(self):
        return (self.nz, self.ny, self.nx)

    def set(self, name, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        assert a.shape == self.shape, (a.shape, self.shape)
        self._check(lib().ho_set_field(self.h, FIELD_IDS[name], _dptr(a)))

    def get(self, name):
        a = np.empty(self.shape, dtype=np.float64)
        self._check(lib().ho_get_field(self.h, FIELD_IDS[name], _dptr(a)))
        return a

    def get_ghosted(self, name):
        a = np.empty((self.nz + 4, self.ny + 4, self.nx + 4), dtype=np.float64)
        self._check(lib().ho_get_ghosted(self.h, FIELD_IDS[name], _dptr(a)))
        return a

    def init(self):
        self._check(lib().ho_init(self.h))

    # calculate_mean_fields statistics, in the order of output_mean_fields
    MEANS = ("u", "v", "w", "t", "uu", "vv", "ww", "uw", "vw", "tt", "tw")

    def mean_enable(self):
        self._check(lib().ho_mean_enable(self.h))

    def get_mean(self, name):
        a = np.empty(self.shape, dtype=np.float64)
        self._check(lib().ho_get_mean(self.h, self.MEANS.index(name), _dptr(a)))
        return a

    def steps(self, n):
        self._check(lib().ho_steps(self.h, int(n)))

    def mesh(self):
        n = self.nz + 4
        out = [np.zeros(n) for _ in range(5)]
        lib().ho_get_mesh(self.h, *[_dptr(a) for a in out])
        return dict(zip(("dz", "l_t", "l_b", "dz_t", "dz_b"), out))

    def coeff(self, eq):
        n = self.nz + 4
        aa, cc = np.zeros(n), np.zeros(n)
        bb = np.zeros((n, self.ny, self.nx), dtype=np.complex128)
        self._check(lib().ho_get_coeff(self.h, {"u": 0, "w": 1, "p": 2, "t": 3}[eq], _dptr(aa),
                                       _dptr(cc), bb.ctypes.data_as(C.POINTER(C.c_double))))
        return aa, bb, cc

    def solve(self, eq, rk_c, src):
        s = np.ascontiguousarray(src, dtype=np.float64)
        out = np.zeros(self.shape)
        self._check(lib().ho_solve(self.h, EQ_IDS[eq], float(rk_c), _dptr(s), _dptr(out)))
        return out

    def spectrum(self, name):
        """C2C spectrum of the spectral path (cds=3), shape (nz, ny, nx),
        forward-normalised as mpi_2d_fft leaves it."""
        a = np.empty(self.shape, dtype=np.complex128)
        self._check(lib().ho_get_spectrum(self.h, SPEC_IDS[name], a.ctypes.data_as(C.POINTER(C.c_double))))
        return a

    def set_spectrum(self, name, arr):
        a = np.ascontiguousarray(arr, dtype=np.complex128)
        self._check(lib().ho_set_spectrum(self.h, SPEC_IDS[name], a.ctypes.data_as(C.POINTER(C.c_double))))

    def spec_divergence(self):
        d = np.empty(self.shape)
        self._check(lib().ho_spec_divergence(self.h, _dptr(d)))
        return d

    def div_stats(self):
        a, m = C.c_double(), C.c_double()
        self._check(lib().ho_div_stats(self.h, C.byref(a), C.byref(m)))
        return a.value, m.value

    def screen(self, isfinal=1):
        """screen_div_error: (div_avg, signed div_max, div_hist_max)."""
        a, m, h = C.c_double(), C.c_double(), C.c_double()
        self._check(lib().ho_screen(self.h, int(isfinal), C.byref(a), C.byref(m), C.byref(h)))
        return a.value, m.value, h.value

    def last_screen(self):
        """(div_avg, div_max, div_hist_max) of the last screen the steps ran."""
        a, m, h = C.c_double(), C.c_double(), C.c_double()
        lib().ho_last_screen(self.h, C.byref(a), C.byref(m), C.byref(h))
        return a.value, m.value, h.value

    def scalar(self, name):
        return lib().ho_get_scalar(self.h, {"dpx_drive": 0, "ubar": 1, "div_hist_max": 2,
                                            "time_sim": 3, "got": 4}[name])

    def set_scalar(self, name, v):
        lib().ho_set_scalar(self.h, {"dpx_drive": 0, "ubar": 1, "got": 4}[name], float(v))


def fft1d(x, sign):
    a = np.ascontiguousarray(x, dtype=np.complex128).copy()
    lib().ho_fft1d(a.ctypes.data_as(C.POINTER(C.c_double)), a.size, int(sign))
    return a
