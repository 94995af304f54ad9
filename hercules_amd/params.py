"""Run parameters: the namelists of HERCULES's ``parameters.input``.

Mirrors the four namelist groups of src/module_parameters.f90:145-157
(``&domain``, ``&modeling``, ``&constants``, ``&io``) and the validation of
initiate_parameters (:256-270).  ``read_namelist`` parses the file the
reference reads, unchanged.
"""
from __future__ import annotations

import ctypes as C
import re
from dataclasses import dataclass, field, asdict

INT_FIELDS = ("nx", "ny", "nz", "ochannel", "dp_opt", "cds", "dts", "issk", "isdamp",
              "nzdamp", "isscalar", "is_ri_var", "imax")
DBL_FIELDS = ("lx", "ly", "lz", "zstretch", "cdamp", "nu", "alpha", "got", "fc", "ug", "vg",
              "dt", "u_mrf", "t_ref", "tbot", "ttop", "ri_str", "ri_end")


class CParams(C.Structure):
    """ctypes image of ``hc_params`` (include/hercules_amd.h)."""
    _fields_ = ([(n, C.c_int) for n in INT_FIELDS] + [(n, C.c_double) for n in DBL_FIELDS] +
                [("wall_rows", C.c_int), ("precision", C.c_int), ("tbc_top", C.c_int),
                 ("dtdz_top", C.c_double), ("zorder", C.c_int)])


@dataclass
class Params:
    # &domain
    p_row: int = 1
    p_col: int = 1
    lx: float = 12.56637061
    ly: float = 4.188790205
    lz: float = 2.0
    nx: int = 256
    ny: int = 192
    nz: int = 128
    zstretch: float = 1.1
    ochannel: int = 0
    # &modeling
    dp_opt: int = 1
    cds: int = 2
    issk: int = 1
    dts: int = 2
    isnoise: int = 0
    noise_mag: float = 0.1
    isdamp: int = 0
    nzdamp: int = 15
    cdamp: float = 0.01
    # &constants
    nu: float = 0.005555556
    alpha: float = 0.007936508
    got: float = 0.0
    fc: float = 0.0
    ug: float = 0.0
    vg: float = 0.0
    dt: float = 0.001
    imax: int = 20000
    u_mrf: float = 0.0
    isscalar: int = 0
    t_ref: float = 0.0
    tbot: float = -0.5
    ttop: float = 0.5
    is_ri_var: int = 0
    ri_str: float = 0.0
    ri_end: float = 0.0
    # &io
    restart: int = 1
    istmsr: int = 1
    ibackup: int = 2000
    iinstfl: int = 2000
    imeanfl: int = 2000
    isxy2d: int = 0
    xy2d_id: list = field(default_factory=list)
    isxz2d: int = 0
    xz2d_id: list = field(default_factory=list)
    isyz2d: int = 0
    yz2d_id: list = field(default_factory=list)
    intv_2d: int = 1000
    # extension, no namelist counterpart: 1 gives every mode the implicit
    # wall rows the reference gives the (0,0) mode only
    wall_rows: int = 0
    # extension, no namelist counterpart: 0 FP64, 1 mixed FP32 + FP64 low
    # wavenumbers (hc_params.precision)
    precision: int = 0
    # extension, no namelist counterpart: top temperature condition, 0 t = ttop
    # (the reference), 1 dT/dz = dtdz_top (hc_params.tbc_top, dtdz_top)
    tbc_top: int = 0
    dtdz_top: float = 0.0
    # order of the z operators: 2 (the reference), 4 or 6 (hc_params.zorder)
    zorder: int = 2

    def validate(self):
        """The checks of initiate_parameters (src/module_parameters.f90:256-270)."""
        if self.nx % self.p_row or self.ny % self.p_row:
            raise ValueError("nx and ny are not divisible by p_row! Quit!")
        if self.ny % self.p_col or self.nz % self.p_col:
            raise ValueError("ny and nz are not divisible by p_col! Quit!")
        if self.nx % 2 or self.ny % 2:
            raise ValueError("Please use even numbers for nx and ny! Quit!")
        return self

    def to_c(self) -> CParams:
        p = CParams()
        for n in INT_FIELDS:
            setattr(p, n, int(getattr(self, n)))
        for n in DBL_FIELDS:
            setattr(p, n, float(getattr(self, n)))
        p.wall_rows = int(self.wall_rows)
        p.precision = int(self.precision)
        p.tbc_top = int(self.tbc_top)
        p.dtdz_top = float(self.dtdz_top)
        p.zorder = int(self.zorder)
        return p

    def as_dict(self):
        return asdict(self)

    def replace(self, **kw) -> "Params":
        d = self.as_dict()
        d.update(kw)
        return Params(**d)


def _parse_value(tok: str):
    t = tok.strip()
    tl = t.lower()
    if tl in (".true.", "t", ".t."):
        return 1
    if tl in (".false.", "f", ".f."):
        return 0
    if re.fullmatch(r"[+-]?\d+", t):
        return int(t)
    return float(tl.replace("d", "e"))


def read_namelist(path: str) -> Params:
    """Read ``parameters.input`` (the four namelists the reference reads at
    src/module_parameters.f90:247-251)."""
    text = open(path).read()
    prm = Params()
    known = set(prm.as_dict())
    for grp in re.finditer(r"&(\w+)(.*?)/", text, re.S):
        body = "\n".join(line.split("!")[0] for line in grp.group(2).splitlines())
        for m in re.finditer(r"(\w+)\s*=\s*([^=]*?)(?=\s*\w+\s*=|\Z)", body, re.S):
            key, val = m.group(1).lower(), m.group(2).strip().rstrip(",")
            if key not in known:
                raise ValueError(f"unknown namelist variable {key!r} in &{grp.group(1)}")
            parts = [v for v in re.split(r"[,\s]+", val) if v]
            vals = [_parse_value(v) for v in parts]
            cur = getattr(prm, key)
            if isinstance(cur, list):
                setattr(prm, key, vals)
            else:
                v = vals[0]
                setattr(prm, key, int(v) if isinstance(cur, int) else float(v))
    return prm.validate()
