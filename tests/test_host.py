"""Host-side checks that need no GPU: the C ABI library loads and exports
every entry point the header declares; the namelist reader; bench.py's
single-GPU contract."""
import ctypes
import os
import re
import subprocess
import sys

import pytest

from hercules_amd import Params, read_namelist
from hercules_amd.channel import ABI_SYMBOLS, LIB_PATH

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "hercules_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(hc_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_the_binding_symbols():
    decl = _declared()
    assert set(ABI_SYMBOLS) <= set(decl), set(ABI_SYMBOLS) - set(decl)


def test_library_exports_every_declared_symbol():
    if not os.path.exists(LIB_PATH):
        pytest.skip("HIP library not built")
    lib = ctypes.CDLL(LIB_PATH)
    missing = [s for s in _declared() if not hasattr(lib, s)]
    assert not missing, missing


NAMELIST = """
&domain
  p_row = 2, p_col = 1
  lx = 12.56637061, ly = 4.188790205, lz = 2.0
  nx = 256, ny = 192, nz = 128
  zstretch = 1.1   ! comment
  ochannel = 0
/
&modeling
  dp_opt = 2, cds = 2, issk = 1, dts = 2, isnoise = 0, noise_mag = 0.1
  isdamp = 0, nzdamp = 15, cdamp = 0.01
/
&constants
  nu = 0.005555556, alpha = 0.007936508, got = 0.0, fc = 0.0, ug = 0.0, vg = 0.0
  dt = 1.0d-3, imax = 20000, u_mrf = 0.0, isscalar = 0, t_ref = 0
  tbot = -0.5, ttop = 0.5, is_ri_var = 0, ri_str = 0.0, ri_end = 0.0
/
&io
  restart = 1, istmsr = 1, ibackup = 2000, iinstfl = 2000, imeanfl = 2000
  isxy2d = 0, xy2d_id = 10,32,64
  isxz2d = 0, xz2d_id = 32,64
  isyz2d = 0, yz2d_id = 32,64, intv_2d = 1000
/
"""


def test_namelist_reader(tmp_path):
    f = tmp_path / "parameters.input"
    f.write_text(NAMELIST)
    p = read_namelist(str(f))
    assert (p.nx, p.ny, p.nz, p.p_row, p.dp_opt) == (256, 192, 128, 2, 2)
    assert p.dt == 1e-3 and p.xy2d_id == [10, 32, 64] and p.tbot == -0.5
    c = p.to_c()
    assert c.nx == 256 and abs(c.nu - 0.005555556) < 1e-15 and c.wall_rows == 0


def test_namelist_validation(tmp_path):
    f = tmp_path / "parameters.input"
    f.write_text(NAMELIST.replace("nx = 256", "nx = 255"))
    with pytest.raises(ValueError, match="Quit!"):
        read_namelist(str(f))


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr


def _run_py(code):
    env = dict(os.environ)
    env["PYTHONPATH"] = ROOT + os.pathsep + env.get("PYTHONPATH", "")
    return subprocess.run([sys.executable, "-W", "always", "-c", code], capture_output=True, text=True, env=env,
                          cwd=ROOT, timeout=300)


@pytest.mark.skipif(not os.path.exists(LIB_PATH), reason="HIP library not built")
def test_torch_binding_warns_when_the_library_was_loaded_first():
    """torch ships a HIP runtime under the library's soname; the binding warns
    when the library's runtime was loaded before torch, and is silent when
    torch came first (no GPU is touched: loading the library creates no
    context)."""
    pytest.importorskip("torch")
    late = _run_py("from hercules_amd.channel import load_library; load_library(); import hercules_amd.autograd")
    assert late.returncode == 0, late.stderr
    assert "torch is imported after the HIP library" in late.stderr
    early = _run_py("import torch; from hercules_amd.channel import load_library; load_library(); "
                    "import hercules_amd.autograd")
    assert early.returncode == 0, early.stderr
    assert "torch is imported after the HIP library" not in early.stderr
