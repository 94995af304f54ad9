"""The allocation-free divergence screen (hc_screen_div_error_noalloc)
through the C ABI against hc_screen_div_error_ref: two calls in a row on two
contexts, the second one smaller, sharing the device's scratch array."""
import ctypes as C

import numpy as np
import pytest

from hercules_amd import Channel, cases
from hercules_amd.channel import load_library
from tests.test_oracle_screen import random_state

pytestmark = pytest.mark.gpu


def noalloc(ch, isfinal, hist):
    L = load_library()
    L.hc_screen_div_error_noalloc.argtypes = [C.c_void_p, C.c_int] + [C.POINTER(C.c_double)] * 3
    a, m, h = C.c_double(), C.c_double(), C.c_double(hist)
    rc = L.hc_screen_div_error_noalloc(ch.h, int(isfinal), C.byref(a), C.byref(m), C.byref(h))
    assert rc == 0, L.hc_last_error().decode()
    return a.value, m.value, h.value


@pytest.mark.parametrize("cds", [2, 1])
def test_noalloc_screen_equals_the_allocating_screen(cds):
    results = []
    for q, (nx, ny, nz) in enumerate(((64, 64, 65), (32, 32, 17))):
        prm = cases.closed_channel(nx=nx, ny=ny, nz=nz, cds=cds)
        st = random_state(prm, 50 + q)
        ch = Channel(prm)
        for k, v in st.items():
            ch.set_field(k, v)
        hist = 0.0
        for fin in (1, 1, 0):
            ref = ch.screen_div_error_ref(fin)
            got = noalloc(ch, fin, hist)
            hist = got[2]
            print("noalloc", cds, (nx, ny, nz), fin, got, ref)
            assert got == ref
            for k in ("u", "v", "w"):  # a different field for the next call
                st[k] = 0.8 * np.roll(st[k], 1, axis=2)
                ch.set_field(k, st[k])
        results.append(got)
        ch.close()
    assert results[0] != results[1]


def test_noalloc_rejects_grids_beyond_the_scratch():
    """16x512x1024: 8 + 512*1024 + 3*32*1024 = 622,600 doubles > 2^19."""
    prm = cases.closed_channel(nx=16, ny=512, nz=1024)
    L = load_library()
    L.hc_screen_div_error_noalloc.argtypes = [C.c_void_p, C.c_int] + [C.POINTER(C.c_double)] * 3
    ch = Channel(prm)
    a, m, h = C.c_double(), C.c_double(), C.c_double(0.0)
    rc = L.hc_screen_div_error_noalloc(ch.h, 1, C.byref(a), C.byref(m), C.byref(h))
    ch.close()
    assert rc == 1 and b"scratch" in L.hc_last_error()
