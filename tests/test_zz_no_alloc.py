"""gpe_get_state_noalloc (no allocation: the level streams through the library's static scratch
array in chunks) against gpe_get_state, through the C ABI: the same bytes on two calls in a
row, the first with a level several times the scratch's size, the second with a smaller one."""
import ctypes

import numpy as np
import pytest

from helpers import random_state

pytestmark = pytest.mark.gpu

gpe = pytest.importorskip("spacetime_amd.gpe")

SCRATCH_VALUES = (32 << 20) // 16  # complex128 values per scratch chunk


def _noalloc(st, level, out):
    lib = gpe.load_library()
    f = lib.gpe_get_state_noalloc
    f.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
    f.restype = ctypes.c_int
    return f(st._h, level, out.ctypes.data if out is not None else None)


def test_noalloc_get_state_equals_get_state():
    big = gpe.Config.centred_box("bcc", (64, 64, 64), 0.3, ncomp=3, precision=gpe.FP32, dt=1e-4, c0=50.0, c1=5.0)
    small = gpe.Config.centred_box("c15", (3, 2, 2), 0.6, ncomp=5, precision=gpe.FP64, dt=1e-4, c0=50.0)
    assert int(np.prod(big.shape)) > 4 * SCRATCH_VALUES > 4 * int(np.prod(small.shape))
    for cfg, seed in ((big, 1), (small, 2)):
        with gpe.Stepper(cfg) as st:
            st.set_state(random_state(cfg.shape, seed))
            st.start(gpe.START_EULER)
            st.step(2)  # two different levels
            for level in (0, 1):
                want = st.get_state(level)
                got = np.full(cfg.shape, np.nan, dtype=np.complex128)
                assert _noalloc(st, level, got) == 0
                assert got.tobytes() == want.tobytes(), (cfg.lattice, level)
            assert _noalloc(st, 2, np.zeros(cfg.shape, np.complex128)) == -22
            assert _noalloc(st, 0, None) == -22
