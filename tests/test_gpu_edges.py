"""Degenerate and ragged boxes through the C ABI vs the FP64 oracle: single blocks, single
rows or columns, x-lines shorter than a wave and one value past a wave / pair boundary, in
both precisions and both FP32 lane widths."""
import numpy as np
import pytest

from helpers import oracle_params, random_state, rel_l2, with_field

pytestmark = pytest.mark.gpu

gpe = pytest.importorskip("spacetime_amd.gpe")

FIELD = dict(bq=0.3, bb=(0.05, -0.02, 0.2))
DIMS = [(1, 1, 1), (1, 5, 3), (7, 1, 1), (65, 2, 1), (129, 1, 2), (2, 1, 9)]


def _run(oracle, cfg, n, traversal=None):
    from spacetime_amd import lattice
    L = oracle.lattice(lattice.get(cfg.lattice))
    p = with_field(oracle_params(oracle, cfg), **FIELD)
    psi0 = random_state(cfg.shape, 4)
    a = psi0.copy()
    b = oracle.forward_euler(L, p, a, cfg.dt)
    oracle.leapfrog3(L, p, a, b, cfg.dt, n, t0=cfg.dt)
    with gpe.Stepper(cfg) as st:
        st.set_field(**FIELD)
        if traversal:
            st.set_traversal(*traversal)
        st.set_state(psi0)
        st.start(gpe.START_EULER)
        st.step(n)
        got = st.get_state(0)
    return got, b


@pytest.mark.parametrize("dims", DIMS)
@pytest.mark.parametrize("precision,xpl,tol", [(gpe.FP64, 1, 1e-12), (gpe.FP32, 1, 2e-5), (gpe.FP32, 2, 2e-5)])
def test_degenerate_boxes(oracle, dims, precision, xpl, tol):
    cfg = gpe.Config.centred_box("bcc", dims, 0.6, ncomp=3, precision=precision, dt=2e-4, trap=(1.0, 0.9, 1.1),
                                 c0=40.0, c1=5.0, q=0.2)
    got, ref = _run(oracle, cfg, 9, traversal=(4, 1, xpl))
    assert rel_l2(got, ref) < tol
