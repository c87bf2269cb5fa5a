"""The C ABI driven from C++ alone: tests/cxx/gpe_driver (built with the library) runs a spin-1
set-up, a device-side start, a knot ramp, leapfrog steps, a checkpoint pair resumed in a second
context and a density picture; its resumed run equals the uninterrupted one bit for bit, and its
observables equal those of the same sequence through the Python binding."""
import json
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

gpe = pytest.importorskip("spacetime_amd.gpe")

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "cxx", "gpe_driver")


def _python_run():
    cfg = gpe.Config(lattice="bcc", ncomp=3, precision=gpe.FP64, scheme=gpe.LEAPFROG3, dims=(12, 10, 8),
                     block_scale=0.8 / (2.0 * np.sqrt(2.0)), p0=(-4.8, -4.0, -3.2), dt=1e-4, trap=(1.0, 1.0, 1.3),
                     c0=1.0e3, c1=50.0, quad_sign=1.0, q=0.1)
    knots = [0.05, -0.02, 0.03, 0.0, -0.04, 0.02, 0.01, -0.03]
    mu = 0.5 * (15.0 * cfg.c0 * 1.3 / (4.0 * np.pi)) ** 0.4
    with gpe.Stepper(cfg) as st:
        st.fill_thomas_fermi(cfg.c0, mu, np.array([0, 1, 0], dtype=complex), noise=0.0, seed=0)
        st.itp(50, 5e-5)
        st.project("polar")
        st.add_noise(0.01, 54363)
        st.scale(1.0 / np.sqrt(st.observables(0).norm))
        psi0 = st.get_state(0)
        st.set_knot_ramp(knots, 200 * cfg.dt, bq=4e-3)
        st.set_state(psi0)
        st.start(gpe.START_EULER)
        st.step(200)
        return st.observables(0)


def test_cxx_driver_resumes_bitwise_and_matches_python(tmp_path):
    assert os.path.exists(DRIVER), "tests/cxx/gpe_driver is built by the build step (spacetime_amd.build)"
    r = subprocess.run([DRIVER, str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    out = json.loads(r.stdout.strip().splitlines()[-1])
    assert out["bitwise_resume"] is True and out["levels"] == 2
    assert (tmp_path / ("drv_%s.dat" % out["tag"])).exists() and (tmp_path / ("drv_even_%s.dat" % out["tag"])).exists()
    from spacetime_amd import pictures
    header, rgb = pictures.read(tmp_path / ("drv_%sms.bmp" % out["tag"]))
    assert (header["width"], header["height"]) == (3 * 2 * 14, 2 * 2 * 12) and rgb.max() == 255
    assert out["norm_resumed"] == out["norm"] and out["energy_resumed"] == out["energy"]
    assert np.isfinite(out["max_hamiltonian"]) and out["max_hamiltonian"] > 0
    o = _python_run()
    assert out["time"] == pytest.approx(o.time, rel=1e-14)
    assert out["norm"] == pytest.approx(o.norm, rel=1e-12)
    assert out["energy"] == pytest.approx(o.energy, rel=1e-12)
    assert out["mz"] == pytest.approx(o.magnetization[2], rel=1e-10, abs=1e-14)
