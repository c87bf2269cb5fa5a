"""State files, checkpoints, spinor states and VTK files through the C ABI on the device.

* Checkpoint pair (Spin1GpeBCC.cu:1228-1307,1668-1713): N steps, save, resume in a fresh
  context, M more steps == N + M uninterrupted steps, bit for bit (FP64 and FP32).
* Device spinor projections (Spin1GpeBCC.cu:463-551, Spin2GpeBCC.cu:417-558) and noise against
  host restatements; knot ramps against the equivalent ramp segments, bit for bit.
* VTK, .dat and BMP files written from a context equal the host writers' files for the same state.
* The adjoint of an FP32 context with FP64 psi and mu equals an FP64 context's adjoint.
"""
import os

import numpy as np
import pytest

from helpers import normalized, random_state, rel_l2

pytestmark = pytest.mark.gpu

gpe = pytest.importorskip("spacetime_amd.gpe")


def _cfg(precision, lat="bcc", ncomp=3, dims=(6, 5, 4), dt=2e-4):
    return gpe.Config.centred_box(lat, dims, 0.6, ncomp=ncomp, precision=precision, dt=dt, c0=60.0, c1=8.0,
                                  c2=-6.0, q=0.2)


def _knots():
    return np.random.default_rng(5).uniform(-0.2, 0.2, 8)


def _stepper(cfg, psi):
    st = gpe.Stepper(cfg)
    st.set_knot_ramp(_knots(), 60 * cfg.dt, bq=0.3)
    st.set_state(psi)
    return st


@pytest.mark.parametrize("precision", [gpe.FP64, gpe.FP32])
def test_checkpoint_resume_is_bitwise(tmp_path, precision):
    cfg = _cfg(precision)
    psi = normalized(random_state(cfg.shape, 21), cfg)
    n, m = 17, 14
    with _stepper(cfg, psi) as a:
        a.start(gpe.START_EULER)
        a.step(n)
        tag = a.save_checkpoint(str(tmp_path), prefix="run_")
        t_saved = a.time()
        a.step(m)
        want0, want1, want_t = a.get_state(0), a.get_state(1), a.time()
    assert os.path.exists(tmp_path / ("run_%s.dat" % tag)) and os.path.exists(tmp_path / ("run_even_%s.dat" % tag))
    assert float(tag) == pytest.approx(t_saved[0], abs=1e-15)
    with gpe.Stepper(cfg) as b:
        b.set_knot_ramp(_knots(), 60 * cfg.dt, bq=0.3)
        assert b.load_checkpoint(str(tmp_path), tag, prefix="run_") == 2
        assert b.time() == t_saved  # the clock comes back exactly
        b.step(m)  # no start: both levels were loaded
        np.testing.assert_array_equal(b.get_state(0), want0)
        np.testing.assert_array_equal(b.get_state(1), want1)
        assert b.time() == want_t
    # only the odd file: level 0 from it, the caller starts again (the reference's forward step)
    os.remove(tmp_path / ("run_even_%s.dat" % tag))
    with gpe.Stepper(cfg) as c:
        assert c.load_checkpoint(str(tmp_path), tag, prefix="run_") == 1
        with pytest.raises(gpe.GpeError, match="ESTATE"):
            c.step(1)
        c.start(gpe.START_EULER)
        c.step(1)


def test_dat_files_from_a_context_match_the_host_writer(tmp_path):
    from spacetime_amd import states
    cfg = _cfg(gpe.FP64, lat="a15", ncomp=5, dims=(3, 2, 2))
    psi = random_state(cfg.shape, 3)
    with gpe.Stepper(cfg) as st:
        st.set_state(psi)
        st.write_dat(str(tmp_path / "dev.dat"))
        st.read_dat(str(tmp_path / "dev.dat"), level=1)
        back = st.get_state(1)
    states.write_reference_dat(tmp_path / "host.dat", cfg, psi)
    assert (tmp_path / "dev.dat").read_bytes() == (tmp_path / "host.dat").read_bytes()
    np.testing.assert_array_equal(back, psi)


@pytest.mark.parametrize("ncomp,phases", [(3, [("polar", 0), ("ferro", 0), ("mixed", 0.3)]),
                                          (5, [("un", 0), ("bn_hori", 0.7), ("bn_vert", -1.1), ("cyclic", 2.0),
                                               ("ferro", 0)])])
@pytest.mark.parametrize("precision", [gpe.FP64, gpe.FP32])
def test_device_projection_matches_host(ncomp, phases, precision):
    from spacetime_amd import states
    cfg = _cfg(precision, ncomp=ncomp, dims=(5, 4, 3))
    psi = random_state(cfg.shape, 8)
    n = np.sum(psi.real ** 2 + psi.imag ** 2, axis=0)
    for name, param in phases:
        with gpe.Stepper(cfg) as st:
            st.set_state(psi)
            stored = st.get_state(0)
            st.project(name, param)
            got = st.get_state(0)
        ns = np.sum(stored.real ** 2 + stored.imag ** 2, axis=0)
        kw = dict(mix=param) if name == "mixed" else dict(rel_phase=param)
        want = states.project(stored, name, **kw)
        tol = 1e-14 if precision == gpe.FP64 else 1e-6
        assert rel_l2(got, want) < tol, name
        # the density of every unknown is kept
        np.testing.assert_allclose(np.sum(np.abs(got) ** 2, axis=0), ns, rtol=10 * tol)
        assert rel_l2(ns, n) < (1e-15 if precision == gpe.FP64 else 1e-6)
    with gpe.Stepper(cfg) as st, pytest.raises(gpe.GpeError, match="EINVAL"):
        st.project("un" if ncomp == 3 else "polar")


def _hash_uniform(seed, idx, s):
    M = np.uint64(0xFFFFFFFFFFFFFFFF)
    with np.errstate(over="ignore"):
        z = (np.uint64(seed) * np.uint64(0x9E3779B97F4A7C15) + idx.astype(np.uint64) * np.uint64(0xD1B54A32D192ED03)
             + np.uint64(s) * np.uint64(0x94D049BB133111EB)) & M
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) / 2.0 ** 53


def test_device_noise_matches_host_restatement():
    cfg = _cfg(gpe.FP64, ncomp=5, lat="fcc", dims=(4, 3, 3))
    psi = random_state(cfg.shape, 2)
    amp, seed = 0.02, 99
    with gpe.Stepper(cfg) as st:
        st.set_state(psi)
        st.add_noise(amp, seed)
        got = st.get_state(0)
    nz, ny, nx, V = cfg.shape[1:]
    node = np.arange(nz * ny * nx * V).reshape(nz, ny, nx, V)
    n = np.sum(psi.real ** 2 + psi.imag ** 2, axis=0)
    want = psi.copy()
    for c in range(5):
        rad = np.sqrt(_hash_uniform(seed, node, 2 * c))
        want[c] += amp * np.sqrt(n) * rad * np.exp(2j * np.pi * _hash_uniform(seed, node, 2 * c + 1))
    assert rel_l2(got, want) < 1e-14
    # every component draws amp sqrt(n) sqrt(u): |noise|^2 / |psi|^2 ~ ncomp amp^2 / 2
    assert 0.5 * amp < rel_l2(got, psi) < amp * np.sqrt(5.0)


def test_knot_ramp_equals_its_segments():
    from spacetime_amd import states
    cfg = _cfg(gpe.FP64)
    psi = random_state(cfg.shape, 4)
    out = []
    for use_knots in (True, False):
        with gpe.Stepper(cfg) as st:
            if use_knots:
                st.set_knot_ramp(_knots(), 60 * cfg.dt, bq=0.3)
            else:
                st.set_ramp(q_segments=[(0.3, 1e300, 0)], b_segments=states.knot_ramp(_knots(), 60 * cfg.dt))
            st.set_state(psi)
            st.start(gpe.START_EULER)
            st.step(70)
            out.append(st.get_state(0))
    np.testing.assert_array_equal(out[0], out[1])


@pytest.mark.parametrize("ncomp,kinds", [(3, ["spinor", "preimage", "volume"]), (5, ["spinor", "volume"])])
def test_vtk_from_a_context_matches_the_host_writer(tmp_path, ncomp, kinds):
    from spacetime_amd import vtk
    cfg = _cfg(gpe.FP64, ncomp=ncomp, dims=(4, 3, 2))
    psi = random_state(cfg.shape, 12)
    if ncomp == 3:  # make part of the state y-polarised for the pre-image
        psi[:, :, :, :2] = np.array([0.5, 0.5j * 2 ** 0.5, -0.5])[:, None, None, None, None]
    with gpe.Stepper(cfg) as st:
        st.set_state(psi)
        for kind in kinds:
            getattr(vtk, "save_" + kind)(st, tmp_path / ("dev_%s.vtk" % kind))
    for kind in kinds:
        getattr(vtk, "write_" + kind)(tmp_path / ("host_%s.vtk" % kind), cfg, psi)
        dev = (tmp_path / ("dev_%s.vtk" % kind)).read_bytes()
        assert dev == (tmp_path / ("host_%s.vtk" % kind)).read_bytes(), kind
        assert len(dev) > 200


@pytest.mark.parametrize("ncomp,styles", [(1, [2]), (3, [0, 1]), (5, [0, 1])])
def test_pictures_from_a_context_match_the_host_writer(tmp_path, ncomp, styles):
    """gpe_write_bmp of a level after steps equals gpe_host_write_bmp of the state read back."""
    from spacetime_amd import pictures
    cfg = _cfg(gpe.FP64, ncomp=ncomp, dims=(5, 4, 3))
    with gpe.Stepper(cfg) as st:
        st.set_state(normalized(random_state(cfg.shape, 21), cfg))
        st.start(gpe.START_EULER)
        st.step(4)
        psi = st.get_state(1)
        for style in styles:
            for basis in ([gpe.BASIS_Z] if ncomp == 1 or style == 1 and ncomp == 5 else [gpe.BASIS_Z, gpe.BASIS_Y]):
                name = "%d_%d" % (style, basis)
                pictures.save(st, tmp_path / ("dev_%s.bmp" % name), level=1, style=style, basis=basis)
                pictures.write(tmp_path / ("host_%s.bmp" % name), cfg, psi, style, basis)
                dev = (tmp_path / ("dev_%s.bmp" % name)).read_bytes()
                assert dev == (tmp_path / ("host_%s.bmp" % name)).read_bytes(), name
                _, rgb = pictures.read(tmp_path / ("dev_%s.bmp" % name))
                assert rgb.max() > 0


def test_series_file_in_the_reference_format(tmp_path):
    """The .m time series (Spin1GpeBCC.cu:1497-1504,1659-1712): eight lines "name = [name, v, ...];"
    holding, per append, the time on the ramp's clock, Bq, Bb x + y + z, and the reference's
    Spin, mag_x, mag_y, mag_z and norm of the level; a save starts a new series."""
    import re
    cfg = _cfg(gpe.FP64)
    psi = normalized(random_state(cfg.shape, 31), cfg)
    ramp_scale = 7.5
    q = [(0.3, 0.001, 0), (0.6, 0.1, 1)]
    b = [((0.01, -0.02, 0.05), 0.002, 0), ((0.03, 0.01, -0.04), 0.1, 1)]
    with gpe.Stepper(cfg) as st, gpe.Series() as ser:
        st.set_ramp(q_segments=q, b_segments=b, time_scale=ramp_scale)
        st.set_state(psi)
        st.start(gpe.START_EULER)
        want = []
        for k in range(3):
            st.step(5)
            o = st.observables(1)
            ser.append(st, 1)
            want.append(o)
        ser.save(tmp_path / "a.m")
        st.step(2)
        ser.append(st, 0)
        o_last = st.observables(0)
        ser.save(tmp_path / "b.m")
    lines = (tmp_path / "a.m").read_text().splitlines()
    names = ["times", "Bq", "Bz", "Spin", "mag_x", "mag_y", "mag_z", "norm"]
    assert [ln.split(" = ")[0] for ln in lines] == names
    cols = {}
    for name, ln in zip(names, lines):
        m = re.fullmatch(r"%s = \[%s((?:, -?\d+\.\d{18})*)\];" % (name, name), ln)
        assert m, ln
        cols[name] = [float(v) for v in m.group(1).split(", ")[1:]]
        assert len(cols[name]) == 3
    for k, o in enumerate(want):
        assert cols["times"][k] == pytest.approx(o.time * ramp_scale, abs=1e-15)
        assert cols["Spin"][k] == pytest.approx(o.ref_spin, rel=1e-15, abs=1e-17)
        assert cols["mag_x"][k] == pytest.approx(o.magnetization[0], abs=1e-17)
        assert cols["mag_y"][k] == pytest.approx(o.magnetization[1], abs=1e-17)
        assert cols["mag_z"][k] == pytest.approx(o.ref_magnetization_z, abs=1e-17)
        assert cols["norm"][k] == pytest.approx(o.norm, rel=1e-15)
    # the field columns: the ramps at those times, inside their second (linear) segments
    for k, tr in enumerate(cols["times"]):
        assert 0.002 < tr < 0.1
        assert cols["Bq"][k] == pytest.approx(0.3 + 0.3 * (tr - 0.001) / 0.1, rel=1e-12)
        assert cols["Bz"][k] == pytest.approx(0.04 - 0.04 * (tr - 0.002) / 0.1, rel=1e-12)
    b_lines = (tmp_path / "b.m").read_text().splitlines()
    assert b_lines[0].startswith("times = [times, ") and b_lines[0].count(",") == 1  # a new series
    assert float(b_lines[7][len("norm = [norm, "):-2]) == pytest.approx(o_last.norm, rel=1e-15)


def test_fp64_trajectory_adjoint_of_an_fp32_context():
    """fp64_trajectory on an FP32 context: the same J and gradients as an FP64 context's
    adjoint started from the state the FP32 context holds (psi rounded to FP32), and the
    context afterwards holds psi_N, psi_{N-1} rounded to FP32."""
    from spacetime_amd import states
    cfg32 = _cfg(gpe.FP32, dims=(6, 5, 4))
    cfg64 = _cfg(gpe.FP64, dims=(6, 5, 4))
    psi = normalized(random_state(cfg32.shape, 30), cfg32).astype(np.complex64).astype(np.complex128)
    target = normalized(random_state(cfg32.shape, 31), cfg32)
    res = {}
    fin = {}
    for name, cfg, flag in (("fp64", cfg64, False), ("mixed", cfg32, True)):
        with gpe.Stepper(cfg) as st:
            st.set_ramp(q_segments=[(0.3, 1e300, 0)], b_segments=states.knot_ramp(_knots(), 40 * cfg.dt))
            st.set_state(psi)
            res[name] = st.adjoint(40, 8, gpe.OBJ_OVERLAP, target, fp64_trajectory=flag)
            fin[name] = (st.get_state(0), st.get_state(1), st.time())
    a, b = res["fp64"], res["mixed"]
    assert b["J"] == pytest.approx(a["J"], rel=1e-13)
    np.testing.assert_allclose(b["grad_trap"], a["grad_trap"], rtol=1e-11)
    np.testing.assert_allclose(b["grad_ramp_b"], a["grad_ramp_b"], rtol=1e-11, atol=1e-14 * np.abs(a["grad_ramp_b"]).max())
    assert b["checkpoint_bytes"] == a["checkpoint_bytes"]
    assert fin["mixed"][2] == fin["fp64"][2]
    for lvl in (0, 1):
        assert rel_l2(fin["mixed"][lvl], fin["fp64"][lvl]) < 1e-7
