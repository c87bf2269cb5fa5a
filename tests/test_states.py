"""Initial states built by the product's host code (libgpe gpe_vortex_*, spacetime_amd.states)
against the compiled reference's VortexState numbers and the oracle's GPBlocks restatement;
closed-form checks of the Thomas-Fermi profile, imprints, noise and projections."""
import copy
import json
import os

import numpy as np
import pytest

from spacetime_amd import gpe, lattice, states

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "gpblocks_scalar_bcc.json")


def _golden():
    with open(GOLDEN) as f:
        return json.load(f)["cases"]


@pytest.mark.parametrize("case", [0, 1])
def test_vortex_relaxation_matches_reference(case):
    c = _golden()[case]
    vs = states.vortex_state(c["kappa"], c["g"], c["rmax"], c["zmax"], c["rstep"], c["zstep"], maxiters=c["maxiters"])
    assert vs["mu"] == pytest.approx(c["vortex_mu"], rel=1e-13)
    assert vs["rsize"] * vs["zsize"] == len(vs["f"])
    # on-grid samples of the profile: the reference printed f(r, 0) at r = 0.15 i
    rstep = vs["params"].rstep
    for r, _z, _f, f0 in c["vortex_f"]:
        k = r / rstep
        if abs(k - round(k)) < 1e-9 and round(k) < vs["rsize"]:
            assert vs["f"][int(round(k)) * vs["zsize"]] == pytest.approx(f0, rel=1e-12, abs=1e-15)


@pytest.mark.parametrize("case", [0, 1])
def test_vortex_fill_matches_oracle_initial_state(oracle, case):
    c = _golden()[case]
    vs = states.vortex_state(c["kappa"], c["g"], c["rmax"], c["zmax"], c["rstep"], c["zstep"], maxiters=c["maxiters"])
    ov = oracle.vortex(c["kappa"], c["g"], c["rmax"], c["zmax"], c["rstep"], c["zstep"], maxiters=c["maxiters"])
    np.testing.assert_allclose(vs["f"], ov["f"], rtol=1e-13, atol=1e-15)
    lat = lattice.get("bcc")
    s = c["block_scale"]
    n = [int((2 * m) / (s * lat.block_width)) + 1 for m in (c["rmax"], c["rmax"], c["zmax"])]
    p0 = [0.5 * (-2 * 0) - 0.5 * s * lat.block_width * k for k in n]
    cfg = gpe.Config(lattice="bcc", ncomp=1, dims=tuple(n), block_scale=s, p0=p0)
    psi = states.vortex_fill(cfg, vs)
    # oracle's VortexState::getPsi at the same positions
    r = states.positions(cfg)
    want = np.zeros(cfg.shape, dtype=np.complex128)
    it = np.nditer(r[..., 0], flags=["multi_index"])
    for _ in it:
        z, y, x, v = it.multi_index
        px, py, pz = r[z, y, x, v]
        rr = np.sqrt(px * px + py * py)
        if rr < 1e-30:
            continue
        want[0, z, y, x, v] = oracle.vortex_function(ov, rr, pz) * ((px + 1j * py) / rr) ** c["kappa"]
    np.testing.assert_allclose(psi, want, rtol=1e-12, atol=1e-14)
    assert np.abs(psi).max() > 0.05


def test_thomas_fermi_norm_and_mu():
    cfg = gpe.Config.centred_box("bcc", (12, 12, 12), 0.8, ncomp=1, trap=(1.0, 1.0, 1.0))
    tf, mu = states.thomas_fermi(cfg, 200.0)
    psi = tf[None].astype(np.complex128)
    assert states.norm(cfg, psi) == pytest.approx(1.0, rel=1e-12)
    # continuum TF chemical potential 0.5 (15 g / 4 pi)^(2/5) for an isotropic trap
    assert mu == pytest.approx(0.5 * (15 * 200.0 / (4 * np.pi)) ** 0.4, rel=0.05)


def test_imprints_noise_and_projections():
    cfg = gpe.Config.centred_box("bcc", (6, 6, 6), 0.8, ncomp=5)
    r = states.positions(cfg)
    line = states.vortex_line_phase(r, 2)
    assert np.allclose(np.abs(line), 1.0)
    x, y = r[..., 0], r[..., 1]
    assert np.allclose(line, ((x + 1j * y) / np.hypot(x, y)) ** 2)
    ring = states.vortex_ring_phase(r, 2.0)
    rho = np.hypot(x, y)
    assert np.allclose(np.angle(ring), np.arctan2(r[..., 2], rho - 2.0))
    psi = np.ones(cfg.shape, dtype=np.complex128)
    a = states.spinor_noise(psi, 0.005, 7)
    b = states.spinor_noise(psi, 0.005, 7)
    assert np.array_equal(a, b) and np.abs(a - psi).max() <= 0.005 * np.sqrt(5) + 1e-15
    n = np.sum(np.abs(a) ** 2, axis=0)
    for name in ("un", "bn_hori", "bn_vert", "cyclic", "ferro"):
        pr = states.project(a, name, rel_phase=0.3)
        assert np.allclose(np.sum(np.abs(pr) ** 2, axis=0), n)
    # cyclic state has zero spin density, ferro has |F| = 2 n
    cyc = states.project(a, "cyclic")
    assert np.allclose(np.abs(cyc[0]) ** 2 * 3, n) and np.allclose(np.abs(cyc[3]) ** 2 * 1.5, n)


def test_reference_dat_roundtrip_and_layout(tmp_path):
    from spacetime_amd import gpe
    rng = np.random.default_rng(9)
    cfg = gpe.Config(lattice="bcc", ncomp=3, dims=(3, 4, 2))
    psi = rng.normal(size=(3, 2, 4, 3, 12)) + 1j * rng.normal(size=(3, 2, 4, 3, 12))
    path = tmp_path / "state.dat"
    states.write_reference_dat(str(path), cfg, psi)
    raw = np.fromfile(str(path), dtype=np.float64)
    assert raw.size == 5 * 6 * 4 * 12 * 3 * 2
    # block (i, j, k) = (2, 1, 0), unknown 7, component m = 0 sits where the reference puts it:
    # dstI = (k+1)(ny+2)(nx+2) + (j+1)(nx+2) + (i+1); Complex3Vec {s1, s0, s_1}
    i, j, k, l = 2, 1, 0, 7
    dst = (k + 1) * 6 * 5 + (j + 1) * 5 + (i + 1)
    off = ((dst * 12 + l) * 3 + 1) * 2
    assert raw[off] == psi[1, k, j, i, l].real and raw[off + 1] == psi[1, k, j, i, l].imag
    assert np.all(raw.reshape(4, 6, 5, 12, 3, 2)[0] == 0)  # halo plane stays zero
    back = states.read_reference_dat(str(path), cfg)
    np.testing.assert_array_equal(back, psi)
    # a file of another grid is refused, not silently misread
    with pytest.raises(ValueError):
        states.read_reference_dat(str(path), gpe.Config(lattice="bcc", ncomp=3, dims=(3, 4, 3)))
