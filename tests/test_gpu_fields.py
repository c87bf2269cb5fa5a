"""Per-unknown field maps (gpe_local_fields) and the spin / first-moment observables against
textbook spin algebra on random states (dense F matrices from the angular-momentum ladder),
the reference's integrated "Spin" and mag_z against its formulas and closed forms on uniform
states, plus closed forms of the spin-1 (u, v, theta) decomposition on polar and ferromagnetic
spinors.  FP64 contexts: 1e-12 relative."""
import numpy as np
import pytest

from helpers import random_state

pytestmark = pytest.mark.gpu

gpe = pytest.importorskip("spacetime_amd.gpe")


def _spin_matrices(F):
    m = np.arange(F, -F - 1, -1).astype(float)
    n = len(m)
    fp = np.zeros((n, n))
    for i in range(n - 1):
        mm = m[i + 1]
        fp[i, i + 1] = np.sqrt(F * (F + 1) - mm * (mm + 1))
    return [0.5 * (fp + fp.T) + 0j, -0.5j * (fp - fp.T), np.diag(m) + 0j]


def _cfg(ncomp, lattice="bcc", dims=(5, 4, 3)):
    return gpe.Config.centred_box(lattice, dims, 0.5, ncomp=ncomp, precision=gpe.FP64, dt=1e-4,
                                  trap=(1.0, 1.1, 0.9), c0=50.0, c1=5.0, c2=-3.0)


@pytest.mark.parametrize("ncomp,lattice", [(3, "bcc"), (5, "fcc"), (1, "a15")])
def test_density_spin_maps_and_integrals(ncomp, lattice):
    from spacetime_amd import states
    cfg = _cfg(ncomp, lattice)
    psi = random_state(cfg.shape, 11)
    psi[:, 0, 0, 0, 0] = 0.0  # one empty unknown: <F> defined as 0 there
    with gpe.Stepper(cfg) as st:
        st.set_state(psi)
        dens = st.local_fields(gpe.FIELD_DENSITY)[..., 0]
        spin = st.local_fields(gpe.FIELD_SPIN)
        snorm = st.local_fields(gpe.FIELD_SPIN_NORM)[..., 0]
        obs = st.observables(0)
        if ncomp != 3:
            with pytest.raises(gpe.GpeError, match="EINVAL"):
                st.local_fields(gpe.FIELD_UV_THETA)
    n = np.sum(np.abs(psi) ** 2, axis=0)
    np.testing.assert_allclose(dens, n, rtol=1e-13, atol=0)
    if ncomp > 1:
        f = _spin_matrices((ncomp - 1) // 2)
        Fv = np.stack([np.einsum("czyxv,cd,dzyxv->zyxv", psi.conj(), fa, psi).real for fa in f], axis=-1)
    else:
        Fv = np.zeros(n.shape + (3,))
    with np.errstate(invalid="ignore", divide="ignore"):
        expect = np.where(n[..., None] > 0, Fv / n[..., None], 0.0)
    np.testing.assert_allclose(spin, expect, rtol=0, atol=1e-12 * max(1.0, np.abs(expect).max()))
    assert np.all(spin[0, 0, 0, 0] == 0.0)
    np.testing.assert_allclose(snorm, np.linalg.norm(Fv, axis=-1), rtol=0, atol=1e-12 * n.max())
    dv = states.dual_volumes(cfg)
    r = states.positions(cfg)
    assert obs.spin == pytest.approx(np.sum(np.linalg.norm(Fv, axis=-1) * dv), rel=1e-12, abs=1e-14)
    np.testing.assert_allclose(obs.moment, np.sum((n * dv)[..., None] * r, axis=(0, 1, 2, 3)), rtol=1e-12,
                               atol=1e-12 * np.sum(n * dv))
    np.testing.assert_allclose(obs.magnetization, np.sum(Fv * dv[:, None], axis=(0, 1, 2, 3)), rtol=1e-12,
                               atol=1e-12 * np.sum(n * dv))
    # the reference's "Spin" and mag_z (localAvgSpinAndDensity + integrateSpinAndDensity): x, y of the
    # spinor divided by sqrt(n) on m = 1, 0, -1, z = |psi_1|^2 - |psi_-1|^2 undivided, weighted by n
    ref_spin = ref_mz = 0.0
    if ncomp > 1:
        o = 1 if ncomp == 5 else 0
        p1, p0, pm = psi[o], psi[o + 1], psi[o + 2]
        with np.errstate(invalid="ignore", divide="ignore"):
            u1, u0, um = (np.where(n > 0, q / np.sqrt(n), 0.0) for q in (p1, p0, pm))
        t = np.sqrt(2.0) * (u1.conj() * u0 + u0.conj() * um)
        zr = np.abs(p1) ** 2 - np.abs(pm) ** 2
        ref_spin = np.sum(n * np.sqrt(t.real ** 2 + t.imag ** 2 + zr ** 2) * dv)
        ref_mz = np.sum(n * zr * dv)
    assert obs.ref_spin == pytest.approx(ref_spin, rel=1e-12, abs=1e-14)
    assert obs.ref_magnetization_z == pytest.approx(ref_mz, rel=1e-12, abs=1e-12 * np.sum(n * n * dv))


@pytest.mark.parametrize("name,spinor,ref_spin_per_n2,ref_mz_per_n2", [
    ("ferro", (0.8, 0.0, 0.0), 1.0, 1.0),                 # |s| = |z| = n: both weighted by n once more
    ("polar", (0.0, 0.8, 0.0), 0.0, 0.0),
    ("ferro_x", (0.4, 0.8 / np.sqrt(2.0), 0.4), None, 0.0),  # transverse: |s| = 1, the physical value
])
def test_reference_spin_integrals_closed_forms(name, spinor, ref_spin_per_n2, ref_mz_per_n2):
    """Uniform spin-1 states: the reference's integrated z spin is sum dv n (|psi_1|^2 - |psi_-1|^2)
    = n^2 V for a ferromagnetic state, where the physical magnetisation is n V; a transverse
    ferromagnetic state gives its "Spin" as n V, as the physical one."""
    from spacetime_amd import states
    cfg = _cfg(3)
    psi = np.zeros(cfg.shape, complex)
    psi[:] = (np.exp(0.3j) * np.array(spinor))[:, None, None, None, None]
    n = float(np.sum(np.abs(np.array(spinor)) ** 2))
    vol = float(np.sum(states.dual_volumes(cfg) * np.ones(cfg.shape[1:])))
    with gpe.Stepper(cfg) as st:
        st.set_state(psi)
        obs = st.observables(0)
    if ref_spin_per_n2 is None:
        assert obs.ref_spin == pytest.approx(n * vol, rel=1e-12)
        assert obs.spin == pytest.approx(n * vol, rel=1e-12)
    else:
        assert obs.ref_spin == pytest.approx(ref_spin_per_n2 * n * n * vol, rel=1e-12, abs=1e-14)
    assert obs.ref_magnetization_z == pytest.approx(ref_mz_per_n2 * n * n * vol, rel=1e-12, abs=1e-14)
    if name == "ferro":
        assert obs.magnetization[2] == pytest.approx(n * vol, rel=1e-12)


def test_uv_theta_spin1():
    cfg = _cfg(3)
    psi = random_state(cfg.shape, 12)
    psi[:, 0, 0, 0, 0] = (0.0, 0.7, 0.0)          # polar: a = (0, 0, 0.7), theta = 0
    psi[:, 0, 0, 0, 1] = (0.6, 0.0, 0.0)          # ferro +1: |m| = |n|, m.n = 0, theta = 0
    with gpe.Stepper(cfg) as st:
        st.set_state(psi)
        uvt = st.local_fields(gpe.FIELD_UV_THETA)
    u, v, th = uvt[..., 0:3], uvt[..., 3:6], uvt[..., 6]
    s = 1 / np.sqrt(2)
    a = np.stack([s * (psi[2] - psi[0]), -1j * s * (psi[2] + psi[0]), psi[1]], axis=-1)
    # closed forms
    np.testing.assert_allclose(uvt[0, 0, 0, 0], [0, 0, 0.7, 0, 0, 0, 0], atol=1e-15)
    np.testing.assert_allclose(uvt[0, 0, 0, 1], [-0.6 * s, 0, 0, 0, -0.6 * s, 0, 0], atol=1e-15)
    # properties: theta in [0, pi), u . v = 0, |u| >= |v|, {u + i v, v + i u...} from a e^{i theta}
    assert np.all((th >= 0) & (th < np.pi))
    np.testing.assert_allclose(np.sum(u * v, axis=-1), 0.0, atol=1e-12)
    assert np.all(np.linalg.norm(u, axis=-1) >= np.linalg.norm(v, axis=-1) - 1e-14)
    rot = a * np.exp(1j * th)[..., None]
    keep = np.linalg.norm(rot.real, axis=-1) >= np.linalg.norm(rot.imag, axis=-1)
    np.testing.assert_allclose(np.where(keep[..., None], rot.real, rot.imag), u, atol=1e-12)
    np.testing.assert_allclose(np.where(keep[..., None], rot.imag, rot.real), v, atol=1e-12)
    # the branch: tan(2 theta) = -2 m.n / (|m|^2 - |n|^2) with the sign of atan2
    m, q = a.real, a.imag
    th_ref = 0.5 * np.arctan2(-2 * np.sum(m * q, -1), np.sum(m * m, -1) - np.sum(q * q, -1))
    th_ref = np.where(th_ref < 0, th_ref + np.pi, th_ref)
    np.testing.assert_allclose(th, th_ref, atol=1e-12)


def test_save_volume_matches_field_maps(tmp_path):
    """The volume file of a spin-1 state, written from the device level in the reference's
    padded layout (points z, x, y, v over the grid with a one-block halo), carries the device
    field maps at the interior points: the density everywhere, the local spin with the
    reference's z part (psi^dag Fz psi, not divided by n) and zeros below the density
    threshold; the halo points are zero."""
    from spacetime_amd import lattice, states, vtk
    cfg = _cfg(3)
    psi = random_state(cfg.shape, 5)
    psi[:, 1, 2, 3, :4] *= 1e-3  # a few unknowns below the density threshold
    with gpe.Stepper(cfg) as st:
        st.set_state(psi)
        dens = st.local_fields(gpe.FIELD_DENSITY)[..., 0]
        spin = st.local_fields(gpe.FIELD_SPIN)
        vtk.save_volume(st, tmp_path / "s.vtk")
    got = vtk.read_volume(tmp_path / "s.vtk")
    nz, ny, nx, V = cfg.shape[1:]
    w = lattice.get(cfg.lattice).block_dims
    z, y, x, v = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), np.arange(V), indexing="ij")
    k = ((((z + 1) * (nx + 2) + x + 1) * (ny + 2) + y + 1) * V + v).reshape(-1)
    npad = (nz + 2) * (nx + 2) * (ny + 2) * V
    assert got["points"].shape == (npad, 3)
    shifted = states.positions(cfg) + cfg.block_scale * w  # one block of halo before index 0
    np.testing.assert_allclose(got["points"][k], shifted.reshape(-1, 3).astype(np.float32), rtol=1e-6, atol=1e-6)
    dens, spin = dens.reshape(-1), spin.reshape(-1, 3)
    live = dens > vtk.DENSITY_THRESHOLD
    assert 0 < live.sum() < live.size
    np.testing.assert_allclose(got["density"][k], dens, rtol=1e-6)
    ref_spin = np.where(live[:, None], spin * np.array([1.0, 1.0, 0.0]) + np.outer(spin[:, 2] * dens, [0, 0, 1]), 0)
    np.testing.assert_allclose(got["localAvgSpin"][k], ref_spin, rtol=1e-5, atol=1e-6)
    np.testing.assert_allclose(got["spinNorm"][k], np.linalg.norm(ref_spin, axis=1), rtol=1e-5, atol=1e-6)
    halo = np.ones(npad, bool)
    halo[k] = False
    for name in ("density", "spinNorm", "theta"):
        assert not np.any(got[name][halo]), name
