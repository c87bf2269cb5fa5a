"""Pins for the FP64 CPU oracle (oracle/gpe_oracle.c).

Each check compares the oracle against something it does not share code with:
  * golden numbers printed by the compiled reference (GPBlocks.cpp + VortexState.cpp),
  * numpy's own MT19937,
  * dense textbook spin matrices built here from angular-momentum algebra,
  * the energy functional's gradient (dE/dpsi* = dv H psi), by finite differences,
  * exact reductions of the spinor equations to the scalar one,
  * a dense Laplacian assembled here from the table by brute force,
  * conservation laws of the 3-level scheme.
"""
import copy
import json
import os

import numpy as np
import pytest

from spacetime_amd import lattice

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "gpblocks_scalar_bcc.json")


def _golden():
    with open(GOLDEN) as f:
        return json.load(f)


# ----------------------------------------------------------------------------- reference --

@pytest.mark.parametrize("case", [0, 1])
def test_vortex_state_matches_reference(oracle, case):
    c = _golden()["cases"][case]
    vs = oracle.vortex(c["kappa"], c["g"], c["rmax"], c["zmax"], c["rstep"], c["zstep"], maxiters=c["maxiters"])
    assert vs["mu"] == pytest.approx(c["vortex_mu"], rel=1e-14)
    for r, z, f, f0 in c["vortex_f"]:
        assert oracle.vortex_function(vs, r, z) == pytest.approx(f, rel=1e-12, abs=1e-15)
        assert oracle.vortex_function(vs, r, 0.0) == pytest.approx(f0, rel=1e-12, abs=1e-15)


@pytest.mark.parametrize("case", [0, 1])
def test_staggered_scalar_run_matches_reference(oracle, case):
    """normsq and overlap error per iteration of GPBlocks.cpp integrateInTime, to 1e-13."""
    c = _golden()["cases"][case]
    lat = copy.copy(lattice.get("bcc"))
    # GPBlocks positions are (1/sqrt 8) * integer vectors; use them exactly (the vortex
    # interpolant of VortexState.cpp:236 is discontinuous in its last z row).
    lat.pos = np.array(c["bcc_pos"])
    L = oracle.lattice(lat)
    vs = oracle.vortex(c["kappa"], c["g"], c["rmax"], c["zmax"], c["rstep"], c["zstep"], maxiters=c["maxiters"])
    ns, er, dt, steps = oracle.gpblocks_integrate(
        L, vs, c["kappa"], c["g"], c["block_scale"], [-c["rmax"], -c["rmax"], -c["zmax"]],
        [c["rmax"], c["rmax"], c["zmax"]], c["period"], c["iterations"])
    assert steps > 1
    np.testing.assert_allclose(ns, c["normsq"], rtol=1e-13)
    np.testing.assert_allclose(er, c["error"], rtol=0, atol=1e-13)
    # the error is not trivially zero: the run really stepped
    assert max(abs(e) for e in c["error"][1:]) > 1e-5


def test_mt19937_matches_numpy(oracle):
    rs = np.random.RandomState(54363)
    raw = rs.randint(0, 2 ** 32, size=4000, dtype=np.uint64).astype(np.uint32)
    signed = raw.view(np.int32).astype(np.float64) / 2147483648.0
    pairs = signed.reshape(-1, 2)
    expected = pairs[(pairs ** 2).sum(1) <= 1.0][:500]
    got = oracle.uniform_circle(54363, 500)
    np.testing.assert_array_equal(got, expected)


# ------------------------------------------------------------------------ spin matrices --

def _spin_matrices(F):
    m = np.arange(F, -F - 1, -1).astype(float)
    n = len(m)
    fz = np.diag(m)
    fp = np.zeros((n, n))
    for i in range(n - 1):  # <m+1| f+ |m> with index i+1 -> i
        mm = m[i + 1]
        fp[i, i + 1] = np.sqrt(F * (F + 1) - mm * (mm + 1))
    fx = 0.5 * (fp + fp.T)
    fy = -0.5j * (fp - fp.T)
    return fx.astype(complex), fy, fz.astype(complex)


def _dense_local_h(F, prm, r, psi, t=0.0):
    n = 2 * F + 1
    dens = np.vdot(psi, psi).real
    V = 0.5 * sum((prm["trap"][a] * r[a]) ** 2 for a in range(3)) + prm["trap_offset"]
    H = (V + prm["c0"] * dens - 1j * prm["alpha"] * dens ** 2) * np.eye(n, dtype=complex)
    if F > 0:
        f = _spin_matrices(F)
        Fv = np.array([np.vdot(psi, fi @ psi).real for fi in f])
        ctr = prm["center"]
        B = np.array([prm["bq"] * (r[0] - ctr[0]), prm["bq"] * (r[1] - ctr[1]), -2 * prm["bq"] * (r[2] - ctr[2])])
        B = B + np.array(prm["bb"])
        s = 1.0 if F == 1 else -1.0
        W = prm["c1"] * Fv + s * B
        H = H + sum(W[a] * f[a] for a in range(3))
        Bq = np.array([prm["bq_quad"] * (r[0] - ctr[0]), prm["bq_quad"] * (r[1] - ctr[1]),
                       -2 * prm["bq_quad"] * (r[2] - ctr[2])]) + np.array(prm["bb_quad"])
        BF = sum(Bq[a] * f[a] for a in range(3))
        H = H + prm["quad_sign"] * (BF @ BF) + prm["q"] * (f[2] @ f[2])
        out = H @ psi
        if F == 2:
            # singlet amplitude A00 = (2 psi2 psi-2 - 2 psi1 psi-1 + psi0^2)/sqrt5 ; E = c2/2 |A00|^2
            theta = 2 * psi[0] * psi[4] - 2 * psi[1] * psi[3] + psi[2] ** 2
            sign = np.array([1, -1, 1, -1, 1])
            out = out + prm["c2"] / 5.0 * theta * sign * np.conj(psi[::-1])
        return out
    return H @ psi


@pytest.mark.parametrize("F", [0, 1, 2])
def test_local_h_equals_dense_spin_matrices(oracle, F):
    rng = np.random.default_rng(11 + F)
    for trial in range(20):
        prm = dict(trap=list(rng.uniform(0.5, 1.5, 3)), trap_offset=rng.uniform(0, 3), c0=rng.uniform(-50, 200),
                   c1=rng.uniform(-30, 30), c2=rng.uniform(-20, 20), alpha=rng.uniform(0, 2) * (trial % 2),
                   bq=rng.uniform(-1, 1), bb=list(rng.uniform(-1, 1, 3)), bq_quad=rng.uniform(-1, 1),
                   bb_quad=list(rng.uniform(-1, 1, 3)), quad_sign=rng.choice([-1.0, 1.0]), q=rng.uniform(-1, 1),
                   center=list(rng.uniform(-0.5, 0.5, 3)))
        p = oracle.params(2 * F + 1, (1, 1, 1), 1.0, (0, 0, 0), **prm)
        psi = rng.normal(size=2 * F + 1) + 1j * rng.normal(size=2 * F + 1)
        r = rng.uniform(-2, 2, 3)
        got = oracle.local_h(p, r, psi)
        want = _dense_local_h(F, prm, r, psi)
        np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12 * np.abs(want).max())


# ----------------------------------------------------------------------- energy gradient --

def _grid_params(oracle, lat_name, ncomp, dims, rng, **kw):
    lat = lattice.get(lat_name)
    edge = 0.9
    s = edge / lat.block_width
    p0 = [-0.5 * edge * d for d in dims]
    return lat, oracle.lattice(lat), oracle.params(ncomp, dims, s, p0, **kw)


@pytest.mark.parametrize("lat_name,F", [("bcc", 0), ("bcc", 1), ("bcc", 2), ("fcc", 1), ("a15", 2), ("cube", 1)])
def test_energy_gradient_is_h(oracle, lat_name, F):
    """d/de E(psi + e eta) at 0 == 2 Re sum dv eta^* (H psi) for the Hermitian part of H."""
    rng = np.random.default_rng(3)
    dims = (2, 2, 2)
    prm = dict(trap=(1.0, 0.8, 1.3), trap_offset=0.7, c0=40.0, c1=7.0, c2=-5.0, bq=0.3, bb=(0.1, -0.2, 0.4),
               bq_quad=0.2, bb_quad=(0.05, 0.1, -0.3), q=0.25, center=(0.05, -0.02, 0.01))
    lat, L, p = _grid_params(oracle, lat_name, 2 * F + 1, dims, rng, **prm)
    shape = (2 * F + 1, dims[2], dims[1], dims[0], lat.V)
    psi = (rng.normal(size=shape) + 1j * rng.normal(size=shape)) * 0.3
    eta = rng.normal(size=shape) + 1j * rng.normal(size=shape)
    hpsi = oracle.apply_h(L, p, psi)
    dv = p.block_scale ** 3 * lat.vol  # per unknown
    analytic = 2 * np.sum(dv[None, None, None, None, :] * (np.conj(eta) * hpsi).real)
    h = 1e-5
    fd = (oracle.energy(L, p, psi + h * eta) - oracle.energy(L, p, psi - h * eta)) / (2 * h)
    assert fd == pytest.approx(analytic, rel=1e-7)


# -------------------------------------------------------------- brute-force Laplacian ---

@pytest.mark.parametrize("lat_name", ["bcc", "fcc", "cube", "a15"])
def test_kinetic_term_matches_dense_assembly(oracle, lat_name):
    rng = np.random.default_rng(5)
    dims = (3, 2, 2)
    lat, L, p = _grid_params(oracle, lat_name, 1, dims, rng, trap=(0.0, 0.0, 0.0))
    nx, ny, nz = dims
    V = lat.V
    N = nx * ny * nz * V
    A = np.zeros((N, N))
    idx = lambda x, y, z, v: ((z * ny + y) * nx + x) * V + v  # noqa: E731
    for z in range(nz):
        for y in range(ny):
            for x in range(nx):
                for f in range(lat.nfaces):
                    v = lat.face_src[f]
                    i = idx(x, y, z, v)
                    w = -0.5 * lat.face_hodge[f] / p.block_scale ** 2
                    X, Y, Z = x + lat.face_d[f][0], y + lat.face_d[f][1], z + lat.face_d[f][2]
                    A[i, i] -= w
                    if 0 <= X < nx and 0 <= Y < ny and 0 <= Z < nz:
                        A[i, idx(X, Y, Z, lat.face_tgt[f])] += w
    psi = rng.normal(size=N) + 1j * rng.normal(size=N)
    got = oracle.apply_h(L, p, psi.reshape(1, nz, ny, nx, V)).reshape(-1)
    np.testing.assert_allclose(got, A @ psi, rtol=1e-12, atol=1e-10)
    # self-adjoint in the volume-weighted inner product
    D = np.tile(lat.vol, nx * ny * nz)
    np.testing.assert_allclose(D[:, None] * A, (D[:, None] * A).T, rtol=1e-12, atol=1e-9)


# ----------------------------------------------------------- reductions to the scalar ---

def _run_pair(oracle, L, p, psi, dt, n):
    a = psi.copy()
    b = oracle.forward_euler(L, p, a, dt)
    oracle.leapfrog3(L, p, a, b, dt, n)
    return a, b


@pytest.mark.parametrize("F,comp,geff", [
    (1, 1, "c0"),          # polar (0, phi, 0): F = 0
    (1, 0, "c0+c1"),       # ferromagnetic (phi, 0, 0): F = n
    (2, 2, "c0+c2/5"),     # uniaxial nematic (0, 0, phi, 0, 0)
    (2, 0, "c0+4*c1"),      # spin-2 ferromagnetic
])
def test_single_component_reduces_to_scalar(oracle, F, comp, geff):
    rng = np.random.default_rng(9)
    dims = (3, 3, 2)
    c0, c1, c2 = 60.0, 9.0, 11.0
    g = eval(geff, {"c0": c0, "c1": c1, "c2": c2})
    bz = 0.37
    m = F - comp
    lat, L, p = _grid_params(oracle, "bcc", 2 * F + 1, dims, rng, c0=c0, c1=c1, c2=c2, bb=(0, 0, bz))
    s = 1.0 if F == 1 else -1.0
    _, Ls, ps = _grid_params(oracle, "bcc", 1, dims, rng, c0=g, trap_offset=s * m * bz)
    shape = (dims[2], dims[1], dims[0], lat.V)
    phi = (rng.normal(size=shape) + 1j * rng.normal(size=shape)) * 0.2
    psi = np.zeros((2 * F + 1,) + shape, dtype=complex)
    psi[comp] = phi
    dt = 2e-4
    a, b = _run_pair(oracle, L, p, psi, dt, 25)
    sa, sb = _run_pair(oracle, Ls, ps, phi[None].copy(), dt, 25)
    np.testing.assert_allclose(b[comp], sb[0], rtol=1e-12, atol=1e-13)
    others = np.delete(b, comp, axis=0)
    assert np.abs(others).max() < 1e-14


def test_leapfrog_conserves_norm_and_mz(oracle):
    rng = np.random.default_rng(4)
    dims = (3, 3, 3)
    lat, L, p = _grid_params(oracle, "bcc", 3, dims, rng, c0=50.0, c1=-8.0, bb=(0, 0, 0.3), trap=(1, 1, 1))
    shape = (3, dims[2], dims[1], dims[0], lat.V)
    psi = (rng.normal(size=shape) + 1j * rng.normal(size=shape)) * 0.2
    oracle.normalize(L, p, psi)
    n0 = oracle.norm(L, p, psi)
    mz0 = oracle.magnetization(L, p, psi)[2]
    e0 = oracle.energy(L, p, psi)
    a, b = _run_pair(oracle, L, p, psi, 1e-4, 400)
    assert oracle.norm(L, p, b) == pytest.approx(n0, rel=2e-4)
    assert oracle.magnetization(L, p, b)[2] == pytest.approx(mz0, rel=2e-3, abs=2e-4)
    assert oracle.energy(L, p, b) == pytest.approx(e0, rel=5e-3)


def test_itp_lowers_energy_and_normalises(oracle):
    rng = np.random.default_rng(2)
    dims = (4, 4, 4)
    lat, L, p = _grid_params(oracle, "bcc", 1, dims, rng, c0=30.0)
    shape = (1, dims[2], dims[1], dims[0], lat.V)
    psi = np.abs(rng.normal(size=shape)).astype(complex)
    oracle.normalize(L, p, psi)
    e = [oracle.energy(L, p, psi)]
    for _ in range(30):
        oracle.itp_step(L, p, psi, 2e-4)
        e.append(oracle.energy(L, p, psi))
    assert oracle.norm(L, p, psi) == pytest.approx(1.0, rel=1e-13)
    assert all(e[i + 1] < e[i] for i in range(len(e) - 1))


# -------------------------------------------------------------------------------- ramps ---

def test_ramp_segments(oracle):
    # spin-2 knot ramp shape (Spin2GpeBCC/KnotRamps.h:23-30) with tau = 0.120
    r = oracle.ramp(q_segments=[(4.3, 1.0, 0), (0.0, 0.5, 1), (0.0, 100.0, 0)],
                    b_segments=[((0, 0, 0.219), 0.1, 0), ((0, 0, 0), 0.0177, 1), ((0, 0, 0), 0.255, 0),
                                ((0, 0, 3.0), 100.0, 3)], tau=0.120)
    bq, bb = oracle.ramp_signal(r, 0.5)
    assert bq == 4.3 and bb[2] == pytest.approx(3.0 * (1 - np.exp(-(0.5 - 0.1 - 0.0177 - 0.255) / 0.12)))
    bq, bb = oracle.ramp_signal(r, 1.25)
    assert bq == pytest.approx(4.3 * 0.5)
    bq, bb = oracle.ramp_signal(r, 0.05)
    assert bb[2] == pytest.approx(0.219)
    bq, bb = oracle.ramp_signal(r, 0.1 + 0.0177 / 2)
    assert bb[2] == pytest.approx(0.219 / 2)
    t = 0.1 + 0.0177 + 0.255 + 0.3
    bq, bb = oracle.ramp_signal(r, t)
    assert bb[2] == pytest.approx(3.0 * (1 - np.exp(-0.3 / 0.12)))
    bq, bb = oracle.ramp_signal(r, 1e4)  # past the end: hold the last value
    assert bq == 0.0 and bb[2] == 3.0
    # sqrt "fast extraction" of the spin-1 ramps
    r1 = oracle.ramp(q_segments=[(1.0, 1.0, 0), (3.0, 4.0, 2)])
    bq, _ = oracle.ramp_signal(r1, 2.0)
    assert bq == pytest.approx(1.0 + 2.0 * np.sqrt(0.25))
    # time scale
    r2 = oracle.ramp(b_segments=[((0, 0, 0.0), 0.0, 0), ((0, 0, 2.0), 10.0, 1)], time_scale=1e3)
    _, bb = oracle.ramp_signal(r2, 2.5e-3)
    assert bb[2] == pytest.approx(0.5)


# ------------------------------------------------------- Hamiltonian bound, weighted diff ---

def test_max_hamiltonian_closed_forms(oracle):
    """maxHamilton (Spin1GpeBCC.cu:144-189) on states where it has a closed form: the trap alone
    (zero state: max V over the unknowns' positions), a uniform density (+ c0 n), z- and
    x-polarised spin-1 states in a uniform bias (|B_a + c1 F_a|, 1/sqrt2 on x and y), and the
    trap switched off."""
    rng = np.random.default_rng(8)
    dims = (3, 2, 4)
    kw = dict(trap=(1.0, 1.3, 0.7), trap_offset=100.0, c0=7.0, c1=-3.0, bb=(0.2, -0.5, 0.9))
    lat, L, p1 = _grid_params(oracle, "bcc", 1, dims, rng, **kw)
    r = oracle.positions(L, p1, lat.V)
    vmax = np.max(0.5 * np.sum((r * np.array(kw["trap"])) ** 2, -1)) + 100.0
    shape1 = (1, dims[2], dims[1], dims[0], lat.V)
    assert oracle.max_hamiltonian(L, p1, np.zeros(shape1, complex)) == pytest.approx(vmax, rel=1e-15)
    a = 0.6
    uni = np.full(shape1, a * np.exp(0.3j))
    assert oracle.max_hamiltonian(L, p1, uni) == pytest.approx(vmax + 7.0 * a * a, rel=1e-14)
    _, L3, p3 = _grid_params(oracle, "bcc", 3, dims, rng, **kw)
    shape3 = (3,) + shape1[1:]
    zpol = np.zeros(shape3, complex)
    zpol[0] = a
    pot = vmax + 7.0 * a * a
    want = pot + max(0.2 / np.sqrt(2), 0.5 / np.sqrt(2), abs(0.9 - 3.0 * a * a))
    assert oracle.max_hamiltonian(L3, p3, zpol) == pytest.approx(want, rel=1e-14)
    xpol = np.zeros(shape3, complex)
    xpol[:] = (a * np.array([0.5, 2 ** -0.5, 0.5]))[:, None, None, None, None]
    want = pot + max(abs(0.2 - 3.0 * a * a) / np.sqrt(2), 0.5 / np.sqrt(2), 0.9)
    assert oracle.max_hamiltonian(L3, p3, xpol) == pytest.approx(want, rel=1e-14)
    _, L3, p3off = _grid_params(oracle, "bcc", 3, dims, rng, **dict(kw, trap_off_time=1.0))
    assert oracle.max_hamiltonian(L3, p3off, zpol, t=2.0) == pytest.approx(
        7.0 * a * a + max(0.2 / np.sqrt(2), 0.5 / np.sqrt(2), abs(0.9 - 3.0 * a * a)), rel=1e-14)
    _, L5, p5 = _grid_params(oracle, "bcc", 5, dims, rng)
    assert np.isnan(oracle.max_hamiltonian(L5, p5, np.zeros((5,) + shape1[1:], complex)))


@pytest.mark.parametrize("lat_name,ncomp", [("bcc", 3), ("fcc", 5), ("a15", 1)])
def test_weighted_diff_closed_forms(oracle, lat_name, ncomp):
    """weightedDiff + integrate (Spin1GpeBCC.cu:436-461,1620-1633): zero at psi = a; psi = 0 on a
    uniform reference gives n0 * norm(ref); a single populated unknown gives dv_v |a|^2 |psi-a|^2
    (index and volume); phase covariance and degree-4 homogeneity."""
    rng = np.random.default_rng(9)
    dims = (2, 3, 2)
    lat, L, p = _grid_params(oracle, lat_name, ncomp, dims, rng)
    shape = (ncomp, dims[2], dims[1], dims[0], lat.V)
    ref = rng.normal(size=shape) + 1j * rng.normal(size=shape)
    psi = rng.normal(size=shape) + 1j * rng.normal(size=shape)
    ph = np.exp(0.7j)
    assert oracle.weighted_diff(L, p, ph * ref, ref, ph) == pytest.approx(0.0, abs=1e-12)
    uni = np.zeros(shape, complex)
    uni[:] = (rng.normal(size=ncomp) + 1j * rng.normal(size=ncomp))[:, None, None, None, None]
    n0 = np.sum(np.abs(uni[:, 0, 0, 0, 0]) ** 2)
    assert oracle.weighted_diff(L, p, np.zeros(shape, complex), uni) == pytest.approx(
        n0 * oracle.norm(L, p, uni), rel=1e-13)
    one = np.zeros(shape, complex)
    at = (1, 2, 0, lat.V - 1)
    one[(slice(None),) + at] = ref[(slice(None),) + at]
    dv = p.block_scale ** 3 * lat.vol[lat.V - 1]
    a2 = np.sum(np.abs(ref[(slice(None),) + at]) ** 2)
    d2 = np.sum(np.abs(psi[(slice(None),) + at] - ref[(slice(None),) + at]) ** 2)
    rest = np.sum(np.abs(psi) ** 2) - np.sum(np.abs(psi[(slice(None),) + at]) ** 2)
    assert rest > 0  # psi elsewhere does not count where a = 0
    assert oracle.weighted_diff(L, p, psi, one) == pytest.approx(dv * a2 * d2, rel=1e-13)
    base = oracle.weighted_diff(L, p, psi, ref)
    assert oracle.weighted_diff(L, p, ph * psi, ref, ph) == pytest.approx(base, rel=1e-13)
    assert oracle.weighted_diff(L, p, 1.5 * psi, 1.5 * ref) == pytest.approx(1.5 ** 4 * base, rel=1e-13)
