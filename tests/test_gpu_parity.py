"""HIP path (libgpe.so through the C ABI) vs the FP64 CPU oracle on the same seeded inputs.

Tolerances: FP64 form 1e-12 relative L2 (different summation order and FMA contraction
only); FP32 form 2e-5 relative L2 over a few tens of sweeps on random states.
"""
import numpy as np
import pytest

from helpers import normalized, oracle_params, random_state, rel_l2, with_field

pytestmark = pytest.mark.gpu

gpe = pytest.importorskip("spacetime_amd.gpe")

FULL = dict(trap=(1.0, 0.9, 1.2), trap_offset=0.5, c0=60.0, c1=8.0, c2=-6.0, q=0.3, center=(0.02, -0.01, 0.03))
FIELD = dict(bq=0.4, bb=(0.1, -0.05, 0.3), bq_quad=0.2, bb_quad=(0.05, 0.02, -0.1))


def _config(lattice, ncomp, precision, dims, dt=2e-4, scheme=0, cell=0.6, **kw):
    phys = dict(FULL)
    phys.update(kw)
    return gpe.Config.centred_box(lattice, dims, cell, ncomp=ncomp, precision=precision, dt=dt, scheme=scheme, **phys)


def _run_both(oracle, cfg, nsteps, field=None, seed=1, traversal=None):
    from spacetime_amd import lattice
    lat = lattice.get(cfg.lattice)
    L = oracle.lattice(lat)
    p = oracle_params(oracle, cfg)
    if field:
        with_field(p, **field)
    psi0 = random_state(cfg.shape, seed)
    a = psi0.copy()
    b = oracle.forward_euler(L, p, a, cfg.dt)
    oracle.leapfrog3(L, p, a, b, cfg.dt, nsteps, t0=cfg.dt)
    with gpe.Stepper(cfg) as st:
        if field:
            st.set_field(**field)
        if traversal:
            st.set_traversal(*traversal)
        st.set_state(psi0)
        st.start(gpe.START_EULER)
        st.step(nsteps)
        g0, g1 = st.get_state(0), st.get_state(1)
        t, sweeps = st.time()
    assert sweeps == nsteps + 1
    assert t == pytest.approx((nsteps + 1) * cfg.dt)
    return (g0, g1), (b, a)


@pytest.mark.parametrize("lattice_name", ["cube", "bcc", "fcc", "a15", "c15", "z"])
@pytest.mark.parametrize("ncomp", [1, 3, 5])
def test_fp64_sweeps_match_oracle(oracle, lattice_name, ncomp):
    dims = (5, 4, 3) if lattice_name not in ("c15", "z") else (3, 2, 2)
    cfg = _config(lattice_name, ncomp, gpe.FP64, dims)
    (g0, g1), (o0, o1) = _run_both(oracle, cfg, 12, field=FIELD if ncomp > 1 else None)
    assert rel_l2(g0, o0) < 1e-12
    assert rel_l2(g1, o1) < 1e-12


@pytest.mark.parametrize("xpl", [1, 2])
@pytest.mark.parametrize("lattice_name", ["cube", "bcc", "fcc", "a15", "c15", "z"])
@pytest.mark.parametrize("ncomp", [1, 3, 5])
def test_fp32_sweeps_match_oracle(oracle, lattice_name, ncomp, xpl):
    """Both FP32 sweep bodies: one x per lane and x pairs per lane (odd nx: a half pair)."""
    dims = (5, 4, 3) if lattice_name not in ("c15", "z") else (3, 2, 2)
    cfg = _config(lattice_name, ncomp, gpe.FP32, dims)
    (g0, g1), (o0, o1) = _run_both(oracle, cfg, 12, field=FIELD if ncomp > 1 else None, traversal=(8, 1, xpl))
    assert rel_l2(g0, o0) < 2e-5
    assert rel_l2(g1, o1) < 2e-5


@pytest.mark.parametrize("precision,tol,xpl,nx,flags", [
    (1, 1e-12, 1, 70, 1), (0, 2e-5, 1, 70, 1), (0, 2e-5, 2, 70, 1), (0, 2e-5, 2, 129, 1), (0, 2e-5, 2, 128, 1),
    (0, 2e-5, 1, 129, 1),
    # the unknown inside the x segment in the traversal order, with and without the XCD remap
    (1, 1e-12, 1, 130, 3), (1, 1e-12, 1, 70, 2), (0, 2e-5, 2, 300, 3), (0, 2e-5, 1, 129, 2)])
def test_wide_rows_span_several_waves(oracle, precision, tol, xpl, nx, flags):
    """nx > 64 (or > 128 with x pairs) exercises several wave segments per x-line, partial last
    waves, the line-end pair clamp (nx = 128: the last pair's +x neighbour pair starts at lx - 1)
    and a lone last value (odd nx); both orders of unknowns and segments."""
    cfg = _config("bcc", 3, precision, (nx, 3, 2))
    (g0, g1), (o0, o1) = _run_both(oracle, cfg, 6, field=FIELD, traversal=(2, flags, xpl))
    assert rel_l2(g0, o0) < tol and rel_l2(g1, o1) < tol


@pytest.mark.parametrize("lattice_name,ncomp,dims,traversal", [
    ("bcc", 3, (5, 6, 3), (4, 1, 2)),   # 4 rows per wave; the second tile's rows 6, 7 are past ny
    ("fcc", 5, (40, 7, 2), (2, 0, 2)),  # 2 rows per wave; row 7 of the last group is past ny
    ("a15", 1, (33, 5, 2), (4, 1, 2)),  # 2 rows per wave over 33 values (17 lanes of 32)
    ("cube", 5, (32, 9, 3), (8, 1, 2)),  # 4 rows per wave; three rows of the last group past ny
    ("c15", 3, (6, 3, 2), (2, 1, 2)),
    ("z", 5, (7, 4, 2), (4, 1, 2)),      # the non-cubic block, 4-row groups
    ("bcc", 5, (130, 5, 2), (4, 1, 2)),  # 4-row groups over 5 waves; the second tile's group past ny
    ("a15", 3, (82, 6, 3), (4, 0, 2)),   # 4-row groups of 41 pairs a row: 3 waves, rows split mid-wave
    ("bcc", 1, (192, 4, 2), (4, 1, 2)),  # 192-value lines: 2-row groups fill all lanes (one row: 3 of 4)
    ("fcc", 1, (160, 6, 2), (4, 1, 2)),  # 160-value lines: 4-row groups of 80 pairs, rows past ny
    ("cube", 3, (293, 10, 2), (8, 1, 2)),  # 8-row groups of 147 pairs; rows 10-15 of tile 2 past ny
    ("bcc", 1, (128, 5, 3), (4, 1, 2)),  # one component, one row per wave; tile 2 holds only row 4
    ("a15", 1, (70, 6, 3), (2, 0, 2)),   # one component, 2-row tiles, no XCD remap
    ("z", 1, (66, 9, 3), (8, 1, 2)),     # one component over 33 pairs; tile 2 holds only row 8
    ("a15", 5, (82, 7, 2), (6, 1, 2)),   # 3-row groups of 41 pairs; tile 2 holds only row 6
    ("z", 3, (68, 16, 2), (14, 1, 2)),   # 7-row groups of 34 pairs over 4 waves; tile 2 holds 2 rows
    ("bcc", 1, (70, 9, 2), (7, 0, 2)),   # 7-row groups of 35 pairs, no XCD remap
    ("c15", 3, (20, 7, 2), (6, 1, 2)),   # 6-row groups of 10 pairs in one wave
    # the unknown inside the x segment: row groups over several waves
    ("bcc", 1, (192, 6, 3), (4, 3, 2)),  # 2-row groups of 96 pairs over 3 waves; tile 2 holds rows 4, 5
    ("a15", 5, (82, 7, 2), (6, 2, 2)),   # 3-row groups of 41 pairs over 2 waves, no XCD remap
    ("z", 3, (68, 16, 2), (14, 3, 2)),   # 7-row groups of 34 pairs over 4 waves
    # slab layout, the tile as the group: one run of pairs from a 128-byte boundary
    ("a15", 5, (82, 13, 2), (8, 1, 2)),  # 8 rows of 41 pairs over 6 waves; tile 2 holds 5 rows
    ("c15", 3, (57, 11, 2), (8, 3, 2)),  # 8 rows of 29 pairs over 4 waves, v inside; tile 2 holds 3 rows
    ("z", 5, (68, 10, 3), (12, 1, 2)),   # one tile of all 10 rows
    ("cube", 1, (90, 20, 3), (24, 0, 2)),  # one tile of 20 rows, no XCD remap
    ("fcc", 5, (45, 17, 3), (16, 1, 2)),  # degree 8, 16 rows of 23 pairs; tile 2 holds 1 row
    ("a15", 1, (82, 41, 2), (41, 0, 2)),  # the whole plane: 41 rows of 41 pairs over 27 waves
])
def test_packed_rows_match_oracle(oracle, lattice_name, ncomp, dims, traversal):
    """x pairs spread over groups of 2-8 y rows or over whole tiles (lines that one row per wave
    leaves partly idle)."""
    cfg = _config(lattice_name, ncomp, gpe.FP32, dims)
    (g0, g1), (o0, o1) = _run_both(oracle, cfg, 10, field=FIELD if ncomp > 1 else None, traversal=traversal)
    assert rel_l2(g0, o0) < 2e-5
    assert rel_l2(g1, o1) < 2e-5


def test_traversal_setter_validates():
    cfg = _config("fcc", 5, gpe.FP64, (4, 4, 4))
    with gpe.Stepper(cfg) as st:
        with pytest.raises(gpe.GpeError, match="EINVAL"):
            st.set_traversal(8, 1, 2)  # FP64 has one x per lane only
        with pytest.raises(gpe.GpeError, match="EINVAL"):
            st.set_traversal(0, 1, 1)
        with pytest.raises(gpe.GpeError, match="EINVAL"):
            st.set_traversal(8, 1, 4)  # one or two x per lane only
        with pytest.raises(gpe.GpeError, match="EINVAL"):
            st.set_traversal(8, 4, 1)  # order flags 0..3
        st.set_traversal(4, 0, 1)
        assert st.traversal() == dict(tile_rows=4, xcd_remap=0, x_per_lane=1)
        st.set_traversal(4, 2, 1)
        assert st.traversal() == dict(tile_rows=4, xcd_remap=2, x_per_lane=1)


def test_loss_term(oracle):
    cfg = _config("bcc", 3, gpe.FP64, (4, 3, 3), alpha=5.0)
    (g0, _), (o0, _) = _run_both(oracle, cfg, 10, field=FIELD)
    assert rel_l2(g0, o0) < 1e-12


def test_staggered_scalar_with_phase_start(oracle):
    """Scalar staggered scheme started with odd = exp(i mu dt/2) even (GPBlocks.cpp:516-523)."""
    from spacetime_amd import lattice
    cfg = _config("bcc", 1, gpe.FP64, (5, 5, 4), dt=1e-3, scheme=gpe.STAGGERED, c1=0.0, c2=0.0, q=0.0)
    lat = lattice.get("bcc")
    L = oracle.lattice(lat)
    p = oracle_params(oracle, cfg)
    even = random_state(cfg.shape, 5)
    mu = 3.7
    odd = np.exp(1j * mu * cfg.dt / 2) * even
    e, o = even.copy(), odd.copy()
    oracle.staggered(L, p, e, o, cfg.dt, 15)
    with gpe.Stepper(cfg) as st:
        st.set_state(even)
        st.start(gpe.START_PHASE, mu)
        st.step(15)
        g_even, g_odd = st.get_state(0), st.get_state(1)
        t, sweeps = st.time()
    assert sweeps == 30 and t == pytest.approx(15 * cfg.dt)
    assert rel_l2(g_even, e) < 1e-12
    assert rel_l2(g_odd, o) < 1e-12


@pytest.mark.parametrize("ncomp", [1, 3, 5])
def test_itp_matches_oracle(oracle, ncomp):
    from spacetime_amd import lattice
    cfg = _config("bcc", ncomp, gpe.FP64, (4, 4, 3))
    L = oracle.lattice(lattice.get("bcc"))
    p = oracle_params(oracle, cfg)
    psi = random_state(cfg.shape, 8)
    ref = psi.copy()
    oracle.normalize(L, p, ref)
    for _ in range(25):
        oracle.itp_step(L, p, ref, 1e-3)
    with gpe.Stepper(cfg) as st:
        st.set_state(psi)
        st.itp(25, 1e-3)
        got = st.get_state(0)
        obs = st.observables(0)
    assert obs.norm == pytest.approx(1.0, rel=1e-12)
    assert rel_l2(got, ref) < 1e-12


@pytest.mark.parametrize("dims,traversal,ncomp", [((4, 4, 3), None, 3), ((6, 7, 3), (4, 1, 2), 3),
                                                  ((6, 7, 3), (4, 1, 2), 1)])
def test_fp32_itp_matches_oracle(oracle, dims, traversal, ncomp):
    """FP32 imaginary time (FP64 norm and renormalisation): the default traversal and x pairs over
    4 rows per wave with rows past ny, three components and one."""
    from spacetime_amd import lattice
    cfg = _config("bcc", ncomp, gpe.FP32, dims)
    L = oracle.lattice(lattice.get("bcc"))
    p = oracle_params(oracle, cfg)
    psi = random_state(cfg.shape, 9)
    ref = psi.copy()
    oracle.normalize(L, p, ref)
    for _ in range(25):
        oracle.itp_step(L, p, ref, 1e-3)
    with gpe.Stepper(cfg) as st:
        if traversal:
            st.set_traversal(*traversal)
        st.set_state(psi)
        st.itp(25, 1e-3)
        got = st.get_state(0)
        obs = st.observables(0)
    assert obs.norm == pytest.approx(1.0, rel=1e-6)
    assert rel_l2(got, ref) < 5e-5


@pytest.mark.parametrize("lattice_name,ncomp,precision", [
    ("bcc", 1, 1), ("bcc", 3, 1), ("bcc", 5, 1), ("fcc", 3, 1), ("a15", 5, 1), ("bcc", 3, 0)])
def test_observables_match_oracle(oracle, lattice_name, ncomp, precision):
    from spacetime_amd import lattice
    cfg = _config(lattice_name, ncomp, precision, (4, 3, 3))
    L = oracle.lattice(lattice.get(lattice_name))
    p = with_field(oracle_params(oracle, cfg), **FIELD)
    psi = random_state(cfg.shape, 12)
    with gpe.Stepper(cfg) as st:
        st.set_field(**FIELD)
        st.set_state(psi)
        o = st.observables(0)
        stored = st.get_state(0)
    tol = 1e-12 if precision == 1 else 1e-6
    # FP32 storage rounds the state; compare against the oracle on the stored values
    assert o.norm == pytest.approx(oracle.norm(L, p, stored), rel=tol)
    assert o.energy == pytest.approx(oracle.energy(L, p, stored), rel=tol)
    assert o.expect_h == pytest.approx(oracle.expect_h(L, p, stored), rel=tol)
    m = oracle.magnetization(L, p, stored)
    np.testing.assert_allclose(o.magnetization, m, rtol=tol, atol=tol * abs(o.norm))
    assert o.max_density == pytest.approx(np.max(np.sum(np.abs(stored) ** 2, axis=0)), rel=1e-12)
    assert o.finite
    if precision == 0:
        assert rel_l2(stored, psi) < 1e-7


@pytest.mark.parametrize("lattice_name,ncomp,precision", [
    ("bcc", 1, 1), ("bcc", 3, 1), ("fcc", 3, 1), ("a15", 5, 1), ("bcc", 3, 0), ("c15", 1, 0)])
def test_max_hamiltonian_and_weighted_diff_match_oracle(oracle, lattice_name, ncomp, precision):
    """gpe_max_hamiltonian (maxHamilton, Spin1GpeBCC.cu:144-189) and gpe_weighted_diff
    (weightedDiff + integrate, :436-461,1620-1633) against the oracle on the stored state, with
    a quadrupole field and a complex phase; the device-pointer form gives the same bits.  The
    device copy of the reference is made with the HIP runtime libgpe itself links (torch ships
    its own, which cannot share the device with it once libgpe has initialised HIP)."""
    import ctypes
    from spacetime_amd import lattice
    cfg = _config(lattice_name, ncomp, precision, (5, 3, 4))
    L = oracle.lattice(lattice.get(lattice_name))
    p = with_field(oracle_params(oracle, cfg), **FIELD)
    psi = random_state(cfg.shape, 21)
    ref = random_state(cfg.shape, 22)
    ph = np.exp(-0.4j)
    with gpe.Stepper(cfg) as st:
        st.set_field(**FIELD)
        st.set_state(psi)
        stored = st.get_state(0)
        if ncomp == 5:
            with pytest.raises(gpe.GpeError):
                st.max_hamiltonian()
        else:
            got = st.max_hamiltonian()
            assert got == pytest.approx(oracle.max_hamiltonian(L, p, stored), rel=1e-13)
        wd = st.weighted_diff(ref, ph)
        hip = gpe.load_library()  # dlsym through libgpe: its own HIP runtime
        hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        hip.hipFree.argtypes = [ctypes.c_void_p]
        host = np.ascontiguousarray(ref, dtype=np.complex128)
        dref = ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(dref), host.nbytes) == 0
        try:
            assert hip.hipMemcpy(dref, host.ctypes.data, host.nbytes, 1) == 0  # host to device
            wd_dev = st.weighted_diff(dref.value, ph)
        finally:
            hip.hipFree(dref)
        assert st.weighted_diff(stored, 1.0) == 0.0
    assert wd == wd_dev
    assert wd == pytest.approx(oracle.weighted_diff(L, p, stored, ref, ph), rel=1e-12)


def test_ramp_matches_oracle(oracle):
    """16-knot piecewise-linear Bz ramp plus a constant quadrupole gradient."""
    from spacetime_amd import lattice
    cfg = _config("bcc", 3, gpe.FP64, (4, 4, 3), q=0.1)
    rng = np.random.default_rng(16)
    knots = rng.uniform(-0.1, 0.1, 16)
    T = 20 * cfg.dt
    b_segments = [((0, 0, knots[0]), 0.0, 0)] + [((0, 0, k), T / 15, 1) for k in knots[1:]]
    q_segments = [(4e-3, 1e9, 0)]
    L = oracle.lattice(lattice.get("bcc"))
    p = oracle_params(oracle, cfg)
    r = oracle.ramp(q_segments=q_segments, b_segments=b_segments)
    psi0 = random_state(cfg.shape, 3)
    a = psi0.copy()
    # forward Euler at t = 0 with the ramp field at t = 0
    bq0, bb0 = oracle.ramp_signal(r, 0.0)
    with_field(p, bq=bq0, bb=bb0)
    b = oracle.forward_euler(L, p, a, cfg.dt)
    oracle.leapfrog3(L, p, a, b, cfg.dt, 22, t0=cfg.dt, ramp=r)
    with gpe.Stepper(cfg) as st:
        st.set_ramp(q_segments=q_segments, b_segments=b_segments)
        st.set_state(psi0)
        st.start(gpe.START_EULER)
        st.step(22)
        g0 = st.get_state(0)
    assert rel_l2(g0, b) < 1e-12


def test_fp32_tracks_fp64_over_many_steps():
    """The FP32 form stays within 1e-4 relative L2 of the FP64 form (small grid)."""
    from spacetime_amd import lattice  # noqa: F401
    cfgs = [_config("bcc", 3, prec, (16, 16, 16), dt=1e-4, cell=0.5, c0=1000.0, c1=50.0, c2=0.0) for prec in (0, 1)]
    psi0 = normalized(random_state(cfgs[0].shape, 4), cfgs[0])
    out = []
    for cfg in cfgs:
        with gpe.Stepper(cfg) as st:
            st.set_field(bq=4e-3, bb=(0, 0, 0.05))
            st.set_state(psi0)
            st.start(gpe.START_EULER)
            st.step(2000)
            out.append(st.get_state(0))
    assert rel_l2(out[0], out[1]) < 1e-4


def test_errors_are_reported():
    cfg = _config("bcc", 3, gpe.FP64, (4, 4, 4))
    with gpe.Stepper(cfg) as st:
        with pytest.raises(gpe.GpeError, match="ESTATE"):
            st.step(3)  # no start yet
        with pytest.raises(ValueError):
            st.set_state(np.zeros((3, 1, 1, 1, 12)))
        bad = np.zeros(cfg.shape, dtype=np.complex128)
        bad[0, 0, 0, 0, 0] = np.nan
        st.set_state(bad)
        o = st.observables(0, check=False)
        assert not o.finite
    with pytest.raises(gpe.GpeError, match="EINVAL"):
        gpe.Stepper(gpe.Config(lattice="bcc", dims=(0, 4, 4)))


def _hash_uniform(seed, idx, s):
    """Host restatement of the fill's counter-based variate (splitmix64 finaliser)."""
    M = np.uint64(0xFFFFFFFFFFFFFFFF)
    with np.errstate(over="ignore"):
        z = (np.uint64(seed) * np.uint64(0x9E3779B97F4A7C15) + idx.astype(np.uint64) * np.uint64(0xD1B54A32D192ED03)
             + np.uint64(s) * np.uint64(0x94D049BB133111EB)) & M
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) / 2.0 ** 53


@pytest.mark.parametrize("lattice_name,ncomp,precision", [("bcc", 5, 1), ("c15", 3, 1), ("fcc", 5, 0), ("z", 3, 1)])
def test_device_thomas_fermi_fill(lattice_name, ncomp, precision):
    from spacetime_amd import states
    cfg = _config(lattice_name, ncomp, precision, (6, 5, 4), cell=0.5)
    r = states.positions(cfg)
    V = states.trap_potential(cfg, r)
    g, mu = 40.0, float(np.median(V))
    tf = np.sqrt(np.maximum(mu - V, 0.0) / g)
    w = np.exp(1j * np.arange(ncomp)) / np.sqrt(ncomp)
    nz, ny, nx, nv = V.shape
    node = np.arange(nz * ny * nx * nv).reshape(nz, ny, nx, nv)
    expect = np.empty(cfg.shape, dtype=np.complex128)
    amp, seed = 0.01, 54363
    for c in range(ncomp):
        rad = np.sqrt(_hash_uniform(seed, node, 2 * c))
        eps = rad * np.exp(2j * np.pi * _hash_uniform(seed, node, 2 * c + 1))
        expect[c] = tf * (w[c] + amp * eps)
    tol = 1e-14 if precision == 1 else 1e-6
    with gpe.Stepper(cfg) as st:
        st.fill_thomas_fermi(g, mu, w, noise=0.0, t=0.25)
        clean = st.get_state(0)
        assert st.time()[0] == pytest.approx(0.25)
        st.fill_thomas_fermi(g, mu, w, noise=amp, seed=seed)
        noisy = st.get_state(0)
    assert rel_l2(clean, w[:, None, None, None, None] * tf[None]) < tol
    assert rel_l2(noisy, expect) < tol
    assert 0.0 < rel_l2(noisy, clean) < amp * np.sqrt(ncomp)


@pytest.mark.parametrize("precision,tol", [(gpe.FP64, 1e-12), (gpe.FP32, 2e-5)])
def test_expanding_grid_matches_oracle(oracle, precision, tol):
    """Grid expansion after trap-off: the block scale of the Laplacian and positions is
    s (1 + rate max(0, t - start)); the oracle is stepped sweep by sweep with that scale."""
    from spacetime_amd import lattice
    dt = 2e-4
    cfg = _config("bcc", 3, precision, (6, 4, 3), dt=dt, expansion_start=3.5 * dt, expansion_rate=60.0,
                  trap_off_time=3.5 * dt)
    L = oracle.lattice(lattice.get("bcc"))
    psi0 = random_state(cfg.shape, 8)
    s0 = cfg.block_scale
    p = with_field(oracle_params(oracle, cfg), **FIELD)

    def at(t):
        p.block_scale = s0 * (1.0 + cfg.expansion_rate * max(0.0, t - cfg.expansion_start))
        return p

    n = 14
    a = psi0.copy()
    b = oracle.forward_euler(L, at(0.0), a, dt)
    for k in range(1, n + 1):
        nxt = a.copy()
        oracle.update(L, at(k * dt), nxt, b, 2 * dt, t=k * dt)
        a, b = b, nxt
    with gpe.Stepper(cfg) as st:
        st.set_field(**FIELD)
        st.set_state(psi0)
        st.start(gpe.START_EULER)
        st.step(n)
        got = st.get_state(0)
        with pytest.raises(gpe.GpeError, match="EINVAL"):
            st.set_state(psi0)
            st.adjoint(4, 2, gpe.OBJ_MAGNETIZATION_Z)
    assert p.block_scale > 1.05 * s0
    assert rel_l2(got, b) < tol


@pytest.mark.parametrize("xpl,nx", [(1, 7), (2, 7), (2, 66), (1, 66), (2, 300)])
def test_fp32_staggered_matches_oracle(oracle, xpl, nx):
    """The FP32 form of the scalar staggered scheme (odd += -i dt H(even); even += -i dt H(odd),
    GPBlocks.cpp:560-606) with one and two x per lane, odd and even line lengths, phase start;
    the 300-block box (two waves per line) has a weaker x trap so that dt V stays below the
    scheme's limit."""
    from spacetime_amd import lattice
    trap = (1.0 if nx <= 100 else 30.0 / (0.6 * nx), 0.9, 1.2)
    cfg = _config("bcc", 1, gpe.FP32, (nx, 5, 4), dt=1e-3, scheme=gpe.STAGGERED, c1=0.0, c2=0.0, q=0.0, trap=trap)
    L = oracle.lattice(lattice.get("bcc"))
    p = oracle_params(oracle, cfg)
    even = random_state(cfg.shape, 6)
    mu = 2.9
    e, o = even.copy(), np.exp(1j * mu * cfg.dt / 2) * even
    oracle.staggered(L, p, e, o, cfg.dt, 15)
    with gpe.Stepper(cfg) as st:
        st.set_traversal(4, 1, xpl)
        st.set_state(even)
        st.start(gpe.START_PHASE, mu)
        st.step(15)
        g_even, g_odd = st.get_state(0), st.get_state(1)
    assert rel_l2(g_even, e) < 2e-5
    assert rel_l2(g_odd, o) < 2e-5


def _golden_cases():
    import json
    import os
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "gpblocks_scalar_bcc.json")
    with open(path) as f:
        return json.load(f)["cases"]


@pytest.mark.parametrize("case", [0, 1])
def test_staggered_run_matches_reference_golden(oracle, case):
    """The HIP FP64 staggered scheme against the compiled reference's own output: the norm and
    overlap error per iteration printed by GrossPitaevskii/GPBlocks.cpp integrateInTime
    (:528-538, loop :560-606) for a stationary vortex state.  Box sizing (:456-460) and the
    first odd level e^{i mu dt/2} even (:516-523) follow the reference; the vortex profile is
    relaxed by the library (gpe_vortex_relax) and sampled at GPBlocks' own positions (the
    reference interpolant is discontinuous in its last z row, so positions one ulp apart can
    sample different rows); dt and the steps per iteration are the reference's rule as restated
    by the oracle."""
    from spacetime_amd import lattice, states
    c = _golden_cases()[case]
    lat = lattice.get("bcc")
    w = lat.block_dims
    s = c["block_scale"]
    minp = np.array([-c["rmax"], -c["rmax"], -c["zmax"]])
    maxp = -minp
    dims = tuple(int((maxp[a] - minp[a]) / (s * w[a])) + 1 for a in range(3))
    p0 = tuple(0.5 * (minp[a] + maxp[a] - s * w[a] * dims[a]) for a in range(3))
    # the reference's dt rule (GPBlocks.cpp:497-511), through the oracle's restatement
    vs = oracle.vortex(c["kappa"], c["g"], c["rmax"], c["zmax"], c["rstep"], c["zstep"], maxiters=c["maxiters"])
    gl = __import__("copy").copy(lat)
    gl.pos = np.array(c["bcc_pos"])
    _, _, dt, steps = oracle.gpblocks_integrate(oracle.lattice(gl), vs, c["kappa"], c["g"], s, list(minp),
                                                list(maxp), c["period"], 0)
    # the library's own relaxation of the same profile, interpolated by the reference's rule
    lib_vs = states.vortex_state(c["kappa"], c["g"], c["rmax"], c["zmax"], c["rstep"], c["zstep"],
                                 maxiters=c["maxiters"])
    assert lib_vs["mu"] == pytest.approx(c["vortex_mu"], rel=1e-12)
    vs = dict(vs, f=lib_vs["f"])
    # psi0 at GPBlocks' positions: p0 + s (w i + pos_v)
    nz, ny, nx = dims[2], dims[1], dims[0]
    iz, iy, ix = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    blk = np.stack([ix, iy, iz], -1)[:, :, :, None, :] * w[None, None, None, None, :]
    r = np.array(p0)[None, None, None, None, :] + s * (blk + gl.pos[None, None, None, :, :])
    # sqrt(x^2 + y^2) as VortexState::getPsi forms it; np.hypot rounds differently and the
    # interpolant's discontinuity turns one ulp of r into a different sample
    rr = np.sqrt(r[..., 0] * r[..., 0] + r[..., 1] * r[..., 1])
    fv = np.vectorize(lambda a, b: oracle.vortex_function(vs, a, b))(rr, r[..., 2])
    with np.errstate(invalid="ignore", divide="ignore"):
        ph = ((r[..., 0] + 1j * r[..., 1]) / rr) ** c["kappa"]
    psi0 = np.where(rr < 1e-30, 0.0, fv * ph)[None].astype(np.complex128)
    cfg = gpe.Config(lattice="bcc", ncomp=1, precision=gpe.FP64, scheme=gpe.STAGGERED, dims=dims, block_scale=s,
                     p0=p0, dt=dt, c0=c["g"])
    vol = s ** 3 * lat.vol[0]
    normsq, error = [], []
    with gpe.Stepper(cfg) as st:
        st.set_state(psi0)
        st.start(gpe.START_PHASE, c["vortex_mu"])
        for it in range(c["iterations"] + 1):
            even = st.get_state(0)
            n = float(np.sum(np.abs(even) ** 2) * vol)
            ov = complex(np.sum(vol * np.conj(psi0) * even))
            normsq.append(n)
            error.append(n - abs(ov))
            if it < c["iterations"]:
                st.step(steps)
    np.testing.assert_allclose(normsq, c["normsq"], rtol=1e-12)
    np.testing.assert_allclose(error, c["error"], rtol=0, atol=1e-12)
    assert max(abs(e) for e in c["error"][1:]) > 1e-5
