"""Batched ensemble stepping (one launch per sweep for all members, eager and hipGraph) and
the traversal autotuner: results must equal each member stepped on its own (same kernel body
and arithmetic; FP32 tolerance 1e-6 allows for different instruction selection), and the oracle for a small member."""
import numpy as np
import pytest

from helpers import normalized, oracle_params, random_state, rel_l2, with_field

pytestmark = pytest.mark.gpu

gpe = pytest.importorskip("spacetime_amd.gpe")


def _members(precision=gpe.FP32, ncomp=5, large=False):
    from spacetime_amd import states
    specs = [("bcc", (9, 7, 5)), ("cube", (70, 6, 4)), ("fcc", (5, 5, 4)), ("a15", (3, 4, 3)), ("c15", (2, 2, 3)),
             ("bcc", (66, 3, 2))]
    if large:  # 4224 workgroups: launched on its own between the batched groups
        specs.insert(1, ("bcc", (128, 64, 22)))
    out = []
    rng = np.random.default_rng(0)
    for i, (lat, dims) in enumerate(specs):
        cfg = gpe.Config.centred_box(lat, dims, 0.7, ncomp=ncomp, precision=precision, dt=2e-4, c0=100.0,
                                     c1=float(rng.choice([-1, 1])) * 20.0, c2=float(rng.choice([-1, 1])) * 5.0)
        psi = normalized(random_state(cfg.shape, 10 + i), cfg)
        knots = rng.uniform(-0.2, 0.2, 8)
        out.append((cfg, psi, states.knot_ramp(knots, 40 * cfg.dt)))
    return out


def _run_single(cfg, psi, ramp, n):
    with gpe.Stepper(cfg) as st:
        st.set_ramp(q_segments=[(0.01, 1e30, 0)], b_segments=ramp)
        st.set_state(psi)
        st.start(gpe.START_EULER)
        st.step(n)
        return st.get_state(0), st.get_state(1)


@pytest.mark.parametrize("graph", [0, 4])
def test_ensemble_equals_individual_runs(graph):
    mem = _members(large=True)
    n = 12
    singles = [_run_single(cfg, psi, ramp, n) for cfg, psi, ramp in mem]
    steppers = []
    for cfg, psi, ramp in mem:
        st = gpe.Stepper(cfg)
        st.set_ramp(q_segments=[(0.01, 1e30, 0)], b_segments=ramp)
        st.set_state(psi)
        st.start(gpe.START_EULER)
        steppers.append(st)
    with gpe.Ensemble(steppers) as ens:
        ens.step(4, graph_sweeps=graph)
        ens.step(8, graph_sweeps=graph)
        ms, unknowns = ens.timing()
    assert ms > 0 and unknowns == sum(np.prod(c.dims) * c.V for c, _, _ in mem)
    for st, (g0, g1) in zip(steppers, singles):
        t, sweeps = st.time()
        assert sweeps == n + 1
        assert rel_l2(st.get_state(0), g0) < 1e-6
        assert rel_l2(st.get_state(1), g1) < 1e-6
        st.close()


def test_ensemble_member_matches_oracle(oracle):
    from spacetime_amd import lattice
    cfg = gpe.Config.centred_box("bcc", (6, 5, 4), 0.7, ncomp=3, precision=gpe.FP64, dt=2e-4, c0=100.0, c1=15.0)
    other = gpe.Config.centred_box("fcc", (4, 4, 4), 0.7, ncomp=3, precision=gpe.FP64, dt=1e-4, c0=50.0)
    psi = normalized(random_state(cfg.shape, 1), cfg)
    psi2 = normalized(random_state(other.shape, 2), other)
    a, b = gpe.Stepper(cfg), gpe.Stepper(other)
    a.set_field(bq=0.02, bb=(0, 0, 0.1))
    a.set_state(psi)
    a.start(gpe.START_EULER)
    b.set_state(psi2)
    b.start(gpe.START_EULER)
    with gpe.Ensemble([a, b]) as ens:
        ens.step(10, graph_sweeps=2)
    got = a.get_state(0)
    L = oracle.lattice(lattice.get("bcc"))
    p = with_field(oracle_params(oracle, cfg), bq=0.02, bb=(0, 0, 0.1))
    x = psi.copy()
    y = oracle.forward_euler(L, p, x, cfg.dt)
    oracle.leapfrog3(L, p, x, y, cfg.dt, 10, t0=cfg.dt)
    assert rel_l2(got, y) < 1e-12
    a.close()
    b.close()


def test_ensemble_rejects_mixed_members():
    c1 = gpe.Config.centred_box("bcc", (4, 4, 4), 0.7, ncomp=3)
    c2 = gpe.Config.centred_box("bcc", (4, 4, 4), 0.7, ncomp=5)
    s1, s2 = gpe.Stepper(c1), gpe.Stepper(c2)
    for s, c in ((s1, c1), (s2, c2)):
        s.set_state(np.zeros(c.shape, complex))
        s.start(gpe.START_COPY)
    with pytest.raises(gpe.GpeError, match="EINVAL"):
        gpe.Ensemble([s1, s2])
    s3 = gpe.Stepper(c1)
    with pytest.raises(gpe.GpeError, match="ESTATE"):
        gpe.Ensemble([s1, s3])  # s3 not started
    for s in (s1, s2, s3):
        s.close()


def test_autotune_keeps_results_identical():
    cfg = gpe.Config.centred_box("bcc", (40, 24, 10), 0.5, ncomp=3, precision=gpe.FP32, dt=1e-4, c0=500.0, c1=20.0)
    psi = normalized(random_state(cfg.shape, 3), cfg)
    ref = _run_single(cfg, psi, [((0, 0, 0.05), 0.0, 0)], 20)
    with gpe.Stepper(cfg) as st:
        st.set_ramp(q_segments=[(0.01, 1e30, 0)], b_segments=[((0, 0, 0.05), 0.0, 0)])
        st.set_state(psi)
        st.start(gpe.START_EULER)
        t = st.autotune(5)
        # the autotuner's tile heights: powers of two and multiples of 3, 5, 6 and 7 up to 32 (the plane, 24 rows, is one of them)
        assert t["tile_rows"] in (1, 2, 3, 4, 5, 6, 7, 8, 10, 12, 14, 16, 20, 24, 28, 32)
        assert t["xcd_remap"] in (0, 1, 2, 3) and t["ms_per_sweep"] > 0  # order flags
        assert t["x_per_lane"] in (1, 2)
        st.step(20)
        got = st.get_state(0)
        t2 = st.autotune(0)  # cached
        assert (t2["tile_rows"], t2["xcd_remap"], t2["x_per_lane"]) == (t["tile_rows"], t["xcd_remap"],
                                                                          t["x_per_lane"])
    assert rel_l2(got, ref[0]) < 1e-6


def test_graph_replays_each_match_eager():
    """The captured K-sweep graph replayed several times, with eager work (state reads,
    observables) and synchronisations between replays: every replay's result equals the same
    member stepped eagerly on its own."""
    mem = _members()[:3]
    steppers, singles = [], []
    for cfg, psi, ramp in mem:
        st = gpe.Stepper(cfg)
        st.set_ramp(q_segments=[(0.01, 1e30, 0)], b_segments=ramp)
        st.set_state(psi)
        st.start(gpe.START_EULER)
        steppers.append(st)
        ref = gpe.Stepper(cfg)
        ref.set_ramp(q_segments=[(0.01, 1e30, 0)], b_segments=ramp)
        ref.set_state(psi)
        ref.start(gpe.START_EULER)
        singles.append(ref)
    with gpe.Ensemble(steppers) as ens:
        for replay in range(4):
            ens.step(6, graph_sweeps=6)
            for st, ref in zip(steppers, singles):
                ref.step(6)
                st.observables(0, check=False)  # eager work between replays
                assert rel_l2(st.get_state(0), ref.get_state(0)) < 1e-6, replay
                assert st.time() == ref.time()
    for s in steppers + singles:
        s.close()


@pytest.mark.parametrize("graph", [0, 4])
def test_mixed_lattice_fp32_spin2_ensemble_matches_oracle(oracle, graph):
    """The configured ensemble's kind at test size: spin-2 FP32 members on the five lattices
    (Cube, BCC, FCC, A15, C15) with seeded signs of c1, c2, their own block counts, dt and 8-knot
    Bz ramps, stepped together; every member against the FP64 oracle stepped alone."""
    from spacetime_amd import lattice, states
    rng = np.random.default_rng(44)
    specs = [("cube", (41, 6, 5)), ("bcc", (7, 6, 5)), ("fcc", (6, 5, 4)), ("a15", (4, 3, 3)), ("c15", (3, 2, 2)),
             ("bcc", (70, 3, 2)), ("fcc", (33, 4, 3))]
    n = 12
    steppers, refs = [], []
    for i, (lat, dims) in enumerate(specs):
        signs = rng.choice([-1.0, 1.0], 2)
        dt = float(rng.choice([1e-4, 2e-4]))
        cfg = gpe.Config.centred_box(lat, dims, 0.7, ncomp=5, precision=gpe.FP32, dt=dt, c0=100.0,
                                     c1=20.0 * signs[0], c2=5.0 * signs[1])
        psi = normalized(random_state(cfg.shape, 30 + i), cfg)
        knots = rng.uniform(-0.2, 0.2, 8)
        b_segments = states.knot_ramp(knots, (n + 1) * dt)
        q_segments = [(0.01, 1e30, 0)]
        st = gpe.Stepper(cfg)
        st.set_ramp(q_segments=q_segments, b_segments=b_segments)
        st.set_state(psi)
        st.start(gpe.START_EULER)
        steppers.append(st)
        L = oracle.lattice(lattice.get(lat))
        p = oracle_params(oracle, cfg)
        r = oracle.ramp(q_segments=q_segments, b_segments=b_segments)
        bq0, bb0 = oracle.ramp_signal(r, 0.0)
        with_field(p, bq=bq0, bb=bb0)
        a = psi.copy()
        b = oracle.forward_euler(L, p, a, cfg.dt)
        oracle.leapfrog3(L, p, a, b, cfg.dt, n, t0=cfg.dt, ramp=r)
        refs.append((b, a))
    with gpe.Ensemble(steppers) as ens:
        ens.step(n, graph_sweeps=graph)
    for (lat, dims), st, (o0, o1) in zip(specs, steppers, refs):
        assert st.time()[1] == n + 1
        assert rel_l2(st.get_state(0), o0) < 2e-5, (lat, dims)
        assert rel_l2(st.get_state(1), o1) < 2e-5, (lat, dims)
        st.close()
