"""Throughput of the DEC GPE step on one MI355X, FP32 form vs FP64 form.

Workload (default, spin-1 BCC, the north-star case): 128^3 BCC blocks (12 unknowns each,
25,165,824 unknowns; 4,194,304 lattice points), box 24^3, trap (1, 1, 166/126), c0 = 1e4,
c1 = +500 (polar), constant quadratic Zeeman q = 0.1, field b'(x, y, -2z) + Bz(t) z with
b' = 4e-3 and Bz piecewise linear through 16 seeded knots U[-0.1, 0.1].  Initial state:
polar Thomas-Fermi profile relaxed by 1000 imaginary-time steps, plus seeded 0.5% noise.
Then a forward-Euler start, W warm-up steps and K timed 3-level leapfrog steps at dt = 1e-4,
in FP32 (FP64 reductions) and in FP64; the two final states are compared.

`--workload scalar` runs the scalar staggered case instead: 192^3 BCC blocks, box 9.6^3,
g = 300, vortex-ring imprint, dt = 1e-4.

`--workload adjoint` runs the discrete adjoint of the spin-1 physics on 64^3 BCC blocks: 4000
sweeps, checkpoints every 200, gradients of the overlap with a target state (made by a seeded
reference ramp) with respect to the 16 Bz knots and 3 trap frequencies, FP32 and FP64, checked
against central finite differences of FP64 forward runs (h = 1e-4).

`--workload ensemble` runs 48 independent spin-2 runs (c0 = 1e4, c1 = +-200, c2 = +-50 with
seeded signs) on mixed Cube/BCC/FCC/A15/C15 meshes of n^3 blocks, n in 32..96 (seeded), each
with its own 8-knot Bz ramp U[-0.2, 0.2] and noisy ferromagnetic Thomas-Fermi state, dt = 1e-4,
FP32 form: first one after another (each member stepped alone), then all together as one
batched launch per sweep replayed from a captured hipGraph, from the same initial states.

One step = one sweep over all unknowns (3-level) or one odd+even pair (staggered); a
node-update = one unknown advanced by one sweep.  Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "node-updates/s on 1 GPU, FP32-form vs FP64 at ≤1e-4 rel L2; % HBM roofline"
PEAK_HBM_GBS = 8000.0

# HBM bytes per launch of the FP32 sweep kernel (traversal 16 rows, XCD remap, two x per lane),
# measured with rocprofv3 --pmc FETCH_SIZE and --pmc WRITE_SIZE on each workload
# (profiles/k_sweep_spin1_bcc_128.md, profiles/k_sweep_scalar_bcc_192.md); FETCH_SIZE is
# scaled by 2.000 and WRITE_SIZE by 1.000, the ratios of bytes moved to counter readings for
# the 1 GiB streaming copy in the same runs.
PMC_TRAFFIC = {
    "spin1_bcc_128^3": 597938.6 * 1024 * 2.000 + 589824.4 * 1024 * 1.000,
    "scalar_bcc_192^3": 673380.3 * 1024 * 2.000 + 663552.0 * 1024 * 1.000,
}


_TORCH = None


def _init_torch():
    """Import torch and create its device context once, at start-up: a first import inside the
    measurement (1-2 s, minutes on a fresh machine) would idle the GPU right before the timed
    sweeps."""
    global _TORCH
    try:
        import torch
        if torch.cuda.is_available():
            torch.cuda.synchronize()
            _TORCH = torch
    except Exception:
        _TORCH = None


def _sync():
    if _TORCH is not None:
        _TORCH.cuda.synchronize()


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def copy_bandwidth(nbytes=1 << 30, reps=20):
    """Achievable HBM rates on this device from libgpe's streaming kernels (16-byte
    non-temporal accesses over two 1 GiB buffers): a copy, and the in-place read-2-write-1
    update that is the sweep's own HBM pattern.  GB/s, or (None, None) without a device."""
    try:
        from spacetime_amd import gpe
        return gpe.stream_bandwidth(nbytes, reps)
    except Exception:
        return None, None


def workload_spin1(seed=2024):
    from spacetime_amd import gpe
    cells = (128, 128, 128)
    box = 24.0
    cfg = gpe.Config.centred_box("bcc", cells, box / cells[0], ncomp=3, dt=1e-4, trap=(1.0, 1.0, 166.0 / 126.0),
                                 c0=1.0e4, c1=500.0, q=0.1)
    rng = np.random.default_rng(seed)
    knots = rng.uniform(-0.1, 0.1, 16)
    info = dict(workload="spin1_bcc_128^3", lattice="bcc", cells=list(cells), box=box, components=3,
                scheme="leapfrog3", dt=cfg.dt, c0=cfg.c0, c1=cfg.c1, q=cfg.q, bq=4e-3, bz_knots=16, seed=seed)
    return cfg, knots, info


def workload_scalar():
    from spacetime_amd import gpe
    cells = (192, 192, 192)
    box = 9.6
    cfg = gpe.Config.centred_box("bcc", cells, box / cells[0], ncomp=1, dt=1e-4, scheme=gpe.STAGGERED, c0=300.0)
    info = dict(workload="scalar_bcc_192^3", lattice="bcc", cells=list(cells), box=box, components=1,
                scheme="staggered", dt=cfg.dt, g=cfg.c0, imprint="vortex ring R=2")
    return cfg, None, info


def ensemble_members(seed=4848, n=48, box=16.0, blocks=(32, 96)):
    """Seeded member list: (config, knots, seed) with c1, c2 signs, lattice and blocks per axis drawn."""
    from spacetime_amd import gpe, lattice
    rng = np.random.default_rng(seed)
    names = rng.choice(["cube", "bcc", "fcc", "a15", "c15"], n)  # the workload's five lattices
    dims = rng.integers(blocks[0], blocks[1] + 1, size=n)
    signs = rng.choice([-1.0, 1.0], size=(n, 2))
    knots = rng.uniform(-0.2, 0.2, size=(n, 8))
    out = []
    for i in range(n):
        d = (int(dims[i]),) * 3
        # box^3 with n^3 blocks; the finest block edge, box / 96, keeps dt * ||H|| < 1 (the
        # 3-level leapfrog's stability bound) for every lattice at dt = 1e-4
        cfg = gpe.Config.centred_box(str(names[i]), d, box / d[0], ncomp=5, precision=gpe.FP32, dt=1e-4, c0=1.0e4,
                                     c1=200.0 * signs[i, 0], c2=50.0 * signs[i, 1])
        out.append((cfg, knots[i], seed + 1 + i))
    return out


def measure_ensemble(members, args, log, label):
    """The members stepped one after another (each alone, eager launches) and together (one
    launch per kernel group per sweep, replayed from a captured hipGraph), from the same states."""
    from spacetime_amd import gpe
    duration = 5000 * 1e-4
    steppers = []
    nodes = 0.0
    for cfg, knots, _ in members:
        st = gpe.Stepper(cfg)
        st.set_knot_ramp(knots, duration)
        steppers.append(st)
        nodes += float(np.prod(cfg.dims) * cfg.V)
    gmu = 1.0e4
    mu = 0.5 * (15.0 * gmu / (4.0 * np.pi)) ** 0.4  # Thomas-Fermi mu of a unit-norm isotropic trap
    ferro = np.array([1, 0, 0, 0, 0], dtype=np.complex128)

    def prepare():
        for (cfg, _, sd), st in zip(members, steppers):
            st.fill_thomas_fermi(gmu, mu, ferro, noise=0.005, seed=sd)
            st.scale(1.0 / np.sqrt(st.observables(0).norm))
            st.start(gpe.START_EULER)

    log("%s: %d members, %.3e unknowns in total (%.2f GB of FP32 levels)" % (
        label, len(members), nodes, nodes * 5 * 8 * 2 / 1e9))
    prepare()
    tune = []
    t0 = time.perf_counter()
    for st in steppers:
        tune.append(st.autotune(args.tune_sweeps))
    log("%s autotune: %d members in %.1fs" % (label, len(steppers), time.perf_counter() - t0))
    prepare()

    # one after another: each member alone, eager launches
    for st in steppers:
        st.step(args.warmup)
        st.sync()
    _sync()
    t0 = time.perf_counter()
    bytes_sweep = 0.0
    per_member = []
    for (cfg, _, _), st in zip(members, steppers):
        st.step(args.steps)
        st.sync()
        ms, nb = st.last_step_timing()
        bytes_sweep += nb
        per_member.append((cfg.lattice, cfg.dims[0], ms, nb))
    _sync()
    wall_seq = time.perf_counter() - t0
    obs_seq = [st.observables(0) for st in steppers]
    log("%s sequential: %d steps in %.3fs, %.3e node-updates/s" % (label, args.steps, wall_seq,
                                                                   nodes * args.steps / wall_seq))
    by_lattice = {}
    for lat, n, ms, nb in per_member:
        a = by_lattice.setdefault(lat, [0.0, 0.0])
        a[0] += ms
        a[1] += nb
    total_ms = sum(a[0] for a in by_lattice.values())
    for lat, (ms, nb) in sorted(by_lattice.items(), key=lambda kv: -kv[1][0]):
        log("%s sequential %-4s: %.3f ms per sweep (%.0f%% of the sum), %.0f GB/s" % (
            label, lat, ms, 100 * ms / total_ms, nb / (ms * 1e-3) / 1e9))

    prepare()
    G = args.graph_sweeps
    with gpe.Ensemble(steppers) as ens:
        ens.step(args.warmup, G if args.warmup % G == 0 else 0)
        _sync()
        t0 = time.perf_counter()
        ens.step(args.steps, G)
        _sync()
        wall_ens = time.perf_counter() - t0
        ms_sweep, _ = ens.timing()
    obs_ens = [st.observables(0) for st in steppers]
    rel_e = max(abs(a.energy - b.energy) / abs(b.energy) for a, b in zip(obs_ens, obs_seq))
    rel_n = max(abs(a.norm - b.norm) / b.norm for a, b in zip(obs_ens, obs_seq))
    drift = max(abs(o.norm - 1.0) for o in obs_ens)
    finite = all(o.finite for o in obs_ens)
    log("%s together: %d steps in %.3fs, %.4f ms/sweep, %.3e node-updates/s (%.3fx); max rel energy diff vs "
        "sequential %.2e" % (label, args.steps, wall_ens, ms_sweep, nodes * args.steps / wall_ens, wall_seq / wall_ens,
                              rel_e))
    for st in steppers:
        st.close()
    counts = {}
    for cfg, _, _ in members:
        counts[cfg.lattice] = counts.get(cfg.lattice, 0) + 1
    return dict(nodes=nodes, wall_seq=wall_seq, wall_ens=wall_ens, ms_sweep=ms_sweep, bytes_sweep=bytes_sweep,
                total_ms=total_ms, per_member=per_member, counts=counts, tune=tune,
                accuracy={"max_rel_energy_diff_vs_sequential": rel_e, "max_rel_norm_diff_vs_sequential": rel_n,
                          "max_norm_drift": drift, "finite": finite})


def run_ensemble(args, log):
    """Aggregate node-updates/s of the batched ensemble vs the same runs one after another: the
    configured 48 spin-2 members of 32^3-96^3 blocks, and 48 small members of 8^3-24^3 blocks,
    where a member alone cannot fill the GPU."""
    big = measure_ensemble(ensemble_members(), args, log, "ensemble")
    small = measure_ensemble(ensemble_members(seed=4849, blocks=(8, 24)), args, log, "small members")
    achieved = big["bytes_sweep"] / (big["ms_sweep"] * 1e-3) / 1e9
    nodes, G = big["nodes"], args.graph_sweeps
    return {
        "metric": "aggregate node-updates/s of a 48-run spin-2 ensemble (batched launches per sweep, hipGraph) "
                  "vs the runs one after another",
        "value": nodes * args.steps / big["wall_ens"],
        "unit": "node-updates/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * big["wall_ens"] / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
        "config": {"workload": "ensemble_spin2_48_mixed", "members": 48, "lattices": big["counts"],
                   "blocks_per_axis": [32, 96], "unknowns": nodes, "components": 5, "dt": 1e-4,
                   "graph_sweeps": G,
                   "autotuned": [[t["tile_rows"], t["xcd_remap"], t["x_per_lane"]] for t in big["tune"]]},
        "sequential": {"value": nodes * args.steps / big["wall_seq"], "ms_per_step": 1e3 * big["wall_seq"] / args.steps,
                       "kernel_ms_per_sweep_sum": big["total_ms"],
                       "members": [[lat, n, round(ms, 5), round(nb / (ms * 1e-3) / 1e9, 1)]
                                   for lat, n, ms, nb in big["per_member"]]},
        "speedup_vs_sequential": big["wall_seq"] / big["wall_ens"],
        "small_members": {"blocks_per_axis": [8, 24], "members": 48, "lattices": small["counts"],
                          "unknowns": small["nodes"],
                          "value": small["nodes"] * args.steps / small["wall_ens"],
                          "sequential_value": small["nodes"] * args.steps / small["wall_seq"],
                          "speedup_vs_sequential": small["wall_seq"] / small["wall_ens"],
                          "accuracy": small["accuracy"]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": achieved / PEAK_HBM_GBS, "traffic": None, "kernel_ms_per_sweep": big["ms_sweep"],
                     "algorithmic_bytes_per_sweep": big["bytes_sweep"]},
        "accuracy": big["accuracy"],
        "cpu_baseline": None,
    }


def run_lattices(args, log):
    """FP32 sweep rate of every lattice x component count at ~25M unknowns (the spin-1 BCC 128^3
    size): cube 293^3, BCC / FCC 128^3, A15 82^3, C15 57^3, Z 68^3 blocks; autotuned traversal.  Each case
    also runs the plain FP64 form from the same start (unless --no-fp64) and reports the relative
    L2 distance of the final states and both forms' norm and energy drift.  --cases picks cases and
    may give a case another box (block scale box / nx)."""
    from spacetime_amd import gpe, lattice
    import dataclasses
    box = 16.0
    gmu = 1.0e4
    mu = 0.5 * (15.0 * gmu / (4.0 * np.pi)) ** 0.4
    target = 25165824
    cases = []
    tot_nodes_steps, tot_s, tot_bytes, tot_kernel_s = 0.0, 0.0, 0.0, 0.0
    if args.cases:  # "lattice:C[:NXxNYxNZ],...": chosen cases, optionally on another box (profiling runs)
        plan = []
        for c in args.cases.split(","):
            parts = c.split(":")
            plan.append((parts[0], int(parts[1]), tuple(int(v) for v in parts[2].split("x")) if len(parts) > 2 else None))
    else:
        plan = [(name, ncomp, None) for name in ("cube", "bcc", "fcc", "a15", "c15", "z") for ncomp in (1, 3, 5)]
    for name, ncomp, dims in plan:
        lat = lattice.get(name)
        n = int(round((target / lat.V) ** (1.0 / 3.0)))
        dims = dims or (n, n, n)
        kw = dict(c0=gmu)
        if ncomp == 3:
            kw.update(c1=500.0, q=0.1)
        elif ncomp == 5:
            kw.update(c1=200.0, c2=50.0)
        cell = box / dims[0]
        probe = gpe.Config.centred_box(name, dims, cell, ncomp=ncomp, precision=gpe.FP32, dt=1e-4, **kw)
        # the 3-level leapfrog needs dt * ||H|| < 1
        dt = min(1e-4, 0.5 / (lat.kinetic_bound(probe.block_scale) + mu))
        cfg = gpe.Config.centred_box(name, dims, cell, ncomp=ncomp, precision=gpe.FP32, dt=dt, **kw)
        weights = np.zeros(ncomp, dtype=np.complex128)
        weights[ncomp // 2] = 1.0
        field = dict(bq=4e-3, bb=(0.0, 0.0, 0.05)) if ncomp > 1 else None
        cfg64 = dataclasses.replace(cfg, precision=gpe.FP64)
        with gpe.Stepper(cfg64) as st:  # the common start: FP64 Thomas-Fermi state with noise
            st.fill_thomas_fermi(gmu, mu, weights, noise=0.005, seed=17)
            st.scale(1.0 / np.sqrt(st.observables(0).norm))
            psi0 = st.get_state(0)

        def run(c, timed):
            with gpe.Stepper(c) as st:
                if field:
                    st.set_field(**field)
                st.set_state(psi0)
                o0 = st.observables(0)
                st.start(gpe.START_EULER)
                if timed and args.traversal:
                    st.set_traversal(*args.traversal)
                    tune = st.traversal()
                else:
                    tune = st.autotune(args.tune_sweeps) if (timed and args.tune_sweeps) else st.traversal()
                st.step(args.warmup)
                st.sync()
                t0 = time.perf_counter()
                st.step(args.steps)
                st.sync()
                wall = time.perf_counter() - t0
                ms, nb = st.last_step_timing()
                o1 = st.observables(0)
                psi = st.get_state(0)
            drift = dict(norm_drift=abs(o1.norm - o0.norm) / o0.norm,
                         energy_drift=abs(o1.energy - o0.energy) / abs(o0.energy))
            return tune, wall, ms, nb, o1, psi, drift

        tune, wall, ms, nb, o, psi32, drift32 = run(cfg, True)
        acc = dict(norm_drift_fp32=drift32["norm_drift"], energy_drift_fp32=drift32["energy_drift"])
        if not args.no_fp64:  # the plain FP64 form from the same start, same steps
            _, _, _, _, _, psi64, drift64 = run(cfg64, False)
            acc.update(rel_l2_fp32_vs_fp64=float(np.linalg.norm((psi32 - psi64).ravel()) /
                                                 np.linalg.norm(psi64.ravel())),
                       norm_drift_fp64=drift64["norm_drift"], energy_drift_fp64=drift64["energy_drift"])
            del psi64
        del psi32, psi0
        nodes = float(np.prod(dims) * lat.V)
        gbs = nb / (ms * 1e-3) / 1e9
        cases.append(dict(lattice=name, components=ncomp, blocks=list(dims), unknowns=nodes, dt=dt,
                          layout=gpe.level_layout(cfg)["layout"],
                          traversal=[tune["tile_rows"], tune["xcd_remap"], tune["x_per_lane"]],
                          kernel_ms_per_sweep=ms, node_updates_per_s=nodes * args.steps / wall,
                          achieved_gbs=gbs, frac=gbs / PEAK_HBM_GBS, norm=o.norm, finite=bool(o.finite),
                          accuracy=acc))
        log("%-4s C=%d %s blocks: %.4f ms/sweep, %.3e node-updates/s, %6.1f GB/s (%.1f%%), traversal %s%s" % (
            name, ncomp, "x".join(str(d) for d in dims), ms, nodes / (ms * 1e-3), gbs, 100 * gbs / PEAK_HBM_GBS,
            cases[-1]["traversal"],
            (", FP32 vs FP64 rel L2 %.2e" % acc["rel_l2_fp32_vs_fp64"]) if "rel_l2_fp32_vs_fp64" in acc else ""))
        tot_nodes_steps += nodes * args.steps
        tot_s += wall
        tot_bytes += nb * args.steps
        tot_kernel_s += ms * 1e-3 * args.steps
    achieved = tot_bytes / tot_kernel_s / 1e9
    return {
        "metric": "FP32 node-updates/s of every lattice x component count at ~25M unknowns",
        "value": tot_nodes_steps / tot_s,
        "unit": "node-updates/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * tot_s / (args.steps * len(cases)),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
        "config": {"workload": "lattices_25M", "box": box, "lattices": ["cube", "bcc", "fcc", "a15", "c15", "z"],
                   "components": [1, 3, 5]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": achieved / PEAK_HBM_GBS, "traffic": None},
        "cases": cases,
        "cpu_baseline": None,
    }


def run_adjoint(args, log):
    """Spin-1 BCC 64^3 discrete adjoint: gradients of |<target|psi(T)>|^2 with respect to the 16
    Bz knots and the 3 trap frequencies, checkpoints every 200 sweeps, in three forms: an FP32
    context (FP32 sweeps and states, FP64 accumulation), the same context with fp64_trajectory
    (trajectory and adjoint state in FP64), and an FP64 context; checked against central finite
    differences of FP64 forward runs (h = 1e-4 and h / 2, Richardson-combined).  The start state
    and the target are rounded to FP32 first, so every form and the finite differences solve the
    problem the FP32 context holds."""
    from spacetime_amd import gpe, states
    from spacetime_amd.adjoint import RampTrapProblem
    import dataclasses
    cells = (64, 64, 64)
    box = 24.0
    cfg = gpe.Config.centred_box("bcc", cells, box / cells[0], ncomp=3, precision=gpe.FP32, dt=1e-4,
                                 trap=(1.0, 1.0, 166.0 / 126.0), c0=1.0e4, c1=500.0, q=0.1)
    cfg64 = dataclasses.replace(cfg, precision=gpe.FP64)
    nsweeps, K = args.steps, 200
    duration = nsweeps * cfg.dt
    knots = np.random.default_rng(2024).uniform(-0.1, 0.1, 16)
    ref_knots = np.random.default_rng(4000).uniform(-0.1, 0.1, 16)
    fp32_round = lambda a: a.astype(np.complex64).astype(np.complex128)  # noqa: E731
    psi0 = fp32_round(prepare_state(cfg, knots, args.itp_steps, args.itp_dt, log))
    with gpe.Stepper(cfg64) as st:  # the target: the same start under the seeded reference ramp
        st.set_ramp(q_segments=[(4e-3, 1e30, 0)], b_segments=states.knot_ramp(ref_knots, duration))
        st.set_state(psi0)
        st.start(gpe.START_EULER)
        st.step(nsweeps - 1)
        target = fp32_round(st.get_state(0))
    out = {}
    for name, c, f64 in (("fp32", cfg, False), ("fp32_fp64_trajectory", cfg, True), ("fp64", cfg64, False)):
        prob = RampTrapProblem(c, psi0, nsweeps, K, duration, objective=gpe.OBJ_OVERLAP, target=target, bq=4e-3,
                               fp64_trajectory=f64)
        prob.value_and_grad(knots, c.trap)  # warm-up (allocations, autotune-free)
        t0 = time.perf_counter()
        J, gk, gt = prob.value_and_grad(knots, c.trap)
        wall = time.perf_counter() - t0
        res = prob.last
        prob.close()
        out[name] = dict(J=J, grad_knots=gk, grad_trap=gt, wall_s=wall, forward_ms=res["forward_ms"],
                         backward_ms=res["backward_ms"], checkpoint_bytes=res["checkpoint_bytes"],
                         segment_bytes=res["segment_bytes"])
        log("adjoint %s: J=%.9f, %.3f s forward+backward (forward %.1f ms, backward %.1f ms), checkpoints %.2f GB, "
            "segment %.2f GB" % (name, J, wall, res["forward_ms"], res["backward_ms"], res["checkpoint_bytes"] / 1e9,
                                  res["segment_bytes"] / 1e9))
    # central differences of FP64 forward runs at h and h / 2; their Richardson combination
    # (4 D(h/2) - D(h)) / 3 removes the O(h^2) term, which for the trap frequencies is not small:
    # a trap change dephases the condensate (phase ~ t omega r^2), so J has large higher
    # derivatives next to its small first derivative near the overlap maximum.  h = 1e-4 for the
    # trap frequencies, 1e-2 for the knots: their gradients (~1e-8) are so small that at 1e-4
    # the rounding of J (~1e-16 / h) is already 1e-3 of them
    h = np.array([1e-2] * 16 + [1e-4] * 3)
    r32 = out["fp32"]
    level_bytes = float(np.prod(cfg.dims) * cfg.V * 3 * 8)
    line = {
        "metric": "spin-1 BCC 64^3 discrete adjoint: seconds per forward+backward (4000 sweeps, checkpoints every "
                  "200), gradient error vs central finite differences",
        "value": r32["wall_s"],
        "unit": "s per forward+backward",
        "n_gpus": 1, "steps": nsweeps, "warmup": 1, "ms_per_step": 1e3 * r32["wall_s"] / nsweeps,
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
        "config": {"workload": "adjoint_spin1_bcc_64^3", "cells": list(cells), "box": box, "sweeps": nsweeps,
                   "checkpoint_every": K, "parameters": "16 Bz knots + 3 trap frequencies", "objective": "overlap",
                   "fd_step": {"knots": 1e-2, "trap": 1e-4}, "fd_scheme": "central at h and h/2, Richardson-combined"},
        "memory_bytes": {"levels": 2 * level_bytes, "checkpoints": r32["checkpoint_bytes"],
                         "segment": r32["segment_bytes"]},
        "forward_ms": r32["forward_ms"], "backward_ms": r32["backward_ms"],
        "fp64_form": {"value": out["fp64"]["wall_s"], "forward_ms": out["fp64"]["forward_ms"],
                      "backward_ms": out["fp64"]["backward_ms"]},
        "fp32_fp64_trajectory_form": {"value": out["fp32_fp64_trajectory"]["wall_s"],
                                      "forward_ms": out["fp32_fp64_trajectory"]["forward_ms"],
                                      "backward_ms": out["fp32_fp64_trajectory"]["backward_ms"]},
        "accuracy": {"J_" + k: r["J"] for k, r in out.items()},
        "roofline": None, "cpu_baseline": None,
    }
    # agreement of the FP32 forms with the FP64 adjoint, per group (relative to the group's largest
    # FP64 gradient): needs no finite differences
    g64 = np.concatenate([out["fp64"]["grad_knots"], out["fp64"]["grad_trap"]])
    grp = {"knots": slice(0, 16), "trap": slice(16, 19)}
    line["accuracy"]["rel_diff_vs_fp64_adjoint"] = {
        k: {gn: float(np.abs(np.concatenate([r["grad_knots"], r["grad_trap"]])[sl] - g64[sl]).max() /
                      np.abs(g64[sl]).max()) for gn, sl in grp.items()}
        for k, r in out.items() if k != "fp64"}
    if args.no_fd:  # timing only (profiling runs): no finite-difference check
        return line
    fd = RampTrapProblem(cfg64, psi0, nsweeps, K, duration, objective=gpe.OBJ_OVERLAP, target=target, bq=4e-3)
    t0 = time.perf_counter()

    def central(steps):
        g = np.zeros(19)
        for j in range(19):
            step = steps[j]
            kp, km = knots.copy(), knots.copy()
            tp, tm = np.array(cfg.trap, float), np.array(cfg.trap, float)
            if j < 16:
                kp[j] += step
                km[j] -= step
            else:
                tp[j - 16] += step
                tm[j - 16] -= step
            g[j] = (fd.value(kp, tp) - fd.value(km, tm)) / (2 * step)
        return g

    g_h = central(h)
    g_h2 = central(h / 2)
    fd.close()
    g_fd = (4 * g_h2 - g_h) / 3
    log("finite differences: 76 FP64 forward runs in %.1f s" % (time.perf_counter() - t0))
    # errors relative to the largest finite difference of each group: the knot gradients
    # (~1e-8) are two orders below the trap gradients (~1e-6)
    groups = {"knots": slice(0, 16), "trap": slice(16, 19)}
    err, err_raw = {}, {}
    for name, r in out.items():
        g = np.concatenate([r["grad_knots"], r["grad_trap"]])
        err[name] = {k: float(np.abs(g[sl] - g_fd[sl]).max() / np.abs(g_fd[sl]).max()) for k, sl in groups.items()}
        err_raw[name] = {k: float(np.abs(g[sl] - g_h[sl]).max() / np.abs(g_h[sl]).max()) for k, sl in groups.items()}
        log("%s: max |adjoint - FD| / max |FD|: knots %.2e, trap %.2e (plain h: %.2e, %.2e)" % (
            name, err[name]["knots"], err[name]["trap"], err_raw[name]["knots"], err_raw[name]["trap"]))
    line["accuracy"].update({
        "max_rel_grad_err_fp32_vs_fd": max(err["fp32"].values()),
        "max_rel_grad_err_fp32_fp64_trajectory_vs_fd": max(err["fp32_fp64_trajectory"].values()),
        "max_rel_grad_err_fp64_vs_fd": max(err["fp64"].values()),
        "per_group": err, "per_group_vs_plain_fd_h": err_raw,
        "grad": {k: np.concatenate([r["grad_knots"], r["grad_trap"]]).tolist() for k, r in out.items()},
        "grad_fd_richardson": g_fd.tolist(), "grad_fd_h": g_h.tolist()})
    return line


def prepare_state(cfg, knots, itp_steps, itp_dt, log):
    """FP64 ground state on the device: Thomas-Fermi fill at the unit-norm chemical potential,
    imaginary time with renormalisation, then for spinors the polar projection and 0.5% spinor
    noise (device kernels), for the scalar the vortex-ring imprint and 1% noise (host)."""
    from spacetime_amd import gpe, lattice, states
    import dataclasses
    lam3 = cfg.trap[0] * cfg.trap[1] * cfg.trap[2]
    mu = 0.5 * (15.0 * cfg.c0 * lam3 / (4.0 * np.pi)) ** 0.4
    cfg64 = dataclasses.replace(cfg, precision=gpe.FP64, scheme=gpe.LEAPFROG3)
    # imaginary-time Euler contracts every mode only for dt * lambda_max < 2: keep dt * ||H|| <= 1
    lam = lattice.get(cfg.lattice).kinetic_bound(cfg.block_scale) + mu
    if itp_dt * lam > 1.0:
        log("itp: dt %g lowered to %g (kinetic bound %.4g)" % (itp_dt, 1.0 / lam, lam))
        itp_dt = 1.0 / lam
    weights = np.zeros(cfg.ncomp, dtype=np.complex128)
    weights[cfg.ncomp // 2] = 1.0
    t0 = time.time()
    with gpe.Stepper(cfg64) as st:
        st.fill_thomas_fermi(cfg.c0, mu, weights)
        st.itp(itp_steps, itp_dt)
        e = st.observables(0)
        if cfg.ncomp > 1:
            st.project("polar")
            st.add_noise(0.005, 54363)
            st.scale(1.0 / np.sqrt(st.observables(0).norm))
        psi = st.get_state(0)
    log("itp: %d steps dt=%g in %.2fs, E=%.6f mu_TF=%.4f" % (itp_steps, itp_dt, time.time() - t0, e.energy, mu))
    if cfg.ncomp == 1:
        r = states.positions(cfg)
        psi = psi * states.vortex_ring_phase(r, 2.0)[None]
        del r
        psi = states.multiplicative_noise(psi, 0.01, 54363)
        psi *= np.sqrt(1.0 / states.norm(cfg, psi))
    return psi


PREROLL_MS = 40.0


def run_form(cfg, psi0, knots, warmup, steps, log, tune_sweeps=0, traversal=None):
    """W warm-up steps and K timed steps of one form.  Right before the timed steps the same
    sweep kernel runs for PREROLL_MS on a scratch copy of the run (untimed; the run's own state
    and step count are untouched): the GPU's clocks go through a transient in the first ~15 ms
    of sustained load after any idle period (sweeps ~7% slower 3-8 ms into a burst, measured
    per launch with rocprofv3), which a 20-step timed region would otherwise measure instead of
    the steady rate that long runs see."""
    from spacetime_amd import gpe, states

    def prepare(st):
        if knots is not None:
            st.set_ramp(q_segments=[(4e-3, 1e30, 0)], b_segments=states.knot_ramp(knots, 2.0))
        st.set_state(psi0)

    with gpe.Stepper(cfg) as st, gpe.Stepper(cfg) as pre:
        prepare(st)
        o0 = st.observables(0)
        st.start(gpe.START_EULER)
        if traversal and (cfg.precision == gpe.FP32 or traversal[2] == 1):
            st.set_traversal(*traversal)
            tune = st.traversal()
        else:
            tune = st.autotune(tune_sweeps) if tune_sweeps else st.traversal()
        log("%s traversal: %s" % ("fp64" if cfg.precision else "fp32", tune))
        prepare(pre)
        pre.start(gpe.START_EULER)
        pre.set_traversal(tune["tile_rows"], tune["xcd_remap"], tune["x_per_lane"])
        if warmup:
            st.step(warmup)
        ow = st.observables(0)  # after the start-up transient, outside the timed region
        est = st.last_step_timing()[0] if warmup else 0.0
        pre.step(max(20, int(PREROLL_MS / max(est, 0.05))))
        st.sync()
        _sync()
        t0 = time.perf_counter()
        st.step(steps)
        st.sync()
        _sync()
        wall = time.perf_counter() - t0
        ms_sweep, bytes_sweep = st.last_step_timing()
        o1 = st.observables(0)
        psi = st.get_state(0)
    per = 2 if cfg.scheme == gpe.STAGGERED else 1
    nodes = np.prod(cfg.dims) * cfg.V
    res = dict(wall_s=wall, ms_per_step=1e3 * wall / steps, kernel_ms_per_sweep=ms_sweep,
               bytes_per_sweep=bytes_sweep, node_updates_per_s=nodes * steps * per / wall,
               kernel_node_updates_per_s=nodes / (ms_sweep * 1e-3),
               achieved_gbs=bytes_sweep / (ms_sweep * 1e-3) / 1e9,
               norm_drift=abs(o1.norm - o0.norm) / o0.norm, energy_drift=abs(o1.energy - o0.energy) / abs(o0.energy),
               energy_drift_timed=abs(o1.energy - ow.energy) / abs(ow.energy),
               energy=o1.energy, norm=o1.norm, traversal=tune)
    log("%s: %d steps in %.3fs, kernel %.4f ms/sweep, %.3e node-updates/s, %.1f GB/s" % (
        "fp64" if cfg.precision else "fp32", steps, wall, ms_sweep, res["node_updates_per_s"], res["achieved_gbs"]))
    return res, psi


def cpu_baseline(cfg, psi0, log, budget_s=15.0):
    """FP64 oracle (C restatement of the reference update) on the host cores, same grid."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_lib import Oracle
    from helpers import oracle_params, with_field
    from spacetime_amd import lattice
    o = Oracle()
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or len(os.sched_getaffinity(0))
    threads = o.set_threads(threads)
    L = o.lattice(lattice.get(cfg.lattice))
    p = oracle_params(o, cfg)
    if cfg.ncomp > 1:
        with_field(p, bq=4e-3, bb=(0.0, 0.0, 0.0))
    a = np.ascontiguousarray(psi0)
    b = a.copy()
    coef = cfg.dt if cfg.scheme == 1 else 2 * cfg.dt
    t0 = time.perf_counter()
    o.update(L, p, b, a, coef)
    t1 = time.perf_counter() - t0
    n = max(1, min(400, int(budget_s / max(t1, 1e-3))))
    t0 = time.perf_counter()
    for k in range(n):
        o.update(L, p, b, a, coef)
        a, b = b, a
    el = time.perf_counter() - t0
    nodes = np.prod(cfg.dims) * cfg.V
    val = nodes * n / el
    log("cpu oracle: %d sweeps of %d unknowns in %.2fs on %d threads: %.3e node-updates/s" % (n, nodes, el, threads,
                                                                                          val))
    return dict(value=val, unit="node-updates/s", cores=threads, kind="port", cpu_model=cpu_model(),
                sample="%d FP64 sweeps of the full %s grid (%d unknowns) with the C oracle" % (n, "x".join(
                    str(d) for d in cfg.dims), nodes))


def cpu_reference(log, iterations=3):
    """The reference's own scalar CPU solver, timed: GrossPitaevskii/GPBlocks.cpp's
    integrateInTime compiled from the reference sources into oracle/_ref (single thread, FP64
    staggered steps), on its CPU-sized configuration: BCC blocks of edge 0.25 over [-4, 4]^3
    (32^3 blocks, 393,216 unknowns), g = 300, vortex-free relaxed initial state.  node-updates/s
    from the wall time between its per-iteration output lines (2 x steps x unknowns each)."""
    import subprocess
    import tempfile
    exe = os.path.join(ROOT, "oracle", "_ref", "gpblocks_driver")
    if not os.path.exists(exe):
        log("cpu reference: %s not built" % exe)
        return None
    args = [exe, "0", "300", "4", "4", "0.05", "0.05", "2000", "%.17g" % (0.25 / 8.0 ** 0.5), "1", str(iterations)]
    grid, steps, stamps = None, None, []
    with tempfile.TemporaryDirectory() as d:  # the solver writes a picture per iteration into its cwd
        proc = subprocess.Popen(args, cwd=d, stdout=subprocess.PIPE, text=True, bufsize=1)
        for line in proc.stdout:
            now = time.perf_counter()
            if line.startswith("GRID "):
                grid = [int(v) for v in line.split()[1:5]]
            elif line.startswith("STEPS_PER_ITERATION "):
                steps = int(line.split()[1])
            elif line.startswith("normsq="):
                stamps.append(now)
        proc.wait()
    if proc.returncode != 0 or not grid or not steps or len(stamps) < 2:
        log("cpu reference: run failed (rc %s)" % proc.returncode)
        return None
    unknowns = grid[0] * grid[1] * grid[2] * grid[3]
    per_iter = min(b - a for a, b in zip(stamps, stamps[1:]))
    val = 2.0 * steps * unknowns / per_iter
    log("cpu reference (GPBlocks.cpp, 1 thread): %d unknowns, %d staggered steps per iteration in %.2fs: "
        "%.3e node-updates/s" % (unknowns, steps, per_iter, val))
    return dict(value=val, unit="node-updates/s", cores=1, kind="reference", cpu_model=cpu_model(),
                sample="GrossPitaevskii/GPBlocks.cpp integrateInTime compiled from the reference sources, "
                       "BCC %dx%dx%d blocks (%d unknowns), g = 300, one iteration of %d staggered steps "
                       "(%d sweeps), fastest of %d" % (grid[0], grid[1], grid[2], unknowns, steps, 2 * steps,
                                                       len(stamps) - 1))


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps (default: 20000 spin1, 10000 scalar, 5000 ensemble)")
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--workload", choices=["spin1", "scalar", "ensemble", "adjoint", "lattices"], default="spin1")
    ap.add_argument("--graph-sweeps", type=int, default=20)
    ap.add_argument("--tune-sweeps", type=int, default=10,
                    help="autotuner trial sweeps per traversal candidate (0: default traversal)")
    ap.add_argument("--traversal", type=lambda v: tuple(int(x) for x in v.split(",")), default=None,
                    help="fixed FP32 sweep traversal 'rows,remap,x_per_lane' instead of the autotuner")
    ap.add_argument("--cases", default=None,
                    help="lattices workload: only these cases, 'lattice:C[:NXxNYxNZ],...' (e.g. 'a15:5,bcc:1' or "
                         "'a15:5:64x82x105' for another box)")
    ap.add_argument("--itp-steps", type=int, default=1000)
    ap.add_argument("--itp-dt", type=float, default=2e-4)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-fd", action="store_true", help="adjoint workload: skip the finite-difference check")
    ap.add_argument("--no-other-configs", action="store_true",
                    help="default workload: skip the short runs of the scalar, adjoint and ensemble configurations")
    ap.add_argument("--quiet", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (run with --gpus 1).",
              file=sys.stderr)
        sys.exit(2)
    log = (lambda *a: None) if args.quiet else (lambda m: print("[bench] " + m, file=sys.stderr, flush=True))
    _init_torch()
    if args.steps is None:
        args.steps = {"spin1": 20000, "scalar": 10000, "ensemble": 5000, "adjoint": 4000,
                      "lattices": 1000}[args.workload]
    if args.workload == "ensemble" and (args.graph_sweeps <= 0 or args.graph_sweeps % 2 or
                                        args.steps % args.graph_sweeps):
        ap.error("--graph-sweeps must be even and divide --steps")

    if args.workload == "ensemble":
        print(json.dumps(run_ensemble(args, log)))
        return
    if args.workload == "adjoint":
        print(json.dumps(run_adjoint(args, log)))
        return
    if args.workload == "lattices":
        print(json.dumps(run_lattices(args, log)))
        return

    from spacetime_amd import gpe
    import dataclasses
    cfg, knots, info = workload_spin1() if args.workload == "spin1" else workload_scalar()
    nodes = int(np.prod(cfg.dims) * cfg.V)
    info.update(unknowns=nodes, lattice_points=int(2 * np.prod(cfg.dims)))
    log("workload %s: %d unknowns" % (info["workload"], nodes))
    copy_gbs, update_gbs = copy_bandwidth()
    if copy_gbs:
        log("streaming kernels (2 x 1 GiB): copy %.1f GB/s, in-place read-2-write-1 %.1f GB/s" % (copy_gbs,
                                                                                            update_gbs))
    psi0 = prepare_state(cfg, knots, args.itp_steps, args.itp_dt, log)

    r32, psi32 = run_form(dataclasses.replace(cfg, precision=gpe.FP32), psi0, knots, args.warmup, args.steps, log,
                          args.tune_sweeps, args.traversal)
    acc = {}
    r64 = None
    if not args.no_fp64:
        r64, psi64 = run_form(dataclasses.replace(cfg, precision=gpe.FP64), psi0, knots, args.warmup, args.steps,
                              log, args.tune_sweeps)
        acc["rel_l2_fp32_vs_fp64"] = float(np.linalg.norm((psi32 - psi64).ravel()) / np.linalg.norm(psi64.ravel()))
        del psi64
    acc.update(norm_drift_fp32=r32["norm_drift"], energy_drift_fp32=r32["energy_drift"],
               energy_drift_timed_fp32=r32["energy_drift_timed"])
    if r64:
        acc.update(norm_drift_fp64=r64["norm_drift"], energy_drift_fp64=r64["energy_drift"],
                   energy_drift_timed_fp64=r64["energy_drift_timed"])
        acc["speedup_fp32_over_fp64"] = r32["node_updates_per_s"] / r64["node_updates_per_s"]
    del psi32
    cpu, cpu_ref = None, None
    if not args.no_cpu:
        cpu = cpu_baseline(cfg, psi0, log)
        cpu_ref = cpu_reference(log)
        if cfg.ncomp == 1 and cpu_ref:  # the reference has a CPU path for the scalar GPE only
            cpu, cpu_ref = cpu_ref, cpu

    line = {
        "metric": METRIC,
        "value": r32["node_updates_per_s"],
        "unit": "node-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r32["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": info,
        "roofline": {
            "bound": "hbm",
            "achieved": r32["achieved_gbs"],
            "peak": PEAK_HBM_GBS,
            "unit": "GB/s",
            "frac": r32["achieved_gbs"] / PEAK_HBM_GBS,
            "traffic": PMC_TRAFFIC.get(info["workload"]),
            "kernel_ms_per_sweep": r32["kernel_ms_per_sweep"],
            "traversal": r32["traversal"],
            "copy_gbs": copy_gbs,
            "stream_update_gbs": update_gbs,
            "frac_of_stream_update": (r32["achieved_gbs"] / update_gbs) if update_gbs else None,
            "algorithmic_bytes_per_sweep": r32["bytes_per_sweep"],
        },
        "cpu_baseline": cpu,
        "cpu_other": cpu_ref,
        "fp64_form": None if r64 is None else {
            "value": r64["node_updates_per_s"], "ms_per_step": r64["ms_per_step"],
            "achieved_gbs": r64["achieved_gbs"], "frac": r64["achieved_gbs"] / PEAK_HBM_GBS},
        "accuracy": acc,
    }
    if args.workload == "spin1" and not args.no_other_configs:
        line["other_configs"] = run_other_configs(log)
    print(json.dumps(line))


OTHER_CONFIGS = (
    # name, arguments of a child bench run, seconds allowed
    ("scalar_bcc_192^3", ["--workload", "scalar", "--steps", "1000", "--warmup", "20", "--no-cpu"], 400),
    ("adjoint_spin1_bcc_64^3", ["--workload", "adjoint", "--no-fd"], 400),
    ("ensemble_spin2_48", ["--workload", "ensemble", "--steps", "100", "--graph-sweeps", "20"], 600),
)


def run_other_configs(log):
    """The other configurations' lines, each from a child `bench.py` run (so that a failure there
    cannot take the headline line with it), in short forms: the scalar BCC 192^3 staggered run
    for 1000 steps, the spin-1 BCC 64^3 adjoint without its finite-difference check (the FP32
    forms' gradients against the FP64 adjoint's instead), the 48-member spin-2 ensemble for 100
    steps.  The full-length runs are `--workload scalar|adjoint|ensemble`."""
    import subprocess
    out = {}
    for name, argv, limit in OTHER_CONFIGS:
        t0 = time.perf_counter()
        log("other configuration %s: bench.py %s" % (name, " ".join(argv)))
        try:
            r = subprocess.run([sys.executable, os.path.abspath(__file__)] + argv, capture_output=True, text=True,
                               timeout=limit)
            lines = [ln for ln in r.stdout.splitlines() if ln.startswith("{")]
            if r.returncode != 0 or not lines:
                tail = (r.stderr or "").strip().splitlines()[-3:]
                out[name] = {"error": "exit %d: %s" % (r.returncode, " | ".join(tail))}
                continue
            res = json.loads(lines[-1])
            keep = {k: res.get(k) for k in ("metric", "value", "unit", "steps", "ms_per_step", "dtype", "config")}
            for k in ("roofline", "accuracy", "fp64_form", "fp32_fp64_trajectory_form", "memory_bytes",
                      "forward_ms", "backward_ms", "small_members", "speedup_vs_sequential"):
                if k in res:
                    keep[k] = res[k]
            keep["wall_s"] = time.perf_counter() - t0
            out[name] = keep
        except subprocess.TimeoutExpired:
            out[name] = {"error": "timed out after %d s" % limit}
    return out


if __name__ == "__main__":
    main()
