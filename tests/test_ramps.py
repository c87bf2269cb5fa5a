"""Reference ramp presets (tables/ramps.json) and unit conversions.

Pinned by (i) re-running the converter on the reference headers, where the checkout exists,
which must reproduce the committed JSON and its source hashes exactly, and (ii) the ramp
semantics of the library (pinned separately against closed forms in test_oracle.py)
evaluated on the presets at times where the reference's getSignal has closed-form values."""
import hashlib
import json
import math
import os

import numpy as np
import pytest

from spacetime_amd import ramps

REF = "/root/reference"
TABLES = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "spacetime_amd", "tables")


def test_presets_and_units_present():
    assert ramps.presets() == ["spin1_alice_ring", "spin1_knot", "spin2_knot"]
    u1, u2 = ramps.units("spin1"), ramps.units("spin2")
    for u in (u1, u2):
        assert u["omega_r"] == pytest.approx(2 * math.pi * 126.0)
        assert u["ms_per_time_unit"] == pytest.approx(1e3 / u["omega_r"])
        assert u["lambda_z"] == pytest.approx(166.0 / 126.0)
    # c0 = 4 pi N (a0 + 2 a2) a_B / (3 a_r) with a_r = sqrt(hbar / (m omega_r)) (Spin1GpeBCC.cu:59-60)
    a_r = math.sqrt(1.05457148e-34 / (1.44316060e-25 * u1["omega_r"]))
    assert u1["c0"] == pytest.approx(4 * math.pi * 2e5 * (101.8 + 2 * 100.4) * 5.2917721092e-11 / (3 * a_r))
    kw = ramps.solver_config("spin2")
    assert kw["ncomp"] == 5 and kw["c1"] == u2["c2"] and kw["c2"] == u2["c4"] and kw["alpha"] == 0.0


@pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "Spin1GpeBCC")), reason="reference checkout not present")
def test_converter_reproduces_committed_json():
    import importlib.util
    spec = importlib.util.spec_from_file_location("convert_ramps", os.path.join(TABLES, "convert_ramps.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    fresh = json.loads(json.dumps(mod.convert(REF)))
    committed = json.load(open(os.path.join(TABLES, "ramps.json")))
    assert fresh == committed
    for path, digest in committed["provenance"].items():
        assert hashlib.sha256(open(os.path.join(REF, path), "rb").read()).hexdigest() == digest


def test_preset_signals_match_get_signal_closed_forms(oracle):
    """Alice ring (spin-1): Bq ramps linearly 0 -> 4.3 over 10 ms, holds 192.03 ms, drops to 0
    over 0.2 ms; the bias holds 1.0 for 10 ms, ramps to 0.045, ..., then sqrt-ramps to 1.2
    over 0.5 ms.  Spin-2 knot: bias 0.219 for 0.1 ms, linear to 0 over 0.0177 ms, ...,
    then 3.0 (1 - exp(-t / 0.12 ms))."""
    kw = ramps.ramp_kwargs("spin1_alice_ring")
    r = oracle.ramp(q_segments=kw["q_segments"], b_segments=kw["b_segments"], tau=kw["tau"])
    bq, bb = oracle.ramp_signal(r, 5.0)
    assert bq == pytest.approx(4.3 * 0.5) and bb[2] == pytest.approx(1.0)
    bq, bb = oracle.ramp_signal(r, 15.0)
    assert bq == pytest.approx(4.3) and bb[2] == pytest.approx(1.0 + 0.5 * (0.045 - 1.0))
    t = 10 + 10 + 2.02 + 180 + 0.125
    bq, bb = oracle.ramp_signal(r, t)
    assert bb[2] == pytest.approx(1.2 * math.sqrt(0.125 / 0.5))
    kw = ramps.ramp_kwargs("spin2_knot")
    r = oracle.ramp(q_segments=kw["q_segments"], b_segments=kw["b_segments"], tau=kw["tau"])
    _, bb = oracle.ramp_signal(r, 0.05)
    assert bb[2] == pytest.approx(0.219)
    _, bb = oracle.ramp_signal(r, 0.1 + 0.0177 / 2)
    assert bb[2] == pytest.approx(0.219 / 2)
    t = 0.1 + 0.0177 + 0.255 + 0.06
    _, bb = oracle.ramp_signal(r, t)
    assert bb[2] == pytest.approx(3.0 * (1 - math.exp(-0.06 / 0.12)))
    assert kw["time_scale"] == pytest.approx(ramps.units("spin2")["ms_per_time_unit"])
    assert kw["scale_b"] == pytest.approx(ramps.units("spin2")["BzScale"]) and kw["qscale_b"] == 0.0
