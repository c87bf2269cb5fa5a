"""The reference's field-ramp presets and unit conversions (data in tables/ramps.json).

Presets (segment tables converted from the reference headers by tables/convert_ramps.py):

* ``spin1_alice_ring`` -- ``Spin1GpeBCC/AliceRingRamps.h:29-45`` (decay bias enabled),
* ``spin1_knot``       -- ``Spin1GpeBCC/KnotRamps.h:29-38``,
* ``spin2_knot``       -- ``Spin2GpeBCC/KnotRamps.h:21-29`` (FAST_EXTRACTION = 1 - exp(-t/tau)).

:func:`ramp_kwargs` returns the keyword arguments of :meth:`gpe.Stepper.set_ramp` that make
the library evaluate the reference's ``getSignal`` at ramp time t_ms = ms_per_time_unit * t
and scale it as the solver does (``Spin1GpeBCC.cu:1577-1587``: B = BqScale * Bq, Bb =
BzScale * Bb; the quadratic-Zeeman field with BqQuadScale, BzQuadScale when enabled).
:func:`solver_config` gives the solver's dimensionless couplings and trap ratios as
:class:`gpe.Config` keyword arguments, with the reference's coupling names mapped to this
library's (spin-1 c2 -> c1; spin-2 c2 -> c1 and c4 -> c2).
"""
import json
import os

_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tables", "ramps.json")
_DATA = None


def _data():
    global _DATA
    if _DATA is None:
        with open(_PATH) as f:
            _DATA = json.load(f)
    return _DATA


def presets():
    return sorted(_data()["presets"])


def units(spin):
    """The solver's unit conversions and couplings ("spin1" or "spin2")."""
    return dict(_data()["units"][spin])


def ramp_kwargs(name, quadratic_zeeman=False):
    """Stepper.set_ramp keyword arguments of a reference ramp preset."""
    p = _data()["presets"][name]
    u = _data()["units"][name.split("_")[0]]
    q = p["q"]
    b = p["b"]
    return dict(
        q_segments=[(v, d, t) for v, d, t in zip(q["values"], q["durations"], q["types"])],
        b_segments=[(tuple(v), d, t) for v, d, t in zip(b["values"], b["durations"], b["types"])],
        tau=p.get("tau", 1.0),
        time_scale=u["ms_per_time_unit"],
        scale_q=u["BqScale"],
        scale_b=u["BzScale"],
        qscale_q=u["BqQuadScale"] if quadratic_zeeman else 0.0,
        qscale_b=u["BzQuadScale"] if quadratic_zeeman else 0.0,
    )


def solver_config(spin, three_body_loss=False):
    """gpe.Config keyword arguments with the reference solver's couplings and trap."""
    u = _data()["units"][spin]
    kw = dict(trap=(u["lambda_x"], u["lambda_y"], u["lambda_z"]), c0=u["c0"],
              alpha=u["alpha"] if three_body_loss else 0.0)
    if spin == "spin1":
        kw.update(ncomp=3, c1=u["c2"])
    else:
        kw.update(ncomp=5, c1=u["c2"], c2=u["c4"])
    return kw
