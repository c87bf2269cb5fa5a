"""Convert the reference's field-ramp tables and unit conversions into repo-owned JSON.

This is the one-off recipe that produced ``ramps.json`` next to it.  It reads the reference
checkout (default ``/root/reference``) as text, so it only runs where that checkout exists;
``spacetime_amd/ramps.py`` uses the JSON.

Extracted:

* the quadrupole (``Bqs``, ``BqDurations``, ``BqTypes``) and bias (``Bbs``, ``BbDurations``,
  ``BbTypes``) segment tables of the Alice-ring and knot ramps of the spin-1 and spin-2
  solvers (``Spin1GpeBCC/AliceRingRamps.h:29-45`` with ``ENABLE_DECAY_BIAS`` = 1,
  ``Spin1GpeBCC/KnotRamps.h:29-38``, ``Spin2GpeBCC/KnotRamps.h:21-29``); durations are in
  ms and may name constants of the solver (``EXPANSION_START``, ``OPT_TRAP_OFF`` ...), which
  are evaluated from its ``constexpr`` definitions.  ``Spin2GpeBCC/AliceRingRamps.h`` is not
  converted: it names ``CREATION_RAMP_START``, which ``Spin2GpeBCC.cu`` does not define (the
  shipped spin-2 solver includes the knot ramps, ``Spin2GpeBCC.cu:68-69``);
* the meaning of ``FAST_EXTRACTION``: sqrt(t/T) in the spin-1 ``getSignal``
  (``AliceRingRamps.h:96-98``), 1 - exp(-t/PROJECTION_RAMP_DURATION) in the spin-2 one
  (``Spin2GpeBCC/KnotRamps.h:80-82``);
* the solver's unit conversions: trap frequencies, omega_r, the ms-per-time-unit factor
  1e3/omega_r (``Spin1GpeBCC.cu:1578``), the field scales ``BqScale``, ``BzScale``,
  ``BqQuadScale``, ``BzQuadScale`` (``Spin1GpeBCC.cu:84-89``, ``Spin2GpeBCC.cu:137-142``) and
  the interaction constants (c0, c2 for spin-1; c0, c2, c4 for spin-2; alpha).

Usage: ``python convert_ramps.py [reference_root]``
"""
import hashlib
import json
import math
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))

PRESETS = {
    "spin1_alice_ring": ("Spin1GpeBCC/AliceRingRamps.h", "Spin1GpeBCC/Spin1GpeBCC.cu", "sqrt"),
    "spin1_knot": ("Spin1GpeBCC/KnotRamps.h", "Spin1GpeBCC/Spin1GpeBCC.cu", "sqrt"),
    "spin2_knot": ("Spin2GpeBCC/KnotRamps.h", "Spin2GpeBCC/Spin2GpeBCC.cu", "exp"),
}
SOLVERS = {"spin1": "Spin1GpeBCC/Spin1GpeBCC.cu", "spin2": "Spin2GpeBCC/Spin2GpeBCC.cu"}
UNIT_NAMES = {
    "spin1": ["trapFreq_r", "trapFreq_z", "omega_r", "lambda_x", "lambda_y", "lambda_z", "c0", "c2", "alpha",
              "BqScale", "BzScale", "BqQuadScale", "BzQuadScale", "N"],
    "spin2": ["trapFreq_r", "trapFreq_z", "omega_r", "lambda_x", "lambda_y", "lambda_z", "c0", "c2", "c4", "alpha",
              "BqScale", "BzScale", "BqQuadScale", "BzQuadScale", "N", "PROJECTION_RAMP_DURATION"],
}
TYPES = {"CONSTANT": 0, "LINEAR": 1}


def _strip_comments(text):
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return re.sub(r"//[^\n]*", "", text)


def _preprocess(text):
    """Keep the active branches of #if NAME / #else / #endif using the file's own #defines."""
    defines = {m.group(1): m.group(2) for m in re.finditer(r"^\s*#define\s+(\w+)\s+(\S+)", text, flags=re.M)}
    out, stack = [], []
    for line in text.splitlines():
        s = line.strip()
        m = re.match(r"#if\s+(\w+)", s)
        if m:
            v = defines.get(m.group(1), "0")
            stack.append(v not in ("0", "false"))
            continue
        if s.startswith("#else"):
            stack[-1] = not stack[-1]
            continue
        if s.startswith("#endif"):
            stack.pop()
            continue
        if all(stack):
            out.append(line)
    return "\n".join(out)


def solver_constants(text):
    """Evaluate the `[static] const[expr] double NAME = expr;` definitions in file order."""
    env = {"PI": math.pi, "sqrt": math.sqrt, "exp": math.exp, "pow": math.pow}
    for m in re.finditer(r"^\s*(?:static\s+)?(?:constexpr|const)\s+double\s+(\w+)\s*=\s*([^;]+);", text,
                         flags=re.M):
        name, expr = m.group(1), m.group(2)
        try:
            env[name] = float(eval(expr, {"__builtins__": {}}, env))
        except Exception:
            pass
    return env


def _array(text, name):
    m = re.search(r"(?:std::array<[^>]*>|std::vector<[^>]*>)\s+%s\s*=\s*\{(.*?)\};" % name, text, flags=re.S)
    if not m:
        raise ValueError("missing table " + name)
    return m.group(1)


def _split(body):
    items, depth, cur = [], 0, ""
    for ch in body:
        if ch == "(":
            depth += 1
        elif ch == ")":
            depth -= 1
        if ch == "," and depth == 0:
            items.append(cur.strip())
            cur = ""
        else:
            cur += ch
    if cur.strip():
        items.append(cur.strip())
    return items


def convert(ref_root):
    out = {"source": "markus-kivioja/SpaceTime ramp tables and solver unit conversions", "presets": {},
           "units": {}, "provenance": {}}
    consts = {}
    for spin, path in SOLVERS.items():
        text = _strip_comments(open(os.path.join(ref_root, path)).read())
        env = solver_constants(text)
        consts[spin] = env
        out["units"][spin] = {k: env[k] for k in UNIT_NAMES[spin]}
        out["units"][spin]["ms_per_time_unit"] = 1e3 / env["omega_r"]
        out["provenance"][path] = hashlib.sha256(open(os.path.join(ref_root, path), "rb").read()).hexdigest()
    for name, (hdr, solver, fast) in PRESETS.items():
        raw = open(os.path.join(ref_root, hdr)).read()
        out["provenance"][hdr] = hashlib.sha256(raw.encode()).hexdigest()
        text = _preprocess(_strip_comments(raw))
        env = consts[name.split("_")[0]]
        ev = lambda e: float(eval(e, {"__builtins__": {}}, env))  # noqa: E731
        types = dict(TYPES, FAST_EXTRACTION=2 if fast == "sqrt" else 3)
        q_val = [ev(e) for e in _split(_array(text, "Bqs"))]
        q_dur = [ev(e) for e in _split(_array(text, "BqDurations"))]
        q_type = [types[e.split("::")[-1]] for e in _split(_array(text, "BqTypes"))]
        b_val = []
        for e in _split(_array(text, "Bbs")):
            m = re.match(r"make_double3\((.*)\)$", e)
            b_val.append([ev(x) for x in _split(m.group(1))])
        b_dur = [ev(e) for e in _split(_array(text, "BbDurations"))]
        b_type = [types[e.split("::")[-1]] for e in _split(_array(text, "BbTypes"))]
        if not (len(q_val) == len(q_dur) == len(q_type) and len(b_val) == len(b_dur) == len(b_type)):
            raise ValueError("inconsistent table lengths in " + hdr)
        preset = {"header": hdr, "solver": solver, "fast_extraction": fast,
                  "q": {"values": q_val, "durations": q_dur, "types": q_type},
                  "b": {"values": b_val, "durations": b_dur, "types": b_type}}
        if fast == "exp":
            preset["tau"] = env["PROJECTION_RAMP_DURATION"]
        out["presets"][name] = preset
    return out


def main():
    ref_root = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    data = convert(ref_root)
    with open(os.path.join(HERE, "ramps.json"), "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
        f.write("\n")
    print("wrote ramps.json:", ", ".join(sorted(data["presets"])))


if __name__ == "__main__":
    main()
