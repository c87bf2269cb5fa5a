"""Write golden vectors from the reference's own scalar CPU solver.

Runs ``oracle/_ref/gpblocks_driver`` -- GrossPitaevskii/GPBlocks.cpp and VortexState.cpp
compiled from the reference checkout by ``oracle/Makefile`` -- and stores what it prints:

* the relaxed axisymmetric vortex profile f(r, z) at sample points and its mu
  (VortexState.cpp:109-238),
* the BCC stencil in the uint-offset form (GPBlocks.cpp:397-449) and positions (:380-396),
* per-iteration ``normsq`` and overlap ``error`` of the staggered leapfrog run
  (GPBlocks.cpp:528-538, loop :560-606).

Usage (in a container where the reference checkout exists):
    make -C oracle ref && python tests/golden/make_gpblocks_golden.py
"""
import json
import os
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
DRIVER = os.path.join(ROOT, "oracle", "_ref", "gpblocks_driver")

# (name, kappa, g, rmax, zmax, rstep, zstep, maxiters, block_scale, period, iterations)
CASES = [
    ("k1_g40", 1, 40.0, 4.5, 4.5, 0.15, 0.15, 20000, 0.35, 1, 6),
    ("k2_g100", 2, 100.0, 5.0, 4.0, 0.2, 0.2, 20000, 0.4, 2, 4),
]


def run_case(case):
    name, kappa, g, rmax, zmax, rstep, zstep, maxiters, s, period, iters = case
    args = [DRIVER] + [str(a) for a in (kappa, g, rmax, zmax, rstep, zstep, maxiters, s, period, iters)]
    with tempfile.TemporaryDirectory() as tmp:
        out = subprocess.run(args, cwd=tmp, check=True, capture_output=True, text=True).stdout
    rec = dict(name=name, kappa=kappa, g=g, rmax=rmax, zmax=zmax, rstep=rstep, zstep=zstep, maxiters=maxiters,
               block_scale=s, period=period, iterations=iters, vortex_f=[], bcc_ind=[], bcc_pos=[],
               normsq=[], error=[])
    for line in out.splitlines():
        tok = line.split()
        if not tok:
            continue
        if tok[0] == "VORTEX_MU":
            rec["vortex_mu"] = float(tok[1])
        elif tok[0] == "VORTEX_F":
            rec["vortex_f"].append([float(x) for x in tok[1:]])
        elif tok[0] == "BCC_IND":
            rec["bcc_ind"].append(int(tok[2]))
        elif tok[0] == "BCC_POS":
            rec["bcc_pos"].append([float(x) for x in tok[2:]])
        elif tok[0].startswith("normsq="):
            rec["normsq"].append(float(tok[0].split("=")[1]))
            rec["error"].append(float(tok[1].split("=")[1]))
    return rec


def main():
    recs = [run_case(c) for c in CASES]
    meta = {"generator": "tests/golden/make_gpblocks_golden.py",
            "source": "GrossPitaevskii/GPBlocks.cpp + GrossPitaevskii/VortexState.cpp (compiled reference)",
            "bcc_ind_strides": {"nx": 12, "ny": 1000, "nz": 1000000}, "cases": recs}
    path = os.path.join(HERE, "gpblocks_scalar_bcc.json")
    with open(path, "w") as f:
        json.dump(meta, f, indent=1)
    print("wrote", path)


if __name__ == "__main__":
    main()
