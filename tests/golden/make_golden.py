"""Write tests/golden/ref_cases.json from the compiled reference.

Runs oracle/_ref/rc_ref (the reference's CPU RayleighChebyshevLM built from
its own headers by `make -C oracle ref`) on each case below and records what
it returns: eigenvalues, residuals, Lanczos iteration count and estimates,
guard value, countData, finalData, the per-iteration verbose trace (degree,
residual max, eigenvalue-difference max, convergence ratio, gap) and the
number of random draws consumed.  Only this container can run it; the fixture
it writes is plain data.

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import json
import os
import subprocess
import sys
import tempfile
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from tests.golden.cases import CASES, build_operator  # noqa: E402

STOP = {"COMBINATION": 1, "EIGENVALUE_ONLY": 2, "RESIDUAL_ONLY": 3}


def run_case(name, c, rc_ref, tmp):
    A = build_operator(c)
    path = os.path.join(tmp, f"{name}.bin")
    out = os.path.join(tmp, f"{name}.json")
    A.write_bin(path)
    args = [rc_ref, path, out, str(c["k"]), repr(c["tol"]), str(c["inc"]), str(c["buf"]),
            str(c.get("seed", 0)), str(STOP[c.get("stop", "COMBINATION")]),
            "interval" if c.get("interval") else "lowest", repr(c.get("lambda_max", 0.0)),
            str(c.get("max_inner", 0))]
    t = time.time()
    subprocess.run(args, check=True, stdout=subprocess.DEVNULL, env=dict(os.environ, OMP_NUM_THREADS="8"))
    dt = time.time() - t
    with open(out) as f:
        g = json.load(f)
    g["wall_s"] = round(dt, 3)
    g["case"] = c
    return g


def main():
    rc_ref = os.path.join(ROOT, "oracle", "_ref", "rc_ref")
    if not os.path.exists(rc_ref):
        sys.exit("build the reference harness first: make -C oracle ref")
    results = {}
    with tempfile.TemporaryDirectory() as tmp:
        for name, c in CASES.items():
            g = run_case(name, c, rc_ref, tmp)
            print(f"{name}: found={g['found']} lanczos={g['lanczos_iters']} "
                  f"iters={len(g['trace'])} wall={g['wall_s']}s", flush=True)
            results[name] = g
    with open(os.path.join(HERE, "ref_cases.json"), "w") as f:
        json.dump(results, f, indent=1)


if __name__ == "__main__":
    main()
