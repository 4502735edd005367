"""Regenerate tests/golden/reference_digests.json by running the REFERENCE's
own per-block solver code (oracle/_ref/ref_harness, built from the sources
under /root/reference by oracle/Makefile) on every scenario of
scenarios.py.  Run in the container that has the reference:

    make -C oracle && python tests/golden/make_golden.py
"""
from __future__ import annotations

import hashlib
import json
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent.parent))

import numpy as np  # noqa: E402

from oracle.binding import ref_available, run_reference  # noqa: E402
from tests.golden.scenarios import pops_for, scenarios  # noqa: E402


def digest(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.float64).tobytes()).hexdigest()


def main():
    if not ref_available():
        raise SystemExit("oracle/_ref/ref_harness not built (needs /root/reference)")
    out = {}
    for name, (factory, seed, steps, t0) in scenarios().items():
        case = factory()
        pops = pops_for(case, seed)
        ref, avg, mass, drag = run_reference(case, pops, steps, t0)
        out[name] = {
            "seed": seed, "steps": steps, "t0": t0,
            "input_sha256": digest(pops),
            "pops_sha256": digest(ref),
            "average_rho": avg.hex(), "mass": mass.hex(),
            "drag": [float(d).hex() for d in drag],
            "pops_abs_sum": float(np.abs(ref).sum()),
        }
        print(name, out[name]["pops_sha256"][:16], flush=True)
    (HERE / "reference_digests.json").write_text(json.dumps(out, indent=1) + "\n")


if __name__ == "__main__":
    main()
