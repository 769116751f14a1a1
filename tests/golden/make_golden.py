"""Writes tests/golden/oracle_tiny.npz: input/output vectors of small IBPM problems.

The reference holds no golden vectors for its solver (test/solver.jl asserts only the
time), so these fixtures are generated by the FP64 CPU oracle, whose functions are pinned
in tests/test_oracle.py. They freeze the oracle's results so that a later change to the
oracle or to the GPU path that moves any number is caught.

Run: python tests/golden/make_golden.py
"""
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (os.path.join(ROOT, "immersedbodies.jl_amd"), os.path.join(ROOT, "oracle"),
          os.path.join(ROOT, "tests")):
    sys.path.insert(0, p)

import numpy as np  # noqa: E402

import cases  # noqa: E402

GOLDEN_CASES = {
    # name: (builder, kwargs, steps)
    "static_cylinder": (cases.cylinder, dict(nx=32, ny=32, nlevel=2, h=0.0625, x0=(-1.0, -1.0),
                                             npts=20, r=0.4), 10),
    "two_lines": (cases.two_lines, dict(nx=40, ny=40, nlevel=2, h=0.05, x0=(-1.0, -1.0)), 8),
    "moving_cylinder": (cases.moving_cylinder, dict(nx=40, ny=40, nlevel=2, h=0.075,
                                                    x0=(-1.5, -1.5)), 5),
}


def run_case(name):
    builder, kw, steps = GOLDEN_CASES[name]
    prob = builder(**kw)
    orc = cases.oracle_for(prob)
    forces = cases.oracle_advance(orc, prob, steps)
    return {f"{name}/forces": forces, f"{name}/Gamma": orc.field("Gamma"),
            f"{name}/Ftilde": orc.field("Ftilde"), f"{name}/xb": orc.field("xb")}


def main():
    out = {}
    for name in GOLDEN_CASES:
        out.update(run_case(name))
    np.savez_compressed(os.path.join(HERE, "oracle_tiny.npz"), **out)
    print("wrote", sorted(out))


if __name__ == "__main__":
    main()
