"""Writes tests/golden/oracle_cylinder1024.npz: the FP64 CPU oracle on the full-size
stationary cylinder (1024 x 1024 cells on [-2,6] x [-4,4], 5 levels, Re 200, dt 0.001,
402 points, impulsive start) for 5 steps.

The oracle's dense B^-1 at this size costs minutes of CPU, so the result is frozen here
as a small fixture instead of being recomputed in every test run: the force history,
per-level 2-norms of Gamma, psi and q, F~b, and the values of Gamma, psi and q at 4096
seeded sample indices. tests/test_gpu_parity.py compares the GPU path against it.

Run: python tests/golden/make_golden_cylinder1024.py   (about 3 minutes on 8 cores)
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

STEPS = 5
NSAMPLE = 4096
SEED = 1024


def sample_indices(n, seed=SEED, k=NSAMPLE):
    return np.sort(np.random.default_rng(seed).choice(n, size=min(k, n), replace=False))


def main():
    import oracle as O

    O.set_threads(max(1, min(16, os.cpu_count() or 1)))
    prob = cases.cylinder_1024()
    orc = cases.oracle_for(prob)
    forces = cases.oracle_advance(orc, prob, STEPS)
    L = prob.fluid.grid.nlevel
    out = {"forces": forces, "Ftilde": orc.field("Ftilde")}
    for name in ("Gamma", "psi", "q"):
        f = orc.field(name)
        out[f"{name}/norms"] = np.linalg.norm(f.reshape(L, -1), axis=1)
        idx = sample_indices(f.size)
        out[f"{name}/idx"] = idx
        out[f"{name}/values"] = f[idx]
    np.savez_compressed(os.path.join(HERE, "oracle_cylinder1024.npz"), **out)
    print("forces", forces[-1])


if __name__ == "__main__":
    main()
