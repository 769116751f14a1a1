"""Writes tests/golden/oracle_ensemble.npz: the FP64 CPU oracle on two members of the
64-problem ensemble (cases.ensemble(64), numpy default_rng(2024)) for 20 steps each:

  member 0  256 x 256 cells on [-2,6] x [-4,4], 5 levels, Re 103.6, 75 points
  member 8  512 x 512 cells, 5 levels, Re 121.6, 183 points

Each is a stationary ellipse (a = 0.5) at incidence, impulsive start, dt = h/5. Stored per
member (prefix m<k>/): the force history, F~b, per-level 2-norms of Gamma, psi and q, and
their values at 4096 seeded sample indices. tests/test_gpu_ensemble.py compares the
batched GPU ensemble against it.

Run: python tests/golden/make_golden_ensemble.py   (about a minute on 8 cores)
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

STEPS = 20
MEMBERS = (0, 8)
NSAMPLE = 4096
SEED = 64


def sample_indices(n, seed=SEED, k=NSAMPLE):
    return np.sort(np.random.default_rng(seed).choice(n, size=min(k, n), replace=False))


def main():
    import oracle as O

    O.set_threads(max(1, min(16, os.cpu_count() or 1)))
    probs = cases.ensemble(64)
    out = {"members": np.array(MEMBERS), "steps": np.array(STEPS)}
    for k in MEMBERS:
        prob = probs[k]
        orc = cases.oracle_for(prob)
        forces = cases.oracle_advance(orc, prob, STEPS)
        L = prob.fluid.grid.nlevel
        out[f"m{k}/forces"] = forces
        out[f"m{k}/Ftilde"] = orc.field("Ftilde")
        for name in ("Gamma", "psi", "q"):
            f = orc.field(name)
            out[f"m{k}/{name}/norms"] = np.linalg.norm(f.reshape(L, -1), axis=1)
            idx = sample_indices(f.size)
            out[f"m{k}/{name}/idx"] = idx
            out[f"m{k}/{name}/values"] = f[idx]
        print("member", k, prob.fluid.grid.base.n, "forces", forces[-1], flush=True)
        orc.close()
    np.savez_compressed(os.path.join(HERE, "oracle_ensemble.npz"), **out)


if __name__ == "__main__":
    main()
