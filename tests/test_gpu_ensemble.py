"""The 64-problem ensemble workload, batched: cases.ensemble(64) (numpy default_rng(2024)):
finest grids N x N with N from {256, 512, 768, 1024} on [-2,6] x [-4,4], 3-5 levels, Re
uniform in [50, 300], one stationary ellipse per problem (a = 0.5, b uniform in [0.1, 0.5],
incidence uniform in [0, 30] deg, points spaced h), dt = h/5.

Every problem advanced in the batch (one graph replay per step) must end bitwise equal to
the same problem advanced alone, and two members (256^2 and 512^2, 5 levels) must match
the FP64 oracle frozen in tests/golden/oracle_ensemble.npz (make_golden_ensemble.py):
forces within 1e-9 and fields within 1e-10 in FP64, forces within 1e-6 in mixed precision.
"""
import os

import numpy as np
import pytest

import cases
import immersedbodies as ib

pytestmark = pytest.mark.gpu

STEPS = 20


def _fixture():
    return np.load(os.path.join(os.path.dirname(__file__), "golden", "oracle_ensemble.npz"))


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_ensemble_64_batched_equals_alone_and_oracle(precision):
    probs = cases.ensemble(64)
    ref = _fixture()
    assert int(ref["steps"]) == STEPS
    print("case ensemble(64)", precision, "alone", flush=True)
    alone = [ib.create_solver(p, precision=precision) for p in probs]
    f_alone = np.stack([s.forces(s.advance(STEPS)) for s in alone], 1)
    print("case ensemble(64)", precision, "batched", flush=True)
    batched = [ib.create_solver(p, precision=precision) for p in probs]
    batch = ib.Batch(batched)
    f_b = batch.advance(STEPS)
    batch.close()
    assert np.array_equal(f_b, f_alone)
    assert np.isfinite(f_b).all()
    for k, (a, b) in enumerate(zip(alone, batched)):
        for name in ("Gamma", "Ftilde"):
            assert np.array_equal(a.read(name), b.read(name)), (k, name)
    for k in (int(m) for m in ref["members"]):
        fr = ref[f"m{k}/forces"]
        e = cases.relerr(f_b[:, k], fr)
        print(f"member {k} {probs[k].fluid.grid.base.n}: force rel err {e:.2e}", flush=True)
        if precision == "fp64":
            assert e <= 1e-9
            s = batched[k]
            assert cases.relerr(s.read("Ftilde"), ref[f"m{k}/Ftilde"]) <= 1e-9
            L = probs[k].fluid.grid.nlevel
            for name in ("Gamma", "psi", "q"):
                f = s.read(name)
                norms = np.linalg.norm(f.reshape(L, -1), axis=1)
                assert np.allclose(norms, ref[f"m{k}/{name}/norms"], rtol=1e-10, atol=0), name
                err = np.linalg.norm(f[ref[f"m{k}/{name}/idx"]] - ref[f"m{k}/{name}/values"])
                assert err <= 1e-10 * np.linalg.norm(ref[f"m{k}/{name}/norms"]), (name, err)
        else:
            assert e <= 1e-6
    for s in alone + batched:
        s.close()
