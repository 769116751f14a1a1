"""Golden fixtures (tests/golden/oracle_tiny.npz, written by make_golden.py from the
pinned FP64 oracle): the oracle must reproduce them, and the GPU FP64 path must match
them within the parity tolerances."""
import numpy as np
import pytest

import cases
from golden.make_golden import GOLDEN_CASES

import os

FIX = np.load(os.path.join(os.path.dirname(__file__), "golden", "oracle_tiny.npz"))


@pytest.mark.parametrize("name", sorted(GOLDEN_CASES))
def test_oracle_reproduces_golden(name):
    builder, kw, steps = GOLDEN_CASES[name]
    prob = builder(**kw)
    orc = cases.oracle_for(prob)
    forces = cases.oracle_advance(orc, prob, steps)
    assert cases.relerr(forces, FIX[f"{name}/forces"]) < 1e-12
    assert cases.relerr(orc.field("Gamma"), FIX[f"{name}/Gamma"]) < 1e-12
    assert cases.relerr(orc.field("Ftilde"), FIX[f"{name}/Ftilde"]) < 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(GOLDEN_CASES))
def test_gpu_matches_golden(name):
    import immersedbodies as ib

    builder, kw, steps = GOLDEN_CASES[name]
    prob = builder(**kw)
    print("golden case", name, flush=True)
    solver = ib.create_solver(prob)
    stats = solver.advance(steps)
    tol_f = 1e-9 if prob.bodies.static else 1e-8
    assert cases.relerr(solver.forces(stats), FIX[f"{name}/forces"]) < tol_f
    assert cases.relerr(solver.read("Gamma"), FIX[f"{name}/Gamma"]) < 1e-10 * (1 if prob.bodies.static else 10)
    assert cases.relerr(solver.read("xb"), FIX[f"{name}/xb"]) < 1e-14
    solver.close()


def test_cylinder1024_fixture_is_symmetric():
    """The frozen 1024^2 oracle run keeps the y-symmetric cylinder's lift at round-off
    against its drag, and its sampled indices are those the generator draws."""
    import os

    from golden.make_golden_cylinder1024 import sample_indices

    ref = np.load(os.path.join(os.path.dirname(__file__), "golden", "oracle_cylinder1024.npz"))
    f = ref["forces"]
    assert np.all(np.isfinite(f)) and np.all(f[:, 0] > 0)
    assert np.abs(f[:, 1]).max() < 1e-9 * np.abs(f[:, 0]).max()
    n = ref["Gamma/norms"].size * 1023 * 1023
    assert np.array_equal(ref["Gamma/idx"], sample_indices(n))
