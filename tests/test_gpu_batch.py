"""GPU tests of batched ensembles: problems of uneven grid sizes, levels, Reynolds
numbers and bodies advanced by one graph replay per step end in the same state as when
advanced alone (every graph replay is checked, over both parities of the step)."""
import numpy as np
import pytest

import cases
import immersedbodies as ib

pytestmark = pytest.mark.gpu

FIELDS = ("q", "Gamma", "psi", "nonlin1", "nonlin2", "Ftilde", "fb")


def _problems():
    return [cases.cylinder(nx=64, ny=64, nlevel=3, h=0.05, x0=(-1.6, -1.6), Re=100.0),
            cases.two_lines(),
            cases.no_body(),
            cases.ellipse_body(nx=96, ny=64, nlevel=2, h=0.05, x0=(-1.2, -1.6), Re=250.0),
            cases.ellipse_body(nx=128, ny=128, nlevel=4, h=0.03, x0=(-1.5, -1.92), Re=60.0,
                               b=0.35, aoa_deg=25.0)]


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_matches_solvers_advanced_alone(precision):
    probs = _problems()
    alone = [ib.create_solver(p, precision=precision) for p in probs]
    batched = [ib.create_solver(p, precision=precision) for p in probs]
    f_alone = []
    for s in alone:
        st = s.advance(12)
        f_alone.append(s.forces(st))
    f_alone = np.stack(f_alone, 1)  # steps x prob x 2
    batch = ib.Batch(batched)
    f1 = batch.advance(7)   # odd: the next call starts on the other graph
    f2 = batch.advance(5)
    f_b = np.concatenate([f1, f2])
    batch.close()
    for k, (a, b) in enumerate(zip(alone, batched)):
        assert a.state.index == b.state.index == 12
        for name in FIELDS:
            x, y = a.read(name), b.read(name)
            assert np.array_equal(x, y), (k, name, np.abs(x - y).max() if x.size else 0)
    assert np.array_equal(f_b, f_alone)


def test_batch_then_alone_and_parity_check():
    probs = _problems()[:3]
    ref = [ib.create_solver(p) for p in probs]
    for s in ref:
        s.advance(6)
    sol = [ib.create_solver(p) for p in probs]
    batch = ib.Batch(sol)
    batch.advance(3)
    for s in sol:  # every solver one step alone: parity stays common
        s.advance(1)
    batch.advance(2)
    for a, b in zip(ref, sol):
        for name in ("Gamma", "psi", "q"):
            assert np.array_equal(a.read(name), b.read(name))
    sol[0].advance(1)  # now out of step with the others
    with pytest.raises(ib._lib.IbpmError, match="out of step"):
        batch.advance(1)
    batch.close()


def test_batch_against_oracle():
    """One batched problem against the FP64 oracle (the batch adds nothing to the step)."""
    p = cases.ellipse_body(nx=96, ny=64, nlevel=2, h=0.05, x0=(-1.2, -1.6), Re=250.0)
    s = ib.create_solver(p)
    batch = ib.Batch([s, ib.create_solver(cases.no_body())])
    f = batch.advance(8)[:, 0]
    batch.close()
    orc = cases.oracle_for(p)
    fo = cases.oracle_advance(orc, p, 8)
    assert cases.relerr(f, fo) < 1e-9
    assert cases.relerr(s.read("Gamma"), orc.field("Gamma")) < 1e-10


def test_batch_with_box_updates_matches_alone():
    """Static problems with the box updates forced (the trial psi on the panels' box, the
    final vort2flux by box products, project_circ on the box) advanced in the batch's
    graphs end bitwise equal to the same solvers advanced alone."""
    probs = _problems()[:2] + _problems()[3:]
    alone = [ib.create_solver(p, box_update="force") for p in probs]
    batched = [ib.create_solver(p, box_update="force") for p in probs]
    assert all(s.box_update(1) is not None for s in alone)
    for s in alone:
        s.advance(6)
    batch = ib.Batch(batched)
    batch.advance(6)
    batch.close()
    for k, (a, b) in enumerate(zip(alone, batched)):
        for name in FIELDS:
            x, y = a.read(name), b.read(name)
            assert np.array_equal(x, y), (k, name)
