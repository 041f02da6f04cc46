"""Pins the oracle's moreDBC handling (dragged vertices: setMDt_2MoreDBC,
computeSn, getErrorKern, updateVelPos; pdUtil.cu:56-95,195-214, 180-193)
by what it must do, not by its formulas: in Jacobi mode a dragged vertex's
prediction is its target and getErrorKern keeps it there, so after a step it
sits exactly at the target with zero velocity; its neighbours are pulled
towards the target; with zero weights the step is the plain one."""
import numpy as np

import cheb_case as C
import oracle as O
from sbs_amd import mesh


def setup(fp64=True):
    s = mesh.cube_scene(5)
    rng = np.random.default_rng(3)
    sel = rng.choice(np.flatnonzero(s.dbc == 0), 6, replace=False)
    more = np.zeros(s.nv)
    more[sel] = 5e3
    targets = s.X + np.array([0.04, -0.02, 0.03])
    p = O.make_params(solver=O.SOLVER_JACOBI, num_iterations=20, rho=C.RHO)
    return s, sel, more, targets, p


def test_dragged_vertices_sit_at_their_targets():
    s, sel, more, targets, p = setup()
    o = O.PdOracle(s, p, fp64=True)
    o.step(1)
    o.set_more_dbc(more, targets)
    for _ in range(3):
        o.step(1)
        X, V = o.get("X"), o.get("V")
        np.testing.assert_allclose(X[sel], targets[sel], rtol=0, atol=1e-12)
        assert np.abs(V[sel]).max() == 0.0
    o.close()


def test_neighbours_are_pulled_towards_the_targets():
    s, sel, more, targets, p = setup()
    plain = O.PdOracle(s, p, fp64=True)
    drag = O.PdOracle(s, p, fp64=True)
    for o in (plain, drag):
        o.step(1)
    drag.set_more_dbc(more, targets)
    plain.step(3)
    drag.step(3)
    Xp, Xd = plain.get("X"), drag.get("X")
    shift = np.array([0.04, -0.02, 0.03])
    # the vertices sharing a tet with a dragged one move along the drag
    nb = np.unique(s.tet[np.isin(s.tet, sel).any(axis=1)])
    nb = np.setdiff1d(nb, np.concatenate([sel, np.flatnonzero(s.dbc != 0)]))
    assert len(nb) > 0
    assert np.all((Xd[nb] - Xp[nb]) @ shift > 0)
    plain.close()
    drag.close()


def test_zero_weights_give_the_plain_step():
    s, sel, more, targets, p = setup()
    a = O.PdOracle(s, p, fp64=True)
    b = O.PdOracle(s, p, fp64=True)
    a.step(1)
    b.step(1)
    b.set_more_dbc(np.zeros(s.nv), targets)
    a.step(2)
    b.step(2)
    np.testing.assert_array_equal(a.get("X"), b.get("X"))
    a.close()
    b.close()
