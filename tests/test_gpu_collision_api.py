"""GPU checks of the rest of CollisionDetection's interface (bvh.h:44-89):
SetBuildType / GetBuildType and UpdateQueries, against the all-pairs oracle.

Bars: query lists bit-exact for every build type; UpdateQueries' query set
and distance types equal to the oracle's, squared distances within 1e-12
relative (the oracle's distances come from another derivation, forward jets),
the order by (distance, type)."""
import numpy as np
import pytest

import oracle as O
from sbs_amd import capi, mesh
from sbs_amd.pd import CollisionDetection
from test_gpu_collision import compare_detect, two_bodies

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("build_type", [0, 1, 2])
@pytest.mark.parametrize("precision", [32, 64])
def test_every_build_type_gives_the_same_queries(gpu, build_type, precision):
    s, X, XT = two_bodies(n=4, gap=0.004, drop=0.05, seed=11)
    fp64 = precision == 64
    cd = CollisionDetection(s.nv, s.tri, s.tri_fathers, precision)
    try:
        assert cd.GetBuildType() == 1
        cd.SetBuildType(build_type)
        assert cd.GetBuildType() == build_type
        tI, _ = cd.DetectCollision(X, XT, True)
        q, toi, _ = cd.GetQueries()
        with pytest.raises(capi.SbsError):
            cd.SetBuildType(3)
    finally:
        cd.close()
    r = O.collision_detect(X, XT, s.tri, s.tri_fathers, ignore_self=True, fp64=fp64)
    np.testing.assert_array_equal(q, r["queries"])
    np.testing.assert_array_equal(toi, r["toi"])
    np.testing.assert_array_equal(tI, r["tI"])
    assert len(q) > 0


def _compare_proximity(q, t, d, qo, to, do_):
    assert len(q) == len(qo) > 0
    # order by (d, type): the GPU's list is sorted that way
    assert np.all(np.diff(d) >= 0)
    kg = {tuple(map(int, row)): (int(tt), float(dd)) for row, tt, dd in zip(q, t, d)}
    ko = {tuple(map(int, row)): (int(tt), float(dd)) for row, tt, dd in zip(qo, to, do_)}
    assert kg.keys() == ko.keys()
    for k in kg:
        assert kg[k][0] == ko[k][0], k
        assert abs(kg[k][1] - ko[k][1]) <= 1e-12 * ko[k][1], (k, kg[k], ko[k])


def test_update_queries_matches_oracle(gpu):
    s = mesh.stack_scene(4, gap=6e-4, jitter=2e-4, seed=3)
    dhat = 1e-3
    cd = CollisionDetection(s.nv, s.tri, s.tri_fathers, 64)
    try:
        # before any detect call the ignore-self flag is false: self pairs too
        q, t, d = cd.UpdateQueries(s.X, dhat)
        qo, to, do_ = O.update_queries(s.X, s.tri, s.tri_fathers, dhat, ignore_self=False)
        _compare_proximity(q, t, d, qo, to, do_)
        cd.DetectCollision(s.X, s.X, True)
        q2, t2, d2 = cd.UpdateQueries(s.X, dhat)
        qo2, to2, do2 = O.update_queries(s.X, s.tri, s.tri_fathers, dhat, ignore_self=True)
        _compare_proximity(q2, t2, d2, qo2, to2, do2)
    finally:
        cd.close()


def test_update_queries_needs_fp64(gpu):
    s = mesh.stack_scene(2, gap=6e-4)
    cd = CollisionDetection(s.nv, s.tri, s.tri_fathers, 32)
    try:
        with pytest.raises(capi.SbsError):
            cd.UpdateQueries(s.X, 1e-3)
    finally:
        cd.close()
