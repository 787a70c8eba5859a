"""Host-side checks of the batched-solve workload (no GPU): the random configurations are
reproducible, inside the stated ranges, and give the stated mesh sizes."""
import numpy as np


def test_draw_batch_params_ranges_and_determinism(sqhost):
    p = sqhost.draw_batch_params(256, seed=0)
    q = sqhost.draw_batch_params(256, seed=0)
    assert len(p) == 256
    for x, y in zip(p, q):
        assert x["a"] == y["a"] and x["b"] == y["b"] and np.array_equal(x["e"], y["e"])
    a = np.array([x["a"] for x in p])
    b = np.array([x["b"] for x in p])
    al = np.array([x["alpha"] for x in p])
    assert a.min() >= 0.1 and a.max() <= 0.7
    assert np.all(b >= 0) and np.all(b <= 0.8 * (1 - a))
    assert al.min() >= -5 and al.max() <= 5
    assert np.allclose([np.linalg.norm(x["e"]) for x in p], 1.0)
    assert {x["level"] for x in p} == {2, 3}
    assert {x["n_theta"] for x in p} == {16, 24, 32}
    assert all(x["n_z"] == 2 * x["n_theta"] and x["n_r"] == x["n_theta"] // 4 and x["L"] == 4.0 for x in p)
    # every sphere stays clear of the wall
    assert np.all(a + b < 1.0)


def test_batch_instances_share_uniform_tubes_and_dof_range(sqhost):
    p = sqhost.draw_batch_params(24, seed=3)
    inst = sqhost.batch_instances(p)
    assert len({id(q.tube) for q in inst}) == len({x["n_theta"] for x in p})
    dof = []
    for q, x in zip(inst, p):
        nth, nz, nr = x["n_theta"], x["n_z"], x["n_r"]
        assert q.tube.P() == 2 * nth * nz + 2 * nth * (2 * nr - 1)
        assert q.sphere.P() == 20 * 4 ** x["level"]
        assert np.allclose(q.sphere.centroid.mean(axis=0), q.center, atol=1e-12)
        dof.append(3 * (q.tube.P() + q.sphere.P()) + 6)
    assert min(dof) >= 4710 and max(dof) <= 19014
