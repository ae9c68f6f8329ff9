"""Immersed-boundary solvers (Luo external, modified Fadlun) on the GPU against the oracle.

Bars: tag and ghost-node index arrays bit-exact; intercept / image points within
1e-12 absolute; per-step relative L2 of u and p within 1e-10; drag and lift within 1e-8.
"""
import numpy as np
import pytest
from helpers import pair, rel_l2

pytestmark = pytest.mark.gpu

TAGS = ("ghostTagsUV", "hybridTagsUV", "hybridTagsUV2", "ghostTagsP", "hybridTagsP")
GEOM = ("body_intercept_x", "body_intercept_y", "image_point_x", "image_point_y",
        "body_intercept_p_x", "body_intercept_p_y", "image_point_p_x", "image_point_p_y",
        "distance_from_intersection_to_node", "distance_between_nodes_at_IB",
        "distance_from_u_to_body", "distance_from_v_to_body", "uv")
TIGHT = {"velocityTol": 1e-12, "poissonTol": 1e-12}


@pytest.mark.parametrize("name", ["cylinderRe40_small_luo", "cylinderRe40_small_fadlun",
                                  "cylinderRe40_luo", "cylinderRe40_fadlun"])
def test_tags_bit_exact(name):
    g, o = pair(name, TIGHT)
    gi, oi = g.info(), o.info()
    for k in ("startI", "startJ", "numCellsX", "numCellsY"):
        assert gi[k] == oi[k], k
    for t in TAGS:
        a, b = g.get(t), o.get(t)
        assert np.array_equal(a, b), (t, np.nonzero(a != b)[0][:10])
    for f in GEOM:
        d = np.abs(g.get(f) - o.get(f)).max()
        assert d < 1e-12, (f, d)


@pytest.mark.parametrize("name", ["cylinderRe40_small_luo", "cylinderRe40_small_fadlun"])
def test_steps_match_oracle(name):
    g, o = pair(name, TIGHT)
    for k in range(10):
        g.step_time()
        o.step()
        for f in ("u", "p"):
            e = rel_l2(g.get(f), o.get(f))
            assert e < 1e-10, (k, f, e)
        gi, oi = g.info(), o.info()
        assert abs(gi["forceX"] - oi["forceX"]) < 1e-8 * max(1.0, abs(oi["forceX"])), (k, gi["forceX"], oi["forceX"])
        assert abs(gi["forceY"] - oi["forceY"]) < 1e-8 * max(1.0, abs(oi["forceY"])), (k, gi["forceY"], oi["forceY"])
