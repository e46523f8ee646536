"""The feature world (tests/features.py) reaches the branches the benchmark worlds never take,
counted by the instrumented oracle, so that the device parity run on it (test_gpu_features.py)
covers them; and the walk that stops at an empty octant renders it as the whole walk does."""

import numpy as np

import features


def test_the_feature_world_reaches_every_rare_branch(oracle):
    fr = features.feature_frame(oracle)
    img, c = oracle.render_hq_counted(fr.view(), 4)
    counts = {k: getattr(c, k) for k in ("see_through_drawn", "see_through_forced",
                                         "translucent_passes", "alpha_zero_rejects",
                                         "shadow_water_exits", "shadow_translucent",
                                         "water_interfaces")}
    print(counts)
    assert all(v >= 100 for v in counts.values()), counts
    assert c.see_through_drawn > c.see_through_forced  # opaque see-through quads drawn too
    whole, cw = oracle.render_hq_counted(fr.view(), 4, device_walk=False)
    assert np.array_equal(whole.view("<u4"), img.view("<u4"))
    assert cw.dda_steps >= c.dda_steps and cw.traces == c.traces


def test_the_feature_world_has_its_see_through_range(oracle):
    fr = features.feature_frame(oracle)
    n = len(fr.scene.quads)
    assert tuple(fr.params["see_through"][:2]) == (n - features.SEE_THROUGH, features.SEE_THROUGH)
    assert fr.params["sun_color"][:3].min() > 0  # the sun is lit
