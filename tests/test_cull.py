"""CPU tests of the blend kernels' wave-level culling test (gs_cull.h).

Skipping a tile-list entry for a whole pixel block is only allowed when every
pixel of the block would skip it anyway, so the test is checked against brute
force over random (often needle-thin) conics, opacities and block positions,
half of them with a border pixel on the power = thr contour.
"""
import oracle as O


def test_culling_never_skips_a_contributing_entry():
    for seed in (11, 21):
        cases, culled, bad, allskip = O.cull_check(2_000_000, seed=seed)
        assert bad == 0
        assert culled <= allskip
        # the test is useful: it removes nearly every entry brute force can
        assert culled > 0.95 * allskip


def test_culling_margin_is_needed():
    # with no rounding margin the boundary cases do fail: the brute force sees
    # rounding-level errors, and the margin is what covers them
    assert O.cull_check(2_000_000, seed=13, margin=0.0)[2] > 0
    assert O.cull_check(300_000, seed=12, margin=-2e-2)[2] > 0
