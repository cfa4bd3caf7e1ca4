"""The oracle's occultation decision at the cut, pinned without the oracle.

Satellites sit on the cut of chosen pixels (tests/cut_cases.py), so s.r - t is
a few ulps of |r| from 0 and the decision depends on the evaluation order.
The oracle's threshold must equal the canonical expression bit for bit and its
decision must be the canonical one for every such pair, computed in exact
rational arithmetic; a different order (no FMAs) must disagree on some of them,
which shows the cases are close enough to the cut to tell orders apart.
"""
import math

import numpy as np
import pytest

import oracle
from tests import cut_cases as cc


@pytest.mark.parametrize("orbit_radius", [cc.FIXED_ORBIT, 0.0])
def test_oracle_decision_at_the_cut_is_canonical(orbit_radius):
    rng = np.random.default_rng(17 if orbit_radius else 18)
    lmn = rng.standard_normal((3, 500))
    lmn /= np.linalg.norm(lmn, axis=0)
    pos, targets = cc.cut_satellites(lmn, 80, 4, rng, orbit_radius)
    n_diff_order = 0
    margins = []
    for j in range(pos.shape[0]):
        for i in range(pos.shape[1]):
            r = pos[j, i]
            s = lmn[:, targets[j, i]]
            t_or = oracle.mask_threshold(*r, cc.R, orbit_radius)
            assert t_or == cc.canonical_threshold(*r, cc.R, orbit_radius)
            dec, margin = cc.canonical_decision(s, r, cc.R, orbit_radius)
            one = s.reshape(3, 1)
            sat = r.reshape(3, 1)
            got = oracle.count_visible(*one, sat, sat, cc.R, orbit_radius)  # both satellites of the row are r
            assert got == int(dec), (j, i, margin)
            margins.append(abs(margin) / math.ulp(np.linalg.norm(r)))
            n_diff_order += int(cc.unfused_decision(s, r, cc.R, orbit_radius) != dec)
    margins = np.array(margins)
    print(f"orbit_radius={orbit_radius}: {margins.size} pairs on the cut, |s.r - t| up to {margins.max():.0f} ulp "
          f"of |r| (median {np.median(margins):.1f}); the unfused order decides {n_diff_order} of them differently")
    assert np.max(margins) <= 64
    assert n_diff_order > 0
