"""CPU checks of the preconditions the ct scan relies on (no GPU).

The single-frame ct scan converts each step d to 32.32 fixed point as
d * 2^32 + 2^52 minus the bits of 2^52 when sigma_s / sigma_r is at most
kFxSmallMaxQ (csrc/dts_kernels.h), instead of __double2ll_rn.  Both round
to nearest, ties to even, as long as d * 2^32 < 2^52.  These tests read the
constants from the header, so an edit that breaks the bound fails here."""
import re
from pathlib import Path

import numpy as np

HEADER = Path(__file__).resolve().parents[1] / "domain_transform_solver_amd" / "csrc" / \
    "dts_kernels.h"


def _constant(name):
    m = re.search(rf"constexpr\s+(?:double|float)\s+{name}\s*=\s*([0-9.eE+-]+)f?;",
                  HEADER.read_text())
    assert m, f"{name} not found in {HEADER.name}"
    return float(m.group(1))


def test_largest_step_below_2_pow_20():
    q_max = _constant("kFxSmallMaxQ")
    inv255sq = np.float64(np.float32(_constant("kINV255SQ_F")))
    # The largest sum of squared channel differences, 4 * 255^2, in the
    # kernel's order of operations (double, contraction off).
    t = np.float64(4 * 255 * 255) * inv255sq
    t = t * np.float64(q_max)
    t = t * np.float64(q_max)
    d_max = np.sqrt(t + 1.0)
    assert d_max * 2.0 ** 32 < 2.0 ** 52, d_max
    # Larger q would break it: the bound is not far from tight.
    assert np.sqrt(4.0 * (2 * q_max) ** 2 + 1.0) * 2.0 ** 32 >= 2.0 ** 52


def test_2_pow_52_conversion_rounds_like_rint():
    rng = np.random.default_rng(7)
    k52 = np.float64(2.0 ** 52)
    d = np.concatenate([
        1.0 + rng.random(200_000) * (2.0 ** 20 - 1.0),
        1.0 + rng.random(100_000) * 4.0,
        1.0 + (np.arange(1, 50_000) + 0.5) / 2.0 ** 32,  # exact ties
        [np.nextafter(2.0 ** 20, 0.0), 1.0],
    ])
    x = d * 2.0 ** 32
    fast = (x + k52).view(np.int64) - k52.view(np.int64)
    np.testing.assert_array_equal(fast, np.rint(x).astype(np.int64))


def test_huber_threshold_float_compare():
    """The HFBS solve kernel clamps when the float |grad| >= 1.345f; the
    reference tests the float against the double 1.345 (src/hfbs.cu:233).
    Both select the same floats: 1.345f is above 1.345 and the float just
    below it is under."""
    t = np.float32(1.345)
    assert float(t) > 1.345
    assert float(np.nextafter(t, np.float32(0))) < 1.345
