"""The site-index quotient of the element passes' shape functions.

The device splits a site index s = (i * n1 + j) * n2 + k with a float
reciprocal estimate of s / d and one integer correction (site_div in
k_pair.hip) instead of a 32-bit integer division.  The host accepts fewer
than 65536 sites per element, so s < 2^16.  The hardware reciprocal
(v_rcp_f32) is within one unit in the last place of 1/d, so the estimate is
taken with the correctly rounded reciprocal and both of its neighbours;
every quotient must equal floor(s / d) for every s below 2^16, for all
divisors up to 2048 and a sample of larger ones.
"""
import numpy as np


def site_div(s, d, rcp):
    q = (s.astype(np.float32) * rcp).astype(np.int64)  # f32 product, truncated
    r = s - q * d
    return q + (r >= d) - (r < 0)


def test_site_div_exact_below_2_16():
    s = np.arange(65536, dtype=np.int64)
    ds = list(range(1, 2049)) + [int(d) for d in np.random.default_rng(7).integers(2049, 65536, 300)]
    for d in ds:
        r0 = np.float32(1.0) / np.float32(d)
        for rcp in (r0, np.nextafter(r0, np.float32(0)), np.nextafter(r0, np.float32(1))):
            assert np.array_equal(site_div(s, d, rcp), s // d), (d, rcp)
