"""The FP32 force loop's reciprocal rounds like IEEE division (R5f's 1.0f / q).

The kernel computes y0 = the hardware estimate of 1/q (within one ulp),
e = fmaf(-q, y0, 1), y1 = fmaf(y0, e, y0), and sets the last significand
bit when q's significand is all ones. This checks, for every float q of one
binade (the rule scales by powers of two, so one binade covers all normal q
with normal reciprocals), that the rule gives exactly 1.0f / q for BOTH
faithful estimates: the float just below 1/q and the one just above. So the
result does not depend on which of them the hardware returns. The fused
operations are evaluated exactly: q y0 has 48 bits and is exact in FP64;
1 - q y0 is exact by Sterbenz's lemma and fits 24 bits; y0 e is exact in
FP64 and the sum y0 + y0 e is carried as an exact two-term sum before its
one rounding to FP32.
"""
import numpy as np


def _round_f32_of_sum(s, err):
    """Round-to-nearest-even FP32 of the exact value s + err (s = RN64 of it,
    |err| <= ulp64(s) / 2)."""
    r = s.astype(np.float32)
    r64 = r.astype(np.float64)
    up = np.nextafter(r, np.float32(np.inf)).astype(np.float64)
    down = np.nextafter(r, np.float32(0)).astype(np.float64)
    # s on a midpoint between r and a neighbour: the rounding of s to FP32
    # went to the even one; the exact value decides by the sign of err
    mid_up = (s - r64) == (up - r64) / 2.0
    mid_down = (r64 - s) == (r64 - down) / 2.0
    out = r.copy()
    out = np.where(mid_up & (err > 0), up.astype(np.float32), out)
    out = np.where(mid_up & (err < 0), r, out)
    out = np.where(mid_down & (err < 0), down.astype(np.float32), out)
    out = np.where(mid_down & (err > 0), r, out)
    # exactly on a midpoint (err == 0): ties to even, which astype did
    return out


def rule(q, y0):
    q64 = q.astype(np.float64)
    y64 = y0.astype(np.float64)
    e64 = 1.0 - q64 * y64                      # exact
    e = e64.astype(np.float32)
    assert np.array_equal(e.astype(np.float64), e64), "1 - q y0 must be exact in FP32"
    p = y64 * e.astype(np.float64)             # exact (24 x 24 bits)
    s = y64 + p
    err = p - (s - y64)                        # Fast2Sum: |y0| >= |p|
    y1 = _round_f32_of_sum(s, err)
    m = q.view(np.uint32) & np.uint32(0x7FFFFF)
    bits = y1.view(np.uint32) | ((m + np.uint32(1)) >> np.uint32(23))
    return bits.view(np.float32)


def test_reciprocal_rule_is_correctly_rounded_for_any_faithful_estimate():
    for exponent in (127, 126):  # q in [1, 2) and [0.5, 1)
        q = ((np.uint32(exponent) << np.uint32(23)) | np.arange(1 << 23, dtype=np.uint32)).view(np.float32)
        want = np.float32(1.0) / q                 # IEEE division
        exact_above = want.astype(np.float64) > 1.0 / q.astype(np.float64)
        exact_equal = want.astype(np.float64) == 1.0 / q.astype(np.float64)
        below = np.where(exact_above, np.nextafter(want, np.float32(0)), want)
        above = np.where(exact_above | exact_equal, want, np.nextafter(want, np.float32(np.inf)))
        for y0 in (below, above):
            got = rule(q, y0)
            bad = np.flatnonzero(got.view(np.uint32) != want.view(np.uint32))
            assert bad.size == 0, f"{bad.size} wrong, first q = {q[bad[:4]]!r}"


def test_the_all_ones_case_needs_the_fix():
    # without the last-bit fix the estimate just below 1/q gives the tie
    # and rounds the wrong way: the case the rule exists for
    q = np.array([np.nextafter(np.float32(2.0), np.float32(0))], dtype=np.float32)
    y0 = np.array([0.5], dtype=np.float32)
    want = np.float32(1.0) / q
    assert rule(q, y0)[0] == want[0]
    e = np.float32(1.0 - float(q[0]) * 0.5)
    y1_plain = np.float32(0.5 + 0.5 * float(e))
    assert y1_plain != want[0]
