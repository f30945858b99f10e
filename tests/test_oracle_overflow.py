"""List overflow in the oracle (lennard_jones.wgsl:494-497): past the
capacity a list keeps its first `capacity` entries, in order, and the
measurement reports the longest list found."""
import numpy as np

from helpers import jittered_lattice


def test_overflowing_lists_keep_their_first_entries(oracle):
    pos, vel = jittered_lattice(oracle)
    full = oracle.Sim(pos, vel, [11.0] * 3)
    assert full.measure().list_overflow == 0
    fc, fe = full.lists()
    with oracle.capacity_override(40):
        cut = oracle.Sim(pos, vel, [11.0] * 3)
    m = cut.measure()
    assert cut.params.capacity == 40
    cc, ce = cut.lists()
    np.testing.assert_array_equal(cc, np.minimum(fc, 40))
    for i in range(len(cc)):
        np.testing.assert_array_equal(ce[i, : cc[i]], fe[i, : cc[i]])
    assert m.list_overflow == fc.max() > 40
