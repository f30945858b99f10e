"""Renderer bindings (lennard_jones.rs:1041-1051, 885-889;
lennard_jones.wgsl:596-603): the float4 render arrays hold the wrapped
positions and the velocities in slot order, with w the neighbours within
1.5 sigma or the local temperature, both pinned by brute force over all
pairs."""
import numpy as np
import pytest

from helpers import sc_liquid

pytestmark = pytest.mark.gpu


def all_pairs(pos, vel, box, rc):
    d = pos[None, :, :] - pos[:, None, :]
    d -= box * np.round(d / box)
    r2 = (d * d).sum(-1)
    np.fill_diagonal(r2, np.inf)
    within = r2 < rc * rc
    near = (within & (r2 < 2.25)).sum(1)
    v2 = (vel * vel).sum(1)
    temp = (v2 + (within * v2[None, :]).sum(1)) / (3.0 * (within.sum(1) + 1))
    return near, temp


@pytest.mark.parametrize("shape", ["single", "batch"])
def test_render_arrays(oracle, lj, shape):
    pos, vel, box = sc_liquid(oracle, n_side=9)
    pos = oracle.jitter(pos, 5, 0.3)
    st = lj.Settings([box] * 3, 2.5, 0.004)
    if shape == "single":
        sim = lj.LennardJones((pos, vel), st)
    else:
        pos2, vel2, box2 = sc_liquid(oracle, n_side=7, density=0.5)
        sim = lj.Batch([((pos, vel), st), ((pos2, vel2), lj.Settings([box2] * 3, 2.5, 0.004))])
    sim.run(40)
    p, v = sim.read_render()
    if shape == "single":
        dp, dv, _, _ = sim.debug_state()
        L = sim.box_size()[0]
        n = dp.shape[0]
    else:
        a = sim.read_particle_array()  # original order; the render arrays are in slot order
        n = 729
        dp, dv = None, None
    if shape == "single":
        np.testing.assert_array_equal(p[:, :3], np.mod(dp, L).astype(np.float32))
        assert (p[:, 3] == 0).all()
        np.testing.assert_array_equal(v[:, :3], dv.astype(np.float32))
        near, _ = all_pairs(dp, dv, L, 2.5)
        np.testing.assert_array_equal(v[:, 3], near.astype(np.float32))
        sim.set_drawn(lj.Drawn.TEMPERATURE)
        _, v = sim.read_render()
        _, temp = all_pairs(dp, dv, L, 2.5)
        np.testing.assert_allclose(v[:, 3], temp, rtol=1e-6)
        with pytest.raises(lj.LennardJonesError, match="drawn is 0"):
            sim.set_drawn(2)
        ptr_p, ptr_v = sim.render_buffers()
        assert ptr_p and ptr_v and ptr_p != ptr_v
    else:
        # every particle of the batch appears once: the multiset of rendered
        # velocities equals the velocities read back, in f32
        got = np.sort(v[:, :3].astype(np.float64).view([("x", "f8"), ("y", "f8"), ("z", "f8")]).ravel())
        want = np.sort(a["vel"].astype(np.float32).astype(np.float64).view(
            [("x", "f8"), ("y", "f8"), ("z", "f8")]).ravel())
        np.testing.assert_array_equal(got, want)
        assert p.shape == (a.shape[0], 4)
