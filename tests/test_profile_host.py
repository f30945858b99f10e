"""Host-side profile analysis (lennard_jones.rs:158-271), checked against the
reference's own unit test (lennard_jones.rs:1447-1497) and against closed
forms for the pressure tensor and surface tension. No GPU needed."""
import math

import pytest


def profile(lj, density):
    """A profile with the given density in each slab of a 1 x 1 x bins box
    (lennard_jones.rs:1447-1466)."""
    n = len(density)
    slabs = [lj.Slab(b + 0.5, rho, [0.0] * 3, [0.0] * 3, 0.0) for b, rho in enumerate(density)]
    return lj.Profile(1, [1.0, 1.0, float(n)], slabs, [0.0] * 3)


def test_phases_of_two_slabs_across_the_boundary(lj):
    # lennard_jones.rs:1468-1497: liquid (0.7) in z 0..10, 50..60 and 25..40,
    # vapour (0.02) elsewhere
    density = [0.7 if (b < 10 or b >= 50 or 25 <= b < 40) else 0.02 for b in range(60)]
    p = profile(lj, density)
    assert sorted(p.interfaces(0.36)) == [10.0, 25.0, 40.0, 50.0]
    phases = p.phases(0.36, 4.0)
    count = lambda ph: sum(1 for q in phases if q == ph)
    assert count(lj.Phase.LIQUID) == 6 + 6 + 7
    assert abs(p.mean_in(density, phases, lj.Phase.LIQUID) - 0.7) < 1e-12
    assert abs(p.mean_in(density, phases, lj.Phase.VAPOUR) - 0.02) < 1e-12
    assert count(lj.Phase.INTERFACE) == 4 * 4
    assert all(q == lj.Phase.BETWEEN for q in profile(lj, [0.3] * 8).phases(0.36, 4.0))


def test_mean_in_an_absent_phase_is_none(lj):
    p = profile(lj, [0.3] * 8)
    phases = p.phases(0.36, 4.0)
    assert p.mean_in(p.density(), phases, lj.Phase.LIQUID) is None


def test_interface_interpolates_between_slab_centres(lj):
    # Density 0.1 then 0.5 in 8 slabs of a 1 x 1 x 8 box: crossing 0.3 half
    # way between the centres 3.5 and 4.5, and again across the boundary
    p = profile(lj, [0.1] * 4 + [0.5] * 4)
    assert sorted(p.interfaces(0.3)) == [0.0, 4.0]


def test_pressure_tensor_and_surface_tension(lj):
    # Two slabs of a 2 x 3 x 4 box (slab volume 12): kinetic and virial sums
    # chosen so every tensor component is a known number
    slabs = [lj.Slab(1.0, 10.0, [12.0, 24.0, 36.0], [12.0, 0.0, -12.0], -5.0),
             lj.Slab(3.0, 20.0, [24.0, 24.0, 24.0], [0.0, 12.0, 24.0], -7.0)]
    p = lj.Profile(2, [2.0, 3.0, 4.0], slabs, [6.0, 6.0, 30.0])
    assert p.slab_volume() == 12.0
    assert p.density() == [10.0 / 12.0, 20.0 / 12.0]
    assert p.temperature() == [72.0 / 30.0, 72.0 / 60.0]
    assert p.pressure_tensor() == [[2.0, 2.0, 2.0], [2.0, 3.0, 4.0]]
    pxx, pyy, pzz = p.total_pressure_tensor()
    assert (pxx, pyy, pzz) == (42.0 / 24.0, 54.0 / 24.0, 90.0 / 24.0)
    assert p.surface_tension() == pytest.approx(0.5 * 4.0 * (pzz - 0.5 * (pxx + pyy)))


def test_profile_from_sums_layout(lj):
    # vlj_read_profile's layout: 8 sums per slab, then the whole-box W_aa
    bins = 3
    sums = [float(k) for k in range(bins * 8 + 3)]
    p = lj.Profile.from_sums(sums, 5, [1.0, 2.0, 6.0])
    assert p.samples == 5 and len(p.slabs) == 3
    assert [s.z for s in p.slabs] == [1.0, 3.0, 5.0]
    s1 = p.slabs[1]
    assert (s1.particles, s1.kinetic, s1.virial, s1.potential) == (8.0, [9.0, 10.0, 11.0], [12.0, 13.0, 14.0], 15.0)
    assert p.virial == [24.0, 25.0, 26.0]
    assert math.isnan(lj.Profile.from_sums([0.0] * 11, 0, [1.0, 1.0, 1.0]).temperature()[0])
