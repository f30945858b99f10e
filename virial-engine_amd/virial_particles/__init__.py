"""MI355X-native `virial_particles`: the Lennard-Jones path only."""
from . import lennard_jones

__all__ = ["lennard_jones"]
