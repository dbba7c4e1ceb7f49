"""MI355X-native engine for PENTrack's particle-tracking hot path.

Host-side mirror of the reference interfaces (TGeometry, TFieldManager,
TParticle species, TTracker) over the C ABI of libpentrack_gpu.so.
"""
from . import constants
from .geometry import Material, Solid, TGeometry, read_stl, write_stl, box_mesh, cylinder_mesh, hollow_box_mesh
from .fields import FieldContainer, TFieldManager, tabfield3, uniform_efield, loop_field
from .particle import SpeciesConfig, make_states, initial_velocity, kinetic_energy
from .tracker import TTracker, Batch, LaunchConfig

__all__ = ["constants", "Material", "Solid", "TGeometry", "read_stl", "write_stl", "box_mesh", "cylinder_mesh",
           "hollow_box_mesh", "FieldContainer", "TFieldManager", "tabfield3", "uniform_efield", "loop_field",
           "SpeciesConfig", "make_states", "initial_velocity", "kinetic_energy", "TTracker", "Batch", "LaunchConfig"]
