"""MI355X-native XPBD cloth substep loop (HIP / gfx950).

The product is the C-ABI shared library libnumi_cloth_hip.so built from
csrc/ by build.py; `cloth` mirrors that ABI for Python callers.
"""

from . import cloth  # noqa: F401
from .cloth import (  # noqa: F401
    ABI_VERSION, CONFIG, PARTICLE, DISTANCE, GRIP, KNOT, BEND, FRUIT, FRUIT_PAIR, YARN_CONTACT, BATCH,
    Batch, ClothError, GripTrajectory, Scene, Solver, library, load_material, material_defaults, state_hash,
)
