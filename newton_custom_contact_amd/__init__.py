"""MI355X-native rigid-contact step with Newton's host API.

``import newton_custom_contact_amd as newton`` provides the names the reference's rigid-contact
loop uses: ``ModelBuilder``, ``Model``, ``State``, ``Control``, ``Contacts``, ``CollisionPipeline``,
``GeoType`` and ``solvers.SolverXPBD`` (reference: newton/__init__.py, newton/solvers.py).  The contact
step itself runs in libnmc.so (HIP kernels for gfx950) through a ctypes C-ABI.
"""

from . import solvers
from ._types import Axis, BodyFlags, GeoType, JointType, ShapeFlags
from .builder import ModelBuilder
from .collide import CollisionPipeline
from .contacts import Contacts
from .model import Control, Model, State

__version__ = "0.1.0"

__all__ = [
    "Axis",
    "BodyFlags",
    "CollisionPipeline",
    "Contacts",
    "Control",
    "GeoType",
    "JointType",
    "Model",
    "ModelBuilder",
    "ShapeFlags",
    "State",
    "solvers",
]
