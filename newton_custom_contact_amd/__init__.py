"""MI355X-native rigid-contact step with Newton's host API.

``import newton_custom_contact_amd as newton`` provides the names the reference's rigid-contact
loop uses: ``ModelBuilder``, ``Model``, ``State``, ``Control``, ``Contacts``, ``CollisionPipeline``,
``GeoType`` and ``solvers.SolverXPBD`` (reference: newton/__init__.py, newton/solvers.py).
"""

from ._types import Axis, BodyFlags, GeoType, JointType, ShapeFlags
from .builder import ModelBuilder
from .model import Control, Model, State

__version__ = "0.1.0"

__all__ = [
    "Axis",
    "BodyFlags",
    "Control",
    "GeoType",
    "JointType",
    "Model",
    "ModelBuilder",
    "ShapeFlags",
    "State",
]
