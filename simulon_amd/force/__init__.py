"""Force fields with the reference's backbone interface (forward() -> dict)."""
from .lennard_jones_force import LennardJonesForce, LennardJonesForceCUDA  # noqa: F401
from .eam_force import EAMForce, EAMForceCUDA, EAMForceCUDAExt  # noqa: F401
from .mlp_force import MLPForce  # noqa: F401
