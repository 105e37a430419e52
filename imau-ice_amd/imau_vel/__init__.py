"""imau_vel -- MI355X-native ice-velocity path of IMAU-ICE (DIVA / SSA Picard solve + SIA).

Host-side mirror of ice_velocity_module's call surface over the C ABI of
libimauvel.so (include/imau_vel.h).
"""
from . import abi, config, geometry, state  # noqa: F401
from .config import default_config, set_config  # noqa: F401
from .state import Grid, IceState  # noqa: F401
