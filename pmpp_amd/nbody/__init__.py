from .particles import Particles  # noqa: F401
from .gravity import gravity  # noqa: F401
