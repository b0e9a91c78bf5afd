"""HIP solver methods; importing this package registers them."""
from . import rk4_hip  # noqa: F401
