"""HIP solver methods; importing this package registers them."""
from . import rk4_hip  # noqa: F401
from . import split_step_hip  # noqa: F401
from . import lawson_hip  # noqa: F401
