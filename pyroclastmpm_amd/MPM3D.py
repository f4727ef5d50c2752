"""3-D module of the PyroclastMPM API (the reference's `pyroclastmpm.MPM3D`)."""
from ._core import MPM3D as _m

globals().update({k: getattr(_m, k) for k in dir(_m) if not k.startswith("__")})
del _m
