"""MI355X-native explicit MPM engine with PyroclastMPM's API.

    import pyroclastmpm_amd.MPM3D as pm
    pm.set_globals(1e-5, 27, "cubic", "./output/")
    nodes = pm.NodesContainer([0, 0, 0], [1.0, 0.5, 0.5], 0.01)
    ...

The classes live in the compiled extension ``_core`` (C++ host objects over
the HIP library ``libpmpm_hip.so``). There is no CPU fallback: importing the
package without the built extension raises ImportError.
"""
import os as _os

_HERE = _os.path.dirname(_os.path.abspath(__file__))

# One HIP runtime per process: torch ships its own libamdhip64 and the
# extension links the system one under the same soname. Loaded in this order
# the extension binds to torch's copy; loaded the other way round the process
# holds two runtimes and the one initialised second finds no device.
try:
    import torch as _torch  # noqa: F401
except ImportError:  # the engine itself does not need torch
    _torch = None

try:
    from . import _core  # noqa: F401
except ImportError as e:  # pragma: no cover - build problem, not a runtime branch
    raise ImportError(
        "pyroclastmpm_amd: the compiled extension is missing; build it with "
        "`python pyroclastmpm_amd/build.py` (needs hipcc for gfx950)") from e

LIBRARY_PATH = _os.path.join(_HERE, "libpmpm_hip.so")
HEADER_PATH = _os.path.join(_os.path.dirname(_HERE), "include", "pmpm.h")
__version__ = _core.version


def device_count() -> int:
    """Number of HIP devices visible to the extension."""
    return _core.device_count()
