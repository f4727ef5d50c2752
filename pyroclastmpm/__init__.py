"""The reference's import path, served by the MI355X engine.

Reference scripts import ``pyroclastmpm.MPM1D`` / ``MPM2D`` / ``MPM3D``
(python/pybind/_pyroclastmpm.cpp:32-40) or names from ``pyroclastmpm``
itself; both resolve here to the modules of ``pyroclastmpm_amd`` (the
package's top level carries the 3-D module's names, with
``global_dimension``), so such a script runs unchanged.
"""
import sys as _sys

from pyroclastmpm_amd import MPM1D, MPM2D, MPM3D  # noqa: F401

for _name, _mod in (("MPM1D", MPM1D), ("MPM2D", MPM2D), ("MPM3D", MPM3D)):
    _sys.modules[__name__ + "." + _name] = _mod

globals().update({_k: getattr(MPM3D, _k) for _k in dir(MPM3D) if not _k.startswith("_")})
