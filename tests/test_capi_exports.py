"""CPU checks of the C ABI library: it loads, exports every function that
include/pmpm.h declares, and rejects invalid arguments before touching the
device. No compute call is made (there is no GPU here)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "pmpm.h")
LIB = os.path.join(ROOT, "pyroclastmpm_amd", "libpmpm_hip.so")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(pmpm_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_entry_points():
    names = declared_functions()
    for required in ["pmpm_partition", "pmpm_shape_function", "pmpm_p2g", "pmpm_grid_bc", "pmpm_g2p",
                     "pmpm_musl_map", "pmpm_musl_g2p", "pmpm_stress_update", "pmpm_step",
                     "pmpm_rigid_levelset"]:
        assert required in names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_ctypes_prototypes_cover_header():
    from pyroclastmpm_amd import capi
    assert set(declared_functions()) == set(capi.PROTOTYPES)


def test_version_string():
    from pyroclastmpm_amd import capi
    assert b"gfx950" in capi.lib().pmpm_version()


def test_invalid_arguments_rejected_on_host():
    from pyroclastmpm_amd import capi
    L = capi.lib()
    g = capi.make_grid([0, 0, 0], [1, 1, 1], 0.1, 3)
    # bad dimension / shape function: rejected before any launch
    assert L.pmpm_shape_function(5, 0, 1, ctypes.byref(g), None, 10, None, None, None) == capi.PMPM_ERR_INVALID
    assert L.pmpm_shape_function(3, 7, 1, ctypes.byref(g), None, 10, None, None, None) == capi.PMPM_ERR_INVALID
    prm = capi.Params()
    prm.dim = 4
    assert L.pmpm_grid_bc(ctypes.byref(prm), ctypes.byref(g), None, 0, 0, 0, None) == capi.PMPM_ERR_INVALID
    prm.dim = 3
    prm.num_bcs = 2
    nodes = capi.Nodes()
    assert L.pmpm_grid_bc(ctypes.byref(prm), ctypes.byref(g), ctypes.byref(nodes), 0, 3, 0, None) == \
        capi.PMPM_ERR_INVALID
    assert L.pmpm_permute(5, 1, 2, None, None, None, 0, None) == capi.PMPM_ERR_INVALID
    assert L.pmpm_rigid_advance(4, 1, None, None, None, None) == capi.PMPM_ERR_INVALID


def test_grid_matches_reference_construction():
    from pyroclastmpm_amd import capi
    # the 2 x 1 x 1 m scoop grid at h = 0.0125 (computed in FP32 as the reference)
    g = capi.make_grid([0, 0, 0], [2.0, 1.0, 1.0], 0.0125, 3)
    assert list(g.num_cells) == [161, 81, 81]
    g = capi.make_grid([0, 0, 0], [1.0, 0.5, 0.5], 0.01, 3)
    assert list(g.num_cells) == [101, 51, 51] and g.num_cells_total == 262701


def test_package_import_has_reference_api():
    import pyroclastmpm_amd.MPM2D as pm2
    import pyroclastmpm_amd.MPM3D as pm3
    for name in ["ParticlesContainer", "NodesContainer", "USL", "MUSL", "Solver", "LinearElastic", "NewtonFluid",
                 "LocalGranularRheology", "DP_rheo", "Gravity", "NodeDomain", "RigidBodyLevelSet", "set_globals",
                 "set_global_shapefunction", "set_global_timestep", "set_global_step",
                 "set_global_output_directory"]:
        assert hasattr(pm3, name), name
        assert hasattr(pm2, name), name
    assert pm3.global_dimension == 3 and pm2.global_dimension == 2
    with pytest.raises(ValueError):
        pm3.set_global_shapefunction("quartic")


def test_reference_import_path():
    """`import pyroclastmpm.MPM3D as pm` and `from pyroclastmpm import USL`
    (the reference's module path) reach the same classes."""
    import pyroclastmpm
    import pyroclastmpm.MPM2D as pm2
    import pyroclastmpm_amd.MPM2D as amd2
    from pyroclastmpm import USL, global_dimension
    from pyroclastmpm_amd import MPM3D

    assert pm2.USL is amd2.USL and pm2.global_dimension == 2
    assert USL is MPM3D.USL and global_dimension == 3
    assert pyroclastmpm.MPM1D.global_dimension == 1
