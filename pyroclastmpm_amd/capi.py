"""ctypes mirror of the C ABI in include/pmpm.h (libpmpm_hip.so).

Used by the differentiable and batched front ends and by the tests that call
the library directly. Device buffers are passed as raw pointers (for example
``tensor.data_ptr()`` of a torch tensor on the HIP device).
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpmpm_hip.so")

PMPM_OK = 0
PMPM_ERR_INVALID = -1
PMPM_ERR_HIP = -2
PMPM_ERR_WORKSPACE = -3
PMPM_ERR_UNSUPPORTED = -4
NOALLOC_SCRATCH_BYTES = 64 << 20  # PMPM_NOALLOC_SCRATCH_BYTES

SF = {"linear": 0, "quadratic": 1, "cubic": 2}
MAT_LINEAR_ELASTIC, MAT_NEWTON_FLUID, MAT_LOCAL_RHEO, MAT_DRUCKER_PRAGER = 1, 2, 3, 4
MAT_VON_MISES, MAT_MOHR_COULOMB, MAT_MODIFIED_CAM_CLAY = 5, 6, 7
BC_GRAVITY, BC_NODE_DOMAIN, BC_RIGID_SDF, BC_BODY_FORCE, BC_PLANAR_DOMAIN = 1, 2, 4, 5, 6
MATF_UPDATE_HISTORY, MATF_VELGRAD_IS_INCREMENT, MATF_APEX_GUARD, MATF_HAS_STRESS_REF = 1, 2, 4, 8
PF_RIGID, PF_ACTIVE = 1, 2
MAX_MATERIALS = 8
MAX_BCS = 8

vp = C.c_void_p


class Grid(C.Structure):
    _fields_ = [("origin", C.c_float * 3), ("cell_size", C.c_float), ("inv_cell_size", C.c_float),
                ("num_cells", C.c_int32 * 3), ("num_cells_total", C.c_int32)]


class Particles(C.Structure):
    _fields_ = [("n", C.c_int32), ("x", vp), ("v", vp), ("stress", vp), ("F", vp), ("L", vp),
                ("mass", vp), ("volume", vp), ("volume0", vp), ("color", vp), ("flags", vp), ("id", vp),
                ("eps_e", vp), ("stress_ref", vp), ("hist", vp)]


def node_vel_stride(dim):
    """Floats per node of pmpm_nodes.vel (PMPM_NODE_VEL_STRIDE)."""
    return 8 if dim == 3 else 2 * dim


class Nodes(C.Structure):
    _fields_ = [("mass", vp), ("mom", vp), ("mom_nt", vp), ("vel", vp), ("clamp", vp)]


class Spatial(C.Structure):
    _fields_ = [("hash", vp), ("hash_sorted", vp), ("sorted", vp), ("cell_start", vp), ("cell_end", vp),
                ("bins", vp)]


class Adjoint(C.Structure):
    _fields_ = [("x", vp), ("v", vp), ("stress", vp), ("F", vp), ("eps_e", vp)]


class Material(C.Structure):
    _fields_ = [("kind", C.c_int32), ("flags", C.c_int32), ("p", C.c_float * 16)]


class BC(C.Structure):
    _fields_ = [("kind", C.c_int32), ("i", C.c_int32 * 6), ("f", C.c_float * 24), ("values", vp), ("mask", vp)]


class Params(C.Structure):
    _fields_ = [("dim", C.c_int32), ("shape_fn", C.c_int32), ("dt", C.c_float), ("alpha", C.c_float),
                ("small_mass_cutoff", C.c_float), ("node_types_built", C.c_int32), ("wrap_quirk", C.c_int32),
                ("num_materials", C.c_int32), ("materials", Material * MAX_MATERIALS),
                ("num_bcs", C.c_int32), ("bcs", BC * MAX_BCS)]


MAX_LANES = 8


class Scene(C.Structure):
    _fields_ = [("ctx", vp), ("prm", Params), ("grid", Grid), ("set", Particles * 2), ("part", Spatial),
                ("nodes", Nodes), ("musl", C.c_int32)]


class AdjointScene(C.Structure):
    _fields_ = [("ctx", vp), ("prm", Params), ("grid", Grid), ("state", Particles), ("out_bar", Adjoint),
                ("in_bar", Adjoint), ("param_bar", vp), ("recompute", C.c_int32), ("part", Spatial),
                ("nodes", Nodes)]


class BatchTuning(C.Structure):
    _fields_ = [("lanes", C.c_int32), ("use_graph", C.c_int32), ("step_ms", C.c_float)]


P = C.POINTER
# name -> (restype, argtypes); every function declared in include/pmpm.h
PROTOTYPES = {
    "pmpm_context_create": (C.c_int, [P(vp)]),
    "pmpm_context_destroy": (C.c_int, [vp]),
    "pmpm_context_reserve": (C.c_int, [vp, C.c_int32, C.c_int32]),
    "pmpm_context_out_of_grid": (C.c_int, [vp, vp, P(C.c_int32)]),
    "pmpm_partition": (C.c_int, [vp, C.c_int32, P(Grid), vp, C.c_int32, P(Spatial), vp]),
    "pmpm_shape_function": (C.c_int, [C.c_int32, C.c_int32, C.c_int32, P(Grid), vp, C.c_int32, vp, vp, vp]),
    "pmpm_p2g": (C.c_int, [vp, P(Params), P(Grid), P(Particles), P(Spatial), P(Nodes), vp]),
    "pmpm_grid_bc": (C.c_int, [P(Params), P(Grid), P(Nodes), C.c_int32, C.c_int32, C.c_int32, vp]),
    "pmpm_rigid_levelset": (C.c_int, [P(Params), P(Grid), P(Particles), P(Spatial), P(Nodes), P(C.c_float),
                                      P(C.c_float), vp, vp, vp]),
    "pmpm_rigid_advance": (C.c_int, [C.c_int32, C.c_int32, vp, vp, vp, vp]),
    "pmpm_permute": (C.c_int, [C.c_int32, C.c_int32, C.c_int32, vp, vp, vp, C.c_int32, vp]),
    "pmpm_g2p": (C.c_int, [P(Params), P(Grid), P(Particles), P(Particles), P(Spatial), P(Nodes), vp]),
    "pmpm_musl_map": (C.c_int, [vp, P(Params), P(Grid), P(Particles), P(Particles), P(Spatial), P(Nodes), vp]),
    "pmpm_musl_g2p": (C.c_int, [P(Params), P(Grid), P(Particles), P(Particles), P(Spatial), P(Nodes), vp]),
    "pmpm_stress_update": (C.c_int, [P(Params), C.c_int32, P(Particles), vp]),
    "pmpm_step": (C.c_int, [vp, P(Params), P(Grid), C.c_int32, P(Particles), P(Particles), P(Spatial), P(Nodes),
                            vp]),
    "pmpm_step_noalloc": (C.c_int, [P(Params), P(Grid), C.c_int32, P(Particles), P(Particles), P(Spatial), P(Nodes),
                                    vp]),
    "pmpm_adjoint_step": (C.c_int, [vp, P(Params), P(Grid), P(Particles), P(Adjoint), P(Adjoint), vp, C.c_int32,
                                    P(Spatial), P(Nodes), vp]),
    "pmpm_batch_create": (C.c_int, [P(vp), C.c_int32]),
    "pmpm_batch_destroy": (C.c_int, [vp]),
    "pmpm_batch_assign": (C.c_int, [P(C.c_float), C.c_int32, C.c_int32, P(C.c_int32)]),
    "pmpm_batch_set_weights": (C.c_int, [vp, P(C.c_float), C.c_int32]),
    "pmpm_batch_step": (C.c_int, [vp, P(Scene), C.c_int32, C.c_int32, vp]),
    "pmpm_batch_capture": (C.c_int, [vp, P(Scene), C.c_int32, C.c_int32, vp]),
    "pmpm_batch_replay": (C.c_int, [vp, C.c_int32, vp]),
    "pmpm_batch_autotune": (C.c_int, [vp, P(Scene), C.c_int32, C.c_int32, C.c_int32, vp, P(BatchTuning)]),
    "pmpm_batch_get_tuning": (C.c_int, [vp, P(BatchTuning)]),
    "pmpm_batch_adjoint_step": (C.c_int, [vp, P(AdjointScene), C.c_int32, vp]),
    "pmpm_timing_enable": (C.c_int, [C.c_int32]),
    "pmpm_timing_read": (C.c_int, [C.c_char_p, C.c_int32, P(C.c_int32), P(C.c_float), C.c_int32, P(C.c_int32)]),
    "pmpm_last_error": (C.c_char_p, []),
    "pmpm_version": (C.c_char_p, []),
}

_lib = None


def lib():
    """Load libpmpm_hip.so once and attach the prototypes."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError(f"{LIB_PATH} is missing; build with `python pyroclastmpm_amd/build.py`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in PROTOTYPES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc, what="pmpm call"):
    if rc != PMPM_OK:
        detail = ""
        if rc == PMPM_ERR_HIP:
            detail = " (" + lib().pmpm_last_error().decode(errors="replace") + ")"
        raise RuntimeError(f"{what} failed with code {rc}{detail}")


def make_grid(origin, end, cell_size, dim):
    """Grid exactly as the reference builds it (types_common.h:262-276), in FP32."""
    import numpy as np

    g = Grid()
    h = np.float32(cell_size)
    g.cell_size = h
    g.inv_cell_size = np.float32(1.0) / h
    total = 1
    for a in range(3):
        g.origin[a] = float(origin[a]) if a < dim else 0.0
        g.num_cells[a] = 1
    for a in range(dim):
        n = int((np.float32(end[a]) - np.float32(origin[a])) / h) + 1
        g.num_cells[a] = n
        total *= n
    g.num_cells_total = total
    return g


class Context:
    """Owns a pmpm_context (sort workspace)."""

    def __init__(self, max_particles=0, max_cells=0):
        self.ptr = vp()
        check(lib().pmpm_context_create(C.byref(self.ptr)), "pmpm_context_create")
        if max_particles:
            self.reserve(max_particles, max_cells)

    def reserve(self, n, cells):
        check(lib().pmpm_context_reserve(self.ptr, int(n), int(cells)), "pmpm_context_reserve")

    def __del__(self):
        try:
            if self.ptr:
                lib().pmpm_context_destroy(self.ptr)
                self.ptr = vp()
        except Exception:
            pass


def timing_enable(on=True):
    """Per-thread kernel timing with HIP events (pmpm_timing_enable)."""
    check(lib().pmpm_timing_enable(1 if on else 0), "pmpm_timing_enable")


def timing_read(max_kernels=64):
    """{kernel name: (launches, total ms)} since the last read
    (pmpm_timing_read; waits for the recorded events)."""
    buf = C.create_string_buffer(1 << 16)
    counts = (C.c_int32 * max_kernels)()
    ms = (C.c_float * max_kernels)()
    nk = C.c_int32(0)
    check(lib().pmpm_timing_read(buf, len(buf), counts, ms, max_kernels, C.byref(nk)), "pmpm_timing_read")
    names = buf.value.decode().split("\n") if nk.value else []
    return {names[k]: (int(counts[k]), float(ms[k])) for k in range(min(nk.value, max_kernels))}
