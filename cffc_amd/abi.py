"""ctypes mirrors of the descriptor structs and constants of
include/cffc_euler2d.h (the C ABI of libcffc_hip.so)."""
import ctypes
from dataclasses import dataclass

import numpy as np

# boundary conditions (src_2D/CFD/CFD.h:273-351)
BC_NONE = 0
BC_CONSTANT_EXTRAPOLATION = 4001
BC_PERIODIC = 6000
BC_REFLECTION = 7001
BC_CHARACTERISTIC = 8001
BC_FIXED = 11000
# time integration (CFD.h:651-659)
TIME_STEPPING_EXPLICIT_EULER = 0
TIME_STEPPING_EXPLICIT_PREDICTOR_CORRECTOR = 3
TIME_STEPPING_MULTISTAGE_OPTIMAL_SMOOTHING = 7
TIME_STEPPING_EXPLICIT_RUNGE_KUTTA = 8
# limiters (CFD.h:749-761)
LIMITER_ONE = -1
LIMITER_ZERO = 0
LIMITER_VENKATAKRISHNAN = 9
FLUX_FUNCTION_ROE = 2
FLUX_FUNCTION_HLLE = 4
MULTIGRID_V_CYCLE = 1
MULTIGRID_W_CYCLE = 2
RECONSTRUCTION_LEAST_SQUARES = 2
GLOBAL_TIME_STEPPING = 0
SCALAR_LOCAL_TIME_STEPPING = 1
FLUX_FP64 = 0
FLUX_FP32 = 1
# runtime errors (CFFC_E_*)
E_ARG, E_HIP, E_NOMEM, E_TOPOLOGY, E_NODEVICE, E_STEP_LIMIT = -1, -2, -3, -4, -5, -6

NBR_N, NBR_S, NBR_E, NBR_W, NBR_NE, NBR_NW, NBR_SE, NBR_SW = range(8)

_i32p = ctypes.POINTER(ctypes.c_int32)
_f64p = ctypes.POINTER(ctypes.c_double)


class MeshDesc(ctypes.Structure):
    _fields_ = [("nblk", ctypes.c_int32), ("nghost", ctypes.c_int32),
                ("ni", _i32p), ("nj", _i32p), ("xnode", _f64p), ("ynode", _f64p),
                ("nbr", _i32p), ("bc", _i32p), ("wo", _f64p), ("level", _i32p), ("nbr2", _i32p),
                ("lattice", _i32p)]


class SchemeDesc(ctypes.Structure):
    _fields_ = [("i_time_integration", ctypes.c_int32), ("n_stage", ctypes.c_int32),
                ("cfl_number", ctypes.c_double), ("i_reconstruction", ctypes.c_int32),
                ("i_limiter", ctypes.c_int32), ("i_flux_function", ctypes.c_int32),
                ("local_time_stepping", ctypes.c_int32), ("flux_precision", ctypes.c_int32),
                ("axisymmetric", ctypes.c_int32)]


class MGDesc(ctypes.Structure):
    _fields_ = [("levels", ctypes.c_int32), ("i_cycle", ctypes.c_int32),
                ("smooths_on_finest_level", ctypes.c_int32), ("pre_smooths", ctypes.c_int32),
                ("post_smooths", ctypes.c_int32), ("smooths_on_coarsest_level", ctypes.c_int32),
                ("defect_correction", ctypes.c_int32), ("prolong_using_injection", ctypes.c_int32),
                ("first_order_coarse", ctypes.c_int32), ("injection_at_dirichlet", ctypes.c_int32),
                ("update_stability_switch", ctypes.c_int32), ("max_update_reductions", ctypes.c_int32),
                ("absolute_tolerance", ctypes.c_double), ("relative_tolerance", ctypes.c_double),
                ("fmg_cycles", ctypes.c_int32), ("pad", ctypes.c_int32),
                ("fmg_absolute_tolerance", ctypes.c_double), ("fmg_relative_tolerance", ctypes.c_double)]


@dataclass
class Scheme:
    """The Euler2D_Input_Parameters fields used on this path
    (Euler2DInput.h:57-118).  Defaults: the 4-stage Runge-Kutta, Roe,
    least-squares, Venkatakrishnan, CFL 0.5, global time stepping scheme."""
    i_time_integration: int = TIME_STEPPING_EXPLICIT_RUNGE_KUTTA
    n_stage: int = 4
    cfl_number: float = 0.5
    i_reconstruction: int = RECONSTRUCTION_LEAST_SQUARES
    i_limiter: int = LIMITER_VENKATAKRISHNAN
    i_flux_function: int = FLUX_FUNCTION_ROE
    local_time_stepping: int = GLOBAL_TIME_STEPPING
    flux_precision: int = FLUX_FP64
    axisymmetric: int = 0


def _keep(arr, dtype, keep):
    a = np.ascontiguousarray(arr, dtype=dtype)
    keep.append(a)
    return a


def mesh_desc(mesh, keep):
    """Build a MeshDesc for a cffc_amd.mesh.Mesh; arrays are appended to keep
    so they outlive the struct."""
    ni = _keep(mesh.ni, np.int32, keep)
    nj = _keep(mesh.nj, np.int32, keep)
    xn = _keep(mesh.xnode, np.float64, keep)
    yn = _keep(mesh.ynode, np.float64, keep)
    nbr = _keep(mesh.nbr, np.int32, keep)
    bc = _keep(mesh.bc, np.int32, keep)
    wo = _keep(mesh.wo, np.float64, keep)
    level = getattr(mesh, "level", None)
    nbr2 = getattr(mesh, "nbr2", None)
    lv = _keep(level, np.int32, keep).ctypes.data_as(_i32p) if level is not None else None
    n2 = _keep(nbr2, np.int32, keep).ctypes.data_as(_i32p) if nbr2 is not None else None
    # lattice positions are passed for meshes that need them (blocks of
    # uneven size); quadtree_box meshes record them but use the neighbour tables
    lat = getattr(mesh, "lattice", None)
    use_lat = lat is not None and (mesh.info.get("kind") != "quadtree" or mesh.info.get("use_lattice", False))
    la = _keep(lat, np.int32, keep).ctypes.data_as(_i32p) if use_lat else None
    return MeshDesc(int(mesh.nblk), 2, ni.ctypes.data_as(_i32p), nj.ctypes.data_as(_i32p),
                    xn.ctypes.data_as(_f64p), yn.ctypes.data_as(_f64p), nbr.ctypes.data_as(_i32p),
                    bc.ctypes.data_as(_i32p), wo.ctypes.data_as(_f64p), lv, n2, la)


def scheme_desc(s):
    return SchemeDesc(s.i_time_integration, s.n_stage, s.cfl_number, s.i_reconstruction,
                      s.i_limiter, s.i_flux_function, s.local_time_stepping, s.flux_precision,
                      s.axisymmetric)
