"""Python host interface of the MI355X 2D Euler path, over the C ABI of
libcffc_hip.so (include/cffc_euler2d.h).

Euler2DQuadBlocks mirrors the reference's multi-block call surface
(src_2D/Euler2D/Euler2DQuadMultiBlock.cc): the methods carry the names of the
reference functions they replace and return the reference's int codes.
Everything runs on the GPU; there is no CPU fallback -- constructing a solver
without the built library or without a HIP device raises.
"""
import ctypes
import os

import numpy as np

from . import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcffc_hip.so")

_lib = None

# tile shapes the stage kernel is built for, in the autotuner's order
# (CFFC_TILE_LIST in csrc/euler2d_kernels.hpp)
TILE_SHAPES = [(16, 16), (32, 8), (8, 32), (16, 8), (8, 8), (32, 4), (32, 16), (16, 4), (8, 16)]


class CffcError(RuntimeError):
    pass


def library_path():
    return LIB_PATH


def load():
    """Load libcffc_hip.so (built in-tree by __graft_entry__.build() or
    `make -C cffc_amd/csrc`)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise CffcError("libcffc_hip.so is not built (run __graft_entry__.build())")
    L = ctypes.CDLL(LIB_PATH)
    vp, i32, i64, d = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_double
    pd = ctypes.POINTER(ctypes.c_double)
    L.cffc_create.argtypes = [vp, vp, ctypes.POINTER(vp)]
    L.cffc_create.restype = i32
    L.cffc_destroy.argtypes = [vp]
    L.cffc_destroy.restype = None
    L.cffc_num_cells.argtypes = [vp]
    L.cffc_num_cells.restype = i64
    for n in ("cffc_set_W", "cffc_set_U", "cffc_get_W", "cffc_get_U", "cffc_get_dt"):
        getattr(L, n).argtypes = [vp, pd, i64]
        getattr(L, n).restype = i32
    L.cffc_get_dUdt.argtypes = [vp, i32, pd, i64]
    L.cffc_exchange_and_bcs.argtypes = [vp]
    L.cffc_apply_flux_corrections.argtypes = [vp, i32]
    L.cffc_cfl.argtypes = [vp, pd, i32]
    L.cffc_set_global_timestep.argtypes = [vp, ctypes.c_double]
    L.cffc_run_time_accurate.argtypes = [vp, ctypes.c_double, i32, ctypes.c_double, pd, ctypes.POINTER(ctypes.c_int)]
    L.cffc_dudt_stage.argtypes = [vp, i32]
    L.cffc_update_stage.argtypes = [vp, i32, ctypes.POINTER(ctypes.c_int64)]
    L.cffc_residual_norms.argtypes = [vp, i32, pd]
    L.cffc_freeze_limiters.argtypes = [vp, i32]
    L.cffc_set_store_limiters.argtypes = [vp, i32]
    L.cffc_run_steps.argtypes = [vp, i32, d, pd, pd, ctypes.POINTER(ctypes.c_int)]
    L.cffc_run_steps_graph.argtypes = [vp, i32, pd, pd]
    L.cffc_run_steps_noalloc.argtypes = [vp, i32, d, pd, pd, ctypes.POINTER(ctypes.c_int)]
    L.cffc_synchronize.argtypes = [vp]
    L.cffc_stream.argtypes = [vp]
    L.cffc_stream.restype = vp
    L.cffc_set_tile.argtypes = [vp, i32, i32]
    L.cffc_get_tile.argtypes = [vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
    L.cffc_set_kernel_timing.argtypes = [vp, i32]
    L.cffc_kernel_times.argtypes = [vp, pd]
    pi32 = ctypes.POINTER(ctypes.c_int32)
    L.cffc_set_parts.argtypes = [vp, i32, pi32]
    L.cffc_num_parts.argtypes = [vp]
    L.cffc_set_stall_freeze.argtypes = [vp, i32]
    L.cffc_run_steady.argtypes = [vp, i32, d, d, i32, pi32, pd, ctypes.POINTER(ctypes.c_int)]
    L.cffc_get_part_history.argtypes = [vp, i32, pd]
    L.cffc_step_times.argtypes = [vp, i32, pd]
    L.cffc_set_residual_smoothing.argtypes = [vp, d, i32]
    L.cffc_residual_smoothing.argtypes = [vp, i32]
    pint = ctypes.POINTER(ctypes.c_int)
    L.cffc_autotune_tile.argtypes = [vp, i32, pint, pint, pd, pint]
    L.cffc_mg_create.argtypes = [vp, i32, vp, vp, ctypes.POINTER(vp)]
    L.cffc_mg_destroy.argtypes = [vp]
    L.cffc_mg_level.argtypes = [vp, i32]
    L.cffc_mg_level.restype = vp
    L.cffc_mg_set_residual_smoothing.argtypes = [vp, d, i32]
    L.cffc_mg_initialize.argtypes = [vp]
    L.cffc_mg_cycle.argtypes = [vp, ctypes.POINTER(ctypes.c_int64)]
    L.cffc_mg_execute.argtypes = [vp, i32, d, pd, pint, pint]
    L.cffc_mg_get_P.argtypes = [vp, i32, pd, ctypes.c_int64]
    L.cffc_mg_set_graph.argtypes = [vp, i32]
    pi32 = ctypes.POINTER(ctypes.c_int32)
    L.cffc_mg_set_parts.argtypes = [vp, i32, pi32, pd, pd]
    L.cffc_mg_part_status.argtypes = [vp, pd, pi32]
    L.cffc_check_arith.argtypes = [pd, pd, i64, pd]
    L.cffc_last_error.argtypes = []
    L.cffc_last_error.restype = ctypes.c_char_p
    _lib = L
    return L


def _p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _check(rc, what):
    if rc < 0:
        raise CffcError("%s failed (%d): %s" % (what, rc, load().cffc_last_error().decode()))
    return rc


def check_arith(a, b):
    """The stage kernels' FP64 division, square root, minimum and maximum
    beside the compiler's, on the device, for operand arrays a and b:
    returns an (n, 8) array (see cffc_check_arith)."""
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    if a.shape != b.shape or a.ndim != 1:
        raise ValueError("a and b must be 1-D arrays of one length")
    out = np.empty((a.size, 8), dtype=np.float64)
    _check(load().cffc_check_arith(_p(a), _p(b), a.size, _p(out)), "cffc_check_arith")
    return out


class Euler2DQuadBlocks:
    """All solution blocks of one mesh resident on the GPU (the reference's
    Euler2D_Quad_Block array plus its AdaptiveBlock2D_List)."""

    def __init__(self, mesh, scheme=None):
        L = load()
        self.mesh = mesh
        self.scheme = scheme or abi.Scheme()
        self._keep = []
        self._md = abi.mesh_desc(mesh, self._keep)
        self._sd = abi.scheme_desc(self.scheme)
        h = ctypes.c_void_p()
        _check(L.cffc_create(ctypes.byref(self._md), ctypes.byref(self._sd), ctypes.byref(h)), "cffc_create")
        self.h = h
        self.ncell = L.cffc_num_cells(h)

    def close(self):
        if getattr(self, "h", None) and getattr(self, "_owned", True):
            load().cffc_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- state -----------------------------------------------------------------
    def set_W(self, W):
        W = np.ascontiguousarray(W, dtype=np.float64).reshape(self.ncell, 4)
        _check(load().cffc_set_W(self.h, _p(W), self.ncell), "cffc_set_W")

    def set_U(self, U):
        U = np.ascontiguousarray(U, dtype=np.float64).reshape(self.ncell, 4)
        _check(load().cffc_set_U(self.h, _p(U), self.ncell), "cffc_set_U")

    def get_W(self):
        out = np.empty((self.ncell, 4))
        _check(load().cffc_get_W(self.h, _p(out), self.ncell), "cffc_get_W")
        return out

    def get_U(self):
        out = np.empty((self.ncell, 4))
        _check(load().cffc_get_U(self.h, _p(out), self.ncell), "cffc_get_U")
        return out

    def get_dUdt(self, k=0):
        out = np.empty((self.ncell, 4))
        _check(load().cffc_get_dUdt(self.h, k, _p(out), self.ncell), "cffc_get_dUdt")
        return out

    def get_dt(self):
        out = np.empty(self.ncell)
        _check(load().cffc_get_dt(self.h, _p(out), self.ncell), "cffc_get_dt")
        return out

    # ---- the reference's multi-block calls ---------------------------------------
    def Send_All_Messages_and_BCs(self):
        """Send_All_Messages(..., NUM_VAR_EULER2D, OFF) followed by BCs()."""
        return _check(load().cffc_exchange_and_bcs(self.h), "cffc_exchange_and_bcs")

    def prime_ghosts(self):
        """Send_All_Messages + BCs repeated until every ghost derives from the
        interior state: once on one level; max_level + 2 times across
        resolution changes, where the coarse-to-fine leg reads coarse BC ghosts
        filled after it.  Returns the number of rounds."""
        lv = getattr(self.mesh, "level", None)
        rounds = int(np.max(lv)) + 2 if lv is not None and np.any(lv != lv[0]) else 1
        for _ in range(rounds):
            self.Send_All_Messages_and_BCs()
        return rounds

    def Apply_Boundary_Flux_Corrections(self, i_stage):
        """Send_Conservative_Flux_Corrections() + Apply_Boundary_Flux_Corrections_
        Multistage_Explicit() (no-op without resolution changes)."""
        return _check(load().cffc_apply_flux_corrections(self.h, i_stage), "cffc_apply_flux_corrections")

    def CFL(self, set_global=True):
        """CFL() + CFFC_Minimum_MPI() (+ Set_Global_TimeStep()); returns dTime."""
        d = ctypes.c_double()
        _check(load().cffc_cfl(self.h, ctypes.byref(d), 1 if set_global else 0), "cffc_cfl")
        return d.value

    def Set_Global_TimeStep(self, dt):
        _check(load().cffc_set_global_timestep(self.h, ctypes.c_double(dt)), "cffc_set_global_timestep")

    def run_time_accurate(self, t_max, max_steps=0, time=0.0, freeze_level=0.0):
        """Time-accurate loop from `time` to t_max (the last step clipped);
        returns (rc, time, steps).  max_steps > 0 caps the steps (rc is then
        CFFC_E_STEP_LIMIT if the cap ends the run first)."""
        t, n = ctypes.c_double(time), ctypes.c_int(0)
        rc = load().cffc_run_time_accurate(self.h, ctypes.c_double(t_max), max_steps, ctypes.c_double(freeze_level),
                                           ctypes.byref(t), ctypes.byref(n))
        if rc != abi.E_STEP_LIMIT:
            _check(rc, "cffc_run_time_accurate")
        return rc, t.value, n.value

    def dUdt_Multistage_Explicit(self, i_stage):
        return _check(load().cffc_dudt_stage(self.h, i_stage), "cffc_dudt_stage")

    def Update_Solution_Multistage_Explicit(self, i_stage):
        """Returns the reference's code: 0, or the i-index of the first cell
        with non-positive density/energy (self.bad_cell then holds
        (block, j, i))."""
        bad = ctypes.c_int64(-1)
        rc = _check(load().cffc_update_stage(self.h, i_stage, ctypes.byref(bad)), "cffc_update_stage")
        v = bad.value
        self.bad_cell = None if v < 0 else (v >> 32, (v >> 16) & 0xFFFF, v & 0xFFFF)
        return rc

    def Residual_Norms(self, var=1):
        """(L1, L2, max) of dUdt[..][0][var] (L1/L2/Max_Norm_Residual)."""
        out = np.zeros(3)
        _check(load().cffc_residual_norms(self.h, var, _p(out)), "cffc_residual_norms")
        return out

    def set_residual_smoothing(self, epsilon, iterations):
        """Residual_Smoothing_Epsilon / _Gauss_Seidel_Iterations (0: off)."""
        _check(load().cffc_set_residual_smoothing(self.h, epsilon, iterations), "cffc_set_residual_smoothing")

    def Residual_Smoothing(self, i_stage):
        return _check(load().cffc_residual_smoothing(self.h, i_stage), "cffc_residual_smoothing")

    def Freeze_Limiters(self, freeze=True):
        return _check(load().cffc_freeze_limiters(self.h, 1 if freeze else 0), "cffc_freeze_limiters")

    def store_limiters(self, on=True):
        return _check(load().cffc_set_store_limiters(self.h, 1 if on else 0), "cffc_set_store_limiters")

    # ---- the fused time-marching loop (Euler2DQuadSolver explicit branch) --------
    def run_steps(self, nsteps, tol=0.0, graph=False):
        """Returns (rc, steps_done, norms[steps,3], dt[steps])."""
        L = load()
        norms = np.zeros((max(nsteps, 1), 3))
        dts = np.zeros(max(nsteps, 1))
        if graph:
            rc = _check(L.cffc_run_steps_graph(self.h, nsteps, _p(norms), _p(dts)), "cffc_run_steps_graph")
            n = nsteps
        else:
            done = ctypes.c_int(0)
            rc = _check(L.cffc_run_steps(self.h, nsteps, tol, _p(norms), _p(dts), ctypes.byref(done)),
                        "cffc_run_steps")
            n = done.value
        return rc, n, norms[:n], dts[:n]

    def run_steps_noalloc(self, nsteps, tol=0.0):
        """cffc_run_steps_noalloc: the same steps as run_steps without any
        allocation in the library (at most 4096 steps).  Returns (rc,
        steps_done, norms[steps,3], dt[steps])."""
        norms = np.zeros((max(nsteps, 1), 3))
        dts = np.zeros(max(nsteps, 1))
        done = ctypes.c_int(0)
        rc = _check(load().cffc_run_steps_noalloc(self.h, nsteps, tol, _p(norms), _p(dts), ctypes.byref(done)),
                    "cffc_run_steps_noalloc")
        n = done.value
        return rc, n, norms[:n], dts[:n]

    # ---- batched steady runs ------------------------------------------------------
    def set_parts(self, first_block):
        """Independent cases: part q owns blocks first_block[q]..first_block[q+1]-1."""
        fb = np.ascontiguousarray(first_block, dtype=np.int32)
        _check(load().cffc_set_parts(self.h, len(fb) - 1, fb.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))),
               "cffc_set_parts")
        self.npart = len(fb) - 1

    def set_stall_freeze(self, window):
        """Freeze a running part's limiters once its L2 has not gone 1% below
        its best for `window` steps (no counterpart in the reference; 0: off)."""
        _check(load().cffc_set_stall_freeze(self.h, int(window)), "cffc_set_stall_freeze")

    def run_steady(self, max_steps, tol, freeze_level=0.0, history=False, resume=False):
        """Every part to its own L2 <= tol (limiters frozen at freeze_level).
        resume continues the previous run for up to max_steps more steps.
        Returns (rc, steps_done, part_steps[npart], part_l2[npart], history
        [steps_done, npart] or None)."""
        L = load()
        npart = max(L.cffc_num_parts(self.h), 1)
        steps = np.zeros(npart, dtype=np.int32)
        l2 = np.zeros(npart)
        done = ctypes.c_int(0)
        flags = (1 if history else 0) | (2 if resume else 0)
        rc = _check(L.cffc_run_steady(self.h, max_steps, tol, freeze_level, flags,
                                      steps.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), _p(l2),
                                      ctypes.byref(done)), "cffc_run_steady")
        npart = L.cffc_num_parts(self.h)
        steps, l2 = steps[:npart], l2[:npart]
        hist = None
        if history:
            hist = np.zeros((max(done.value, 1), npart))
            _check(L.cffc_get_part_history(self.h, done.value, _p(hist)), "cffc_get_part_history")
            hist = hist[:done.value]
        return rc, done.value, steps, l2, hist

    def step_times(self, nsteps):
        """Seconds from the start of the last run to the end of each step."""
        out = np.zeros(max(nsteps, 1))
        _check(load().cffc_step_times(self.h, nsteps, _p(out)), "cffc_step_times")
        return out[:nsteps]

    def synchronize(self):
        _check(load().cffc_synchronize(self.h), "cffc_synchronize")

    def stream(self):
        return load().cffc_stream(self.h)

    def set_tile(self, ti, tj):
        _check(load().cffc_set_tile(self.h, ti, tj), "cffc_set_tile")

    def get_tile(self):
        a, b = ctypes.c_int(), ctypes.c_int()
        load().cffc_get_tile(self.h, ctypes.byref(a), ctypes.byref(b))
        return a.value, b.value

    def autotune(self, nsteps=4):
        """Pick the fastest tile shape for this mesh (state is preserved).
        Returns ((ti, tj), {(ti, tj): ms per step})."""
        a, b, n = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        ms = np.zeros(16)
        _check(load().cffc_autotune_tile(self.h, nsteps, ctypes.byref(a), ctypes.byref(b), _p(ms), ctypes.byref(n)),
               "cffc_autotune_tile")
        return (a.value, b.value), dict(zip(TILE_SHAPES[:n.value], ms[:n.value].tolist()))

    def kernel_timing(self, on=True):
        _check(load().cffc_set_kernel_timing(self.h, 1 if on else 0), "cffc_set_kernel_timing")

    def kernel_times(self):
        """{stage_ms, stage_launches, exchange_ms, exchange_launches} since
        kernel_timing(True)."""
        out = np.zeros(4)
        _check(load().cffc_kernel_times(self.h, _p(out)), "cffc_kernel_times")
        return dict(stage_ms=out[0], stage_launches=int(out[1]), exchange_ms=out[2], exchange_launches=int(out[3]))

    def step_reference_order(self, nsteps):
        """The reference's loop (Euler2DQuadSolvers.cc:549-898) driven call by
        call through the granular entry points; returns (rc, norms)."""
        s = self.scheme
        norms = []
        for n in range(nsteps + 1):
            self.CFL(set_global=(s.local_time_stepping == abi.GLOBAL_TIME_STEPPING))
            if n > 0:
                norms.append(self.Residual_Norms(1))
            if n == nsteps:
                break
            for st in range(1, s.n_stage + 1):
                self.Send_All_Messages_and_BCs()
                self.dUdt_Multistage_Explicit(st)
                self.Apply_Boundary_Flux_Corrections(st)
                rc = self.Update_Solution_Multistage_Explicit(st)
                if rc:
                    return rc, np.array(norms)
        return 0, np.array(norms)
