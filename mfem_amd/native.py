"""ctypes binding of libmfem_hip.so (the C ABI in include/mfem_hip.h).

The library must have been built (python -m mfem_amd.build); loading fails
loudly otherwise -- there is no CPU fallback for any entry point.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libmfem_hip.so")
_lib = None

dp = C.POINTER(C.c_double)
ip = C.POINTER(C.c_int)

EXPORTED = [
    "mfem_last_error", "mfem_version", "mfem_sim_create_json", "mfem_sim_create_json_2d",
    "mfem_sim_create", "mfem_sim_write_obj", "mfem_sim_write_mesh", "mfem_sim_write_state",
    "mfem_boundary_faces", "mfem_sim_newton_history",
    "mfem_sim_destroy", "mfem_sim_step", "mfem_sim_update_system", "mfem_sim_solve",
    "mfem_sim_energy", "mfem_sim_stats", "mfem_sim_get_positions", "mfem_sim_size",
    "mfem_sim_get", "mfem_sim_get_int", "mfem_sim_set_profiling", "mfem_sim_matrix_info",
    "mfem_sim_time_spmv", "mfem_sim_spmv_tuning", "mfem_sim_set_cg_shape", "mfem_sim_cg_tuning",
    "mfem_sim_set_element_kernel",
    "mfem_sim_kkt_size", "mfem_sim_kkt_apply",
    "mfem_material_s6_batch", "mfem_snh_f9_batch", "mfem_polar_batch",
    "mfem_batch_create", "mfem_batch_destroy", "mfem_batch_step", "mfem_batch_sizes",
    "mfem_batch_get_positions", "mfem_batch_stats", "mfem_batch_set_cg_shape",
    "mfem_batch_cg_shape", "mfem_sim_set_callback", "mfem_pcg_bsr3", "mfem_pcg_bsr3_noalloc",
    "mfem_batch_set_graphs", "mfem_batch_launch_counts",
]

# scalar-row limit of mfem_pcg_bsr3_noalloc (MFEM_PCG_NOALLOC_MAX_ROWS)
PCG_NOALLOC_MAX_ROWS = 524288

# void (*)(mfem_sim*, int iteration, void* user)
NEWTON_CALLBACK = C.CFUNCTYPE(None, C.c_void_p, C.c_int, C.c_void_p)


class SimParams(C.Structure):
    _fields_ = [
        ("nV", C.c_int), ("nT", C.c_int),
        ("V", dp), ("T", ip), ("mu", dp), ("la", dp), ("mat", ip), ("fixed", ip), ("v0", dp),
        ("density", C.c_double), ("dt", C.c_double), ("gravity", C.c_double * 3),
        ("bdf_order", C.c_int), ("mixed", C.c_int), ("linear_solver", C.c_char_p),
        ("itr_tol", C.c_double), ("max_iters", C.c_int), ("max_newton", C.c_int),
        ("newton_tol", C.c_double), ("ls_iters", C.c_int), ("precision", C.c_int),
        ("bc_type", C.c_int), ("bc_axis", C.c_int), ("bc_ratio", C.c_double),
        ("dim", C.c_int), ("bc_velocity", C.c_double), ("x0", dp),
    ]


class StepStats(C.Structure):
    _fields_ = [
        ("newton_iters", C.c_int), ("pcg_iters", C.c_int), ("refinements", C.c_int),
        ("ls_iters", C.c_int), ("ls_fail", C.c_int),
        ("decrement", C.c_double), ("energy", C.c_double), ("alpha", C.c_double),
        ("res", C.c_double), ("rel_residual", C.c_double),
        ("update_ms", C.c_double), ("solve_ms", C.c_double), ("ls_ms", C.c_double),
        ("step_ms", C.c_double), ("spmv_ms", C.c_double), ("spmv_count", C.c_int),
    ]


class MfemError(RuntimeError):
    pass


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise MfemError(f"{LIB_PATH} is missing: build it with `python -m mfem_amd.build`")
        # One HIP runtime per process: torch bundles its own libamdhip64 (same
        # soname). Loading torch first lets this library bind to that instance
        # by soname; loading it second would bring up a second runtime, and
        # whichever of the two initialises later finds no device.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.mfem_last_error.restype = C.c_char_p
        L.mfem_sim_create_json.argtypes = [C.c_char_p, C.POINTER(vp)]
        L.mfem_sim_create_json_2d.argtypes = [C.c_char_p, C.POINTER(vp)]
        L.mfem_sim_write_obj.argtypes = [vp, C.c_char_p]
        L.mfem_sim_write_mesh.argtypes = [vp, C.c_char_p]
        L.mfem_sim_write_state.argtypes = [vp, C.c_char_p, C.c_char_p]
        L.mfem_boundary_faces.argtypes = [dp, C.c_int, ip, C.c_int, ip, C.POINTER(C.c_int)]
        L.mfem_sim_newton_history.argtypes = [vp, dp, C.POINTER(C.c_int)]
        L.mfem_sim_create.argtypes = [C.POINTER(SimParams), C.POINTER(vp)]
        for f in ("mfem_sim_destroy", "mfem_sim_step", "mfem_sim_update_system"):
            getattr(L, f).argtypes = [vp]
        L.mfem_sim_solve.argtypes = [vp, dp]
        L.mfem_sim_energy.argtypes = [vp, C.c_double, dp]
        L.mfem_sim_stats.argtypes = [vp, C.POINTER(StepStats)]
        L.mfem_sim_get_positions.argtypes = [vp, dp]
        L.mfem_sim_size.argtypes = [vp, C.c_char_p]
        L.mfem_sim_size.restype = C.c_long
        L.mfem_sim_get.argtypes = [vp, C.c_char_p, dp]
        L.mfem_sim_get_int.argtypes = [vp, C.c_char_p, ip]
        L.mfem_sim_set_profiling.argtypes = [vp, C.c_int]
        L.mfem_sim_matrix_info.argtypes = [vp, ip, ip, ip]
        L.mfem_sim_time_spmv.argtypes = [vp, C.c_int, dp]
        L.mfem_sim_spmv_tuning.argtypes = [vp, dp, ip, ip, ip, ip, ip]
        L.mfem_sim_set_cg_shape.argtypes = [vp, C.c_int, C.c_int]
        L.mfem_sim_set_element_kernel.argtypes = [vp, C.c_int]
        L.mfem_sim_cg_tuning.argtypes = [vp, dp, ip, ip, ip, ip, ip, ip, ip, ip]
        L.mfem_sim_kkt_size.argtypes = [vp, C.POINTER(C.c_long)]
        L.mfem_sim_kkt_apply.argtypes = [vp, C.c_int, dp, dp]
        L.mfem_material_s6_batch.argtypes = [C.c_int, vp, vp, vp, vp, C.c_int, vp, vp, vp, vp]
        L.mfem_snh_f9_batch.argtypes = [C.c_int, vp, vp, vp, vp, vp, vp, vp]
        L.mfem_polar_batch.argtypes = [C.c_int, vp, vp, vp, vp, vp, vp]
        L.mfem_batch_create.argtypes = [vp, C.c_int, vp]
        L.mfem_batch_destroy.argtypes = [vp]
        L.mfem_batch_step.argtypes = [vp]
        L.mfem_batch_sizes.argtypes = [vp, ip, ip, ip]
        L.mfem_batch_get_positions.argtypes = [vp, C.c_int, dp]
        L.mfem_batch_stats.argtypes = [vp, ip, ip, ip, ip, dp, dp]
        L.mfem_batch_set_cg_shape.argtypes = [vp, C.c_int, C.c_int]
        L.mfem_batch_cg_shape.argtypes = [vp, ip, ip]
        L.mfem_batch_set_graphs.argtypes = [vp, C.c_int]
        L.mfem_batch_launch_counts.argtypes = [vp, C.POINTER(C.c_long), C.POINTER(C.c_long),
                                               C.POINTER(C.c_long)]
        L.mfem_sim_set_callback.argtypes = [vp, NEWTON_CALLBACK, vp]
        L.mfem_pcg_bsr3.argtypes = [C.c_int, C.c_int, vp, vp, vp, vp, vp, C.c_double, C.c_int,
                                    ip, dp, vp]
        L.mfem_pcg_bsr3_noalloc.argtypes = L.mfem_pcg_bsr3.argtypes
        _lib = L
    return _lib


def check(rc):
    if rc != 0:
        raise MfemError(lib().mfem_last_error().decode())


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _params(sc, precision, p=None):
    """mfem_sim_params for a Scene; returns (params, arrays to keep alive)."""
    keep = dict(V=_f64(sc.V), T=_i32(sc.T), mu=_f64(sc.mu), la=_f64(sc.la), mat=_i32(sc.mat),
                fixed=_i32(sc.fixed), solver=sc.solver.encode())
    bc = getattr(sc, "bc", None)
    p = SimParams() if p is None else p
    p.nV, p.nT = sc.nV, sc.nT
    p.V = keep["V"].ctypes.data_as(dp)
    p.T = keep["T"].ctypes.data_as(ip)
    p.mu = keep["mu"].ctypes.data_as(dp)
    p.la = keep["la"].ctypes.data_as(dp)
    p.mat = keep["mat"].ctypes.data_as(ip)
    if bc:  # the library derives the pins from the boundary condition (and moves them)
        from .scenes import BCS
        p.fixed = ip()
        p.bc_type, p.bc_axis = BCS[bc["type"]], int(bc.get("axis", 0))
        p.bc_ratio, p.bc_velocity = float(bc.get("ratio", 0.1)), float(bc.get("velocity", 0.1))
    else:
        p.fixed = keep["fixed"].ctypes.data_as(ip)
    if sc.v0 is not None:
        keep["v0"] = _f64(sc.v0)
        p.v0 = keep["v0"].ctypes.data_as(dp)
    else:
        p.v0 = dp()
    if getattr(sc, "x0", None) is not None:
        keep["x0"] = _f64(sc.x0)
        p.x0 = keep["x0"].ctypes.data_as(dp)
    else:
        p.x0 = dp()
    p.density, p.dt = sc.density, sc.dt
    g = list(sc.gravity) + [0.0] * 3
    p.gravity = (C.c_double * 3)(*g[:3])
    p.bdf_order = sc.bdf_order
    p.mixed = int(sc.mixed)
    p.linear_solver = keep["solver"]
    p.itr_tol, p.max_iters = sc.itr_tol, sc.max_iters
    p.max_newton, p.newton_tol, p.ls_iters = sc.max_newton, sc.newton_tol, sc.ls_iters
    p.precision = {"fp64": 0, "mixed": 1}[precision]
    p.dim = sc.dim
    return p, keep


class Sim:
    """One simulation on the GPU: SimState<3> + NewtonOptimizer<3> behind the C ABI."""

    def __init__(self, scene=None, precision: str = "fp64", json_path: str | None = None,
                 dim: int = 3):
        """From a Scene, or from a scene JSON file (dim 2: the 2D
        application's loader)."""
        self.h = C.c_void_p()
        self.scene = scene
        if json_path is not None:
            create = lib().mfem_sim_create_json_2d if dim == 2 else lib().mfem_sim_create_json
            check(create(json_path.encode(), C.byref(self.h)))
            self.nV = int(lib().mfem_sim_size(self.h, b"x") // 3)
            self.dim = dim
            return
        p, self._keep = _params(scene, precision)
        check(lib().mfem_sim_create(C.byref(p), C.byref(self.h)))
        self.nV = scene.nV
        self.dim = scene.dim

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            lib().mfem_sim_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self):
        check(lib().mfem_sim_step(self.h))

    def set_callback(self, fn):
        """fn(iteration) after every Newton iteration of step() (the
        reference's Optimizer::callback); None removes it."""
        if fn is None:
            self._cb = None
            check(lib().mfem_sim_set_callback(self.h, NEWTON_CALLBACK(), None))
            return
        self._cb = NEWTON_CALLBACK(lambda _sim, it, _user: fn(it))
        check(lib().mfem_sim_set_callback(self.h, self._cb, None))

    HISTORY_COLUMNS = ("iteration", "energy", "energy_res", "newton_dec", "alpha", "update_ms",
                       "solve_ms", "ls_ms", "solver_iters", "ls_halvings")

    def newton_history(self):
        """The last step's Newton iterations: an (iterations x 10) array with
        the columns of HISTORY_COLUMNS."""
        n = C.c_int(0)
        check(lib().mfem_sim_newton_history(self.h, dp(), C.byref(n)))
        out = np.zeros((n.value, 10))
        check(lib().mfem_sim_newton_history(self.h, out.ctypes.data_as(dp), C.byref(n)))
        return out

    # per-step output (the reference application's OBJ / .mesh / DMAT export)
    def write_obj(self, path):
        check(lib().mfem_sim_write_obj(self.h, str(path).encode()))

    def write_mesh(self, path):
        check(lib().mfem_sim_write_mesh(self.h, str(path).encode()))

    def write_state(self, x_path=None, v_path=None):
        enc = lambda p: None if p is None else str(p).encode()
        check(lib().mfem_sim_write_state(self.h, enc(x_path), enc(v_path)))

    def update_system(self):
        check(lib().mfem_sim_update_system(self.h))

    def solve(self):
        d = C.c_double(0)
        check(lib().mfem_sim_solve(self.h, C.byref(d)))
        return d.value

    def energy(self, alpha):
        e = C.c_double(0)
        check(lib().mfem_sim_energy(self.h, float(alpha), C.byref(e)))
        return e.value

    def stats(self):
        st = StepStats()
        check(lib().mfem_sim_stats(self.h, C.byref(st)))
        return {k: getattr(st, k) for k, _ in StepStats._fields_}

    def positions(self):
        d = getattr(self, "dim", 3)
        out = np.zeros(self.nV * d)
        check(lib().mfem_sim_get_positions(self.h, out.ctypes.data_as(dp)))
        return out.reshape(-1, d)

    def get(self, name):
        n = lib().mfem_sim_size(self.h, name.encode())
        if n < 0:
            raise MfemError(lib().mfem_last_error().decode())
        out = np.zeros(n)
        check(lib().mfem_sim_get(self.h, name.encode(), out.ctypes.data_as(dp)))
        return out

    def get_int(self, name):
        n = lib().mfem_sim_size(self.h, name.encode())
        if n < 0:
            raise MfemError(lib().mfem_last_error().decode())
        out = np.zeros(n, dtype=np.int32)
        check(lib().mfem_sim_get_int(self.h, name.encode(), out.ctypes.data_as(ip)))
        return out

    def set_profiling(self, flags=1):
        check(lib().mfem_sim_set_profiling(self.h, int(flags)))

    def matrix_info(self):
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        check(lib().mfem_sim_matrix_info(self.h, C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, c.value

    def time_spmv(self, reps=50):
        ms = C.c_double()
        check(lib().mfem_sim_time_spmv(self.h, int(reps), C.byref(ms)))
        return ms.value

    def spmv_tuning(self):
        """(per-shape mean ms, chosen shape index, (lanes, unroll, nontemporal))."""
        n = C.c_int()
        check(lib().mfem_sim_spmv_tuning(self.h, dp(), C.byref(n), None, None, None, None))
        t = np.zeros(n.value)
        ch, G, U, NT = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        check(lib().mfem_sim_spmv_tuning(self.h, t.ctypes.data_as(dp), C.byref(n), C.byref(ch),
                                         C.byref(G), C.byref(U), C.byref(NT)))
        return t, ch.value, (G.value, U.value, NT.value)

    def set_element_kernel(self, variant):
        """0: stiffness-block products on the vector units (default); 1: on
        the matrix cores (FP32 storage, 3D)."""
        check(lib().mfem_sim_set_element_kernel(self.h, int(variant)))

    def set_cg_shape(self, variant, blocks=0):
        """Force the CG kernel shape and grid (variant -1: autotune again)."""
        check(lib().mfem_sim_set_cg_shape(self.h, int(variant), int(blocks)))

    def kkt_size(self):
        n = C.c_long(0)
        check(lib().mfem_sim_kkt_size(self.h, C.byref(n)))
        return n.value

    def kkt_apply(self, which, v=None):
        """KKT system of "minres-indefinite-block" at the current state: which
        0 -> K v, 1 -> P v (block preconditioner), 2 -> right-hand side."""
        N = self.kkt_size()
        out = np.zeros(N)
        vin = _f64(v) if v is not None else np.zeros(N)
        check(lib().mfem_sim_kkt_apply(self.h, int(which), vin.ctypes.data_as(dp),
                                       out.ctypes.data_as(dp)))
        return out

    def cg_tuning(self):
        """List of (variant, grid, ms per CG iteration) the tuner timed, the
        chosen index, and the chosen (lanes, blocks in flight, workgroup size,
        grid)."""
        n = C.c_int()
        check(lib().mfem_sim_cg_tuning(self.h, None, None, None, C.byref(n), None, None, None,
                                       None, None))
        t = np.zeros(n.value)
        v = np.zeros(n.value, dtype=np.int32)
        g = np.zeros(n.value, dtype=np.int32)
        ch, G, U, BS, NB = C.c_int(), C.c_int(), C.c_int(), C.c_int(), C.c_int()
        check(lib().mfem_sim_cg_tuning(self.h, t.ctypes.data_as(dp), v.ctypes.data_as(ip),
                                       g.ctypes.data_as(ip), C.byref(n), C.byref(ch), C.byref(G),
                                       C.byref(U), C.byref(BS), C.byref(NB)))
        trials = [(int(a), int(b), float(c)) for a, b, c in zip(v, g, t)]
        return trials, ch.value, (G.value, U.value, BS.value, NB.value)


class Batch:
    """Independent mixed-stretch simulations advanced together on the GPU
    (mfem_batch_*): one launch per Newton phase for all of them, each with
    its own CG scalars, line search and Newton termination."""

    def __init__(self, scenes, precision: str = "mixed"):
        self.h = C.c_void_p()
        params = (SimParams * len(scenes))()
        self._keep = []
        for k, sc in enumerate(scenes):
            _, keep = _params(sc, precision, params[k])
            self._keep.append(keep)
        check(lib().mfem_batch_create(params, len(scenes), C.byref(self.h)))
        self.count = len(scenes)
        nV = np.zeros(self.count, dtype=np.int32)
        nT = np.zeros(self.count, dtype=np.int32)
        cnt = C.c_int()
        check(lib().mfem_batch_sizes(self.h, C.byref(cnt), nV.ctypes.data_as(ip), nT.ctypes.data_as(ip)))
        self.nV, self.nT = nV, nT

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            lib().mfem_batch_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self):
        check(lib().mfem_batch_step(self.h))

    def positions(self, k):
        out = np.zeros(int(self.nV[k]) * 3)
        check(lib().mfem_batch_get_positions(self.h, int(k), out.ctypes.data_as(dp)))
        return out.reshape(-1, 3)

    def stats(self):
        n = self.count
        ni, pi, rf, lf = (np.zeros(n, dtype=np.int32) for _ in range(4))
        dec = np.zeros(n)
        ms = C.c_double()
        check(lib().mfem_batch_stats(self.h, ni.ctypes.data_as(ip), pi.ctypes.data_as(ip),
                                     rf.ctypes.data_as(ip), lf.ctypes.data_as(ip),
                                     dec.ctypes.data_as(dp), C.byref(ms)))
        return dict(newton_iters=ni, pcg_iters=pi, refinements=rf, ls_fail=lf, decrement=dec,
                    step_ms=ms.value)

    def set_cg_shape(self, variant, rows_per_workgroup=256):
        check(lib().mfem_batch_set_cg_shape(self.h, int(variant), int(rows_per_workgroup)))

    def cg_shape(self):
        v, r = C.c_int(), C.c_int()
        check(lib().mfem_batch_cg_shape(self.h, C.byref(v), C.byref(r)))
        return v.value, r.value

    def set_graphs(self, on):
        check(lib().mfem_batch_set_graphs(self.h, int(bool(on))))

    def launch_counts(self):
        """(update-phase submissions, CG submissions, Newton iterations)."""
        a, b, c = C.c_long(), C.c_long(), C.c_long()
        check(lib().mfem_batch_launch_counts(self.h, C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, c.value


# ---- batched element entry points (torch tensors as device memory) ---------

def _dev(a, dtype):
    import torch
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype).cuda()


def _stream():
    import torch
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def polar_batch(F, R0=None):
    """F: (n,3,3). Returns R (n,3,3), s (n,6), dsdF (n,6,9), iterations (n)."""
    import torch
    F = np.asarray(F, dtype=np.float64)
    n = F.shape[0]
    Fc = _dev(np.transpose(F, (0, 2, 1)).reshape(n, 9), torch.float64)
    R = np.tile(np.eye(3), (n, 1, 1)) if R0 is None else np.asarray(R0, dtype=np.float64)
    Rc = _dev(np.transpose(R, (0, 2, 1)).reshape(n, 9), torch.float64)
    s = torch.zeros((n, 6), dtype=torch.float64, device="cuda")
    d = torch.zeros((n, 54), dtype=torch.float64, device="cuda")
    it = torch.zeros(n, dtype=torch.int32, device="cuda")
    check(lib().mfem_polar_batch(n, C.c_void_p(Fc.data_ptr()), C.c_void_p(Rc.data_ptr()),
                                 C.c_void_p(s.data_ptr()), C.c_void_p(d.data_ptr()),
                                 C.c_void_p(it.data_ptr()), _stream()))
    torch.cuda.synchronize()
    Rh = Rc.cpu().numpy().reshape(n, 3, 3).transpose(0, 2, 1)
    dsdF = d.cpu().numpy().reshape(n, 9, 6).transpose(0, 2, 1)
    return Rh, s.cpu().numpy(), dsdF, it.cpu().numpy()


def material_s6_batch(mtype, mu, la, s, psd_fix=True):
    """s: (n,6). Returns E (n), g (n,6), H (n,6,6)."""
    import torch
    s = np.asarray(s, dtype=np.float64)
    n = s.shape[0]
    t = _dev(np.broadcast_to(np.asarray(mtype, dtype=np.int32), (n,)), torch.int32)
    m = _dev(np.broadcast_to(np.asarray(mu, dtype=np.float64), (n,)), torch.float64)
    l = _dev(np.broadcast_to(np.asarray(la, dtype=np.float64), (n,)), torch.float64)
    sd = _dev(s, torch.float64)
    E = torch.zeros(n, dtype=torch.float64, device="cuda")
    g = torch.zeros((n, 6), dtype=torch.float64, device="cuda")
    H = torch.zeros((n, 36), dtype=torch.float64, device="cuda")
    check(lib().mfem_material_s6_batch(n, C.c_void_p(t.data_ptr()), C.c_void_p(m.data_ptr()),
                                       C.c_void_p(l.data_ptr()), C.c_void_p(sd.data_ptr()),
                                       int(psd_fix), C.c_void_p(E.data_ptr()),
                                       C.c_void_p(g.data_ptr()), C.c_void_p(H.data_ptr()),
                                       _stream()))
    torch.cuda.synchronize()
    return E.cpu().numpy(), g.cpu().numpy(), H.cpu().numpy().reshape(n, 6, 6).transpose(0, 2, 1)


def snh_f9_batch(mu, la, F9):
    """F9: (n,9) column-major vec(F). Returns E (n), g (n,9), H (n,9,9) PSD-projected."""
    import torch
    F9 = np.asarray(F9, dtype=np.float64)
    n = F9.shape[0]
    m = _dev(np.broadcast_to(np.asarray(mu, dtype=np.float64), (n,)), torch.float64)
    l = _dev(np.broadcast_to(np.asarray(la, dtype=np.float64), (n,)), torch.float64)
    Fd = _dev(F9, torch.float64)
    E = torch.zeros(n, dtype=torch.float64, device="cuda")
    g = torch.zeros((n, 9), dtype=torch.float64, device="cuda")
    H = torch.zeros((n, 81), dtype=torch.float64, device="cuda")
    check(lib().mfem_snh_f9_batch(n, C.c_void_p(m.data_ptr()), C.c_void_p(l.data_ptr()),
                                  C.c_void_p(Fd.data_ptr()), C.c_void_p(E.data_ptr()),
                                  C.c_void_p(g.data_ptr()), C.c_void_p(H.data_ptr()), _stream()))
    torch.cuda.synchronize()
    return E.cpu().numpy(), g.cpu().numpy(), H.cpu().numpy().reshape(n, 9, 9).transpose(0, 2, 1)


def boundary_faces(V, T):
    """Boundary triangles of a tetrahedral mesh (host code of the library)."""
    V, T = _f64(V), _i32(T)
    n = C.c_int(0)
    check(lib().mfem_boundary_faces(V.ctypes.data_as(dp), V.shape[0], T.ctypes.data_as(ip),
                                    T.shape[0], ip(), C.byref(n)))
    F = np.zeros((n.value, 3), dtype=np.int32)
    check(lib().mfem_boundary_faces(V.ctypes.data_as(dp), V.shape[0], T.ctypes.data_as(ip),
                                    T.shape[0], F.ctypes.data_as(ip), C.byref(n)))
    return F


def pcg_bsr3(rowptr, col, vals, b, tol, max_iters, noalloc=False, x=None):
    """mfem_pcg_bsr3 (or, with noalloc, mfem_pcg_bsr3_noalloc) on torch device
    tensors (rowptr, col int32; vals 9 nnzb and b 3 nb float64; x, 3 nb
    float64, is allocated when not given): returns (x, iterations, relative
    residual)."""
    import torch
    nb = int(rowptr.numel()) - 1
    if x is None:
        x = torch.empty(3 * nb, dtype=torch.float64, device=b.device)
    it = C.c_int(0)
    rr = C.c_double(0)
    fn = lib().mfem_pcg_bsr3_noalloc if noalloc else lib().mfem_pcg_bsr3
    check(fn(nb, int(col.numel()), rowptr.data_ptr(), col.data_ptr(), vals.data_ptr(),
             b.data_ptr(), x.data_ptr(), float(tol), int(max_iters), C.byref(it), C.byref(rr),
             _stream()))
    return x, it.value, rr.value
