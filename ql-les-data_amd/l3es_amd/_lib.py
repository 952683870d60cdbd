"""ctypes binding of libqlles.so (declared in include/qlles.h).

The library is loaded from ``ql-les-data_amd/lib/libqlles.so`` (built by
``__graft_entry__.build()`` or ``make -C ql-les-data_amd/csrc``).  There is no
fallback: if the library is missing every entry point raises.
"""

import ctypes
import os

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
# QLLES_LIB names another build of the same library (e.g. one compiled with other tuning macros)
LIB_PATH = os.environ.get("QLLES_LIB") or os.path.normpath(os.path.join(_HERE, "..", "lib", "libqlles.so"))

QLLES_OK = 0
QLLES_ERR_ARG = 1
QLLES_ERR_SHAPE = 2
QLLES_ERR_HIP = 3
QLLES_ERR_NONFINITE = 4

QLLES_F32, QLLES_F64 = 0, 1
FORCING = {"none": 0, "ekin_tot": 1, "power": 2, "kolmogorov": 3}
VISCOUS = {"explicit": 0, "IF": 1}
INTERP = {"nufft": 0, "dft": 1}

# every function exported through include/qlles.h
EXPORTS = (
    "qlles_plan_create", "qlles_plan_destroy", "qlles_plan_workspace_bytes", "qlles_step",
    "qlles_rhs", "qlles_rfftn", "qlles_irfftn", "qlles_diagnostics", "qlles_dissipation",
    "qlles_energy_hat", "qlles_filter", "qlles_interp_create", "qlles_interp_destroy",
    "qlles_interp_field", "qlles_interp_hat", "qlles_last_error", "qlles_profile", "qlles_profile_read",
    "qlles_step_advect", "qlles_batch_create", "qlles_batch_destroy", "qlles_batch_add",
    "qlles_batch_configure", "qlles_batch_tune", "qlles_batch_get_config", "qlles_batch_kinds", "qlles_batch_run",
    "qlles_rhs_jvp", "qlles_rhs_vjp", "qlles_step_jvp", "qlles_step_vjp", "qlles_irfftn_adjoint",
    "qlles_rfftn_adjoint", "qlles_interp_hat_vjp", "qlles_plan_set_variant", "qlles_plan_get_variant",
    "qlles_plan_tune", "qlles_energy_spectrum", "qlles_sph_create", "qlles_sph_destroy", "qlles_sph_density",
    "qlles_sph_relax_acc", "qlles_sph_forces", "qlles_sph_relax_advect", "qlles_sph_mls_grid", "qlles_sph_dft_hat",
    "qlles_step_record_bytes", "qlles_step_record", "qlles_step_vjp_record",
    "qlles_stats_create", "qlles_stats_destroy", "qlles_stats_moments", "qlles_stats_moments_noalloc",
)
KINDS = ("x-pass", "y-inverse", "z-pass", "y-forward", "other")


class QllesError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"qlles error {code}: {msg}")
        self.code = code


class Desc(ctypes.Structure):
    _fields_ = [
        ("dim", ctypes.c_int), ("N", ctypes.c_int), ("dtype", ctypes.c_int),
        ("nu", ctypes.c_double), ("dt", ctypes.c_double),
        ("forcing", ctypes.c_int), ("kf", ctypes.c_int),
        ("e_kin_init", ctypes.c_double), ("power", ctypes.c_double),
        ("kolm_k", ctypes.c_int), ("kolm_amp", ctypes.c_double), ("drag", ctypes.c_double),
        ("viscous", ctypes.c_int),
    ]


_lib = None


def load():
    """Load libqlles.so once; raises if it is missing (no CPU fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} not found: build it with __graft_entry__.build() "
                          "or `make -C ql-les-data_amd/csrc`")
    lib = ctypes.CDLL(LIB_PATH)
    vp, ci, cd = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
    pd = ctypes.POINTER(ctypes.c_double)
    sig = {
        "qlles_plan_create": (ci, [ctypes.POINTER(Desc), ctypes.POINTER(vp)]),
        "qlles_plan_destroy": (ci, [vp]),
        "qlles_plan_workspace_bytes": (ctypes.c_size_t, [vp]),
        "qlles_step": (ci, [vp, vp, vp, vp, ci, vp, vp]),
        "qlles_rhs": (ci, [vp, vp, vp, ci, vp, vp]),
        "qlles_rfftn": (ci, [vp, vp, vp, vp]),
        "qlles_irfftn": (ci, [vp, vp, vp, vp]),
        "qlles_diagnostics": (ci, [vp, vp, vp, vp]),
        "qlles_dissipation": (ci, [vp, vp, vp, vp]),
        "qlles_energy_hat": (ci, [vp, vp, vp, vp]),
        "qlles_filter": (ci, [vp, vp, ci, vp, vp]),
        "qlles_interp_create": (ci, [ci, ci, ci, ci, cd, ctypes.POINTER(vp)]),
        "qlles_interp_destroy": (ci, [vp]),
        "qlles_interp_field": (ci, [vp, vp, vp, ci, vp, vp]),
        "qlles_interp_hat": (ci, [vp, ci, vp, vp, ci, vp, cd, vp]),
        "qlles_last_error": (ctypes.c_char_p, []),
        "qlles_profile": (ci, [vp, ci]),
        "qlles_step_advect": (ci, [vp, vp, vp, vp, ci, vp, ci, vp, vp, vp]),
        "qlles_profile_read": (ci, [vp, vp, vp, vp]),
        "qlles_batch_create": (ci, [ctypes.POINTER(vp)]),
        "qlles_batch_destroy": (ci, [vp]),
        "qlles_batch_add": (ci, [vp, vp, vp, vp, vp, ci, vp, ctypes.POINTER(ci)]),
        "qlles_batch_configure": (ci, [vp, ci, ci, ci]),
        "qlles_batch_tune": (ci, [vp, ci, vp]),
        "qlles_batch_get_config": (ci, [vp, ctypes.POINTER(ci), ctypes.POINTER(ci), ctypes.POINTER(ci), pd]),
        "qlles_batch_kinds": (ci, [vp, ctypes.POINTER(ci), vp, vp, ci]),
        "qlles_batch_run": (ci, [vp, vp, ci, vp]),
        "qlles_rhs_jvp": (ci, [vp, vp, vp, vp, vp]),
        "qlles_rhs_vjp": (ci, [vp, vp, vp, vp, vp, vp]),
        "qlles_step_jvp": (ci, [vp, vp, vp, vp, vp]),
        "qlles_step_vjp": (ci, [vp, vp, vp, vp, vp, vp]),
        "qlles_irfftn_adjoint": (ci, [vp, vp, vp, vp]),
        "qlles_rfftn_adjoint": (ci, [vp, vp, vp, vp]),
        "qlles_interp_hat_vjp": (ci, [vp, ci, vp, vp, ci, vp, vp, vp, vp]),
        "qlles_plan_set_variant": (ci, [vp, ci, ci]),
        "qlles_plan_get_variant": (ci, [vp, ctypes.POINTER(ci), ctypes.POINTER(ci), ctypes.POINTER(ci)]),
        "qlles_plan_tune": (ci, [vp, vp, ci, pd, vp]),
        "qlles_energy_spectrum": (ci, [vp, vp, cd, vp, vp]),
        "qlles_sph_create": (ci, [ci, ci, cd, ci, cd, ctypes.POINTER(vp)]),
        "qlles_sph_destroy": (ci, [vp]),
        "qlles_sph_density": (ci, [vp, vp, ci, vp, vp]),
        "qlles_sph_relax_acc": (ci, [vp, vp, ci, vp, vp, vp]),
        "qlles_sph_forces": (ci, [vp, vp, vp, ci, vp, vp, vp, vp]),
        "qlles_sph_relax_advect": (ci, [vp, vp, ci, vp, ci, cd, cd, vp, vp]),
        "qlles_sph_mls_grid": (ci, [vp, vp, vp, ci, vp, vp]),
        "qlles_sph_dft_hat": (ci, [vp, vp, vp, ci, vp, vp]),
        "qlles_step_record_bytes": (ctypes.c_size_t, [vp]),
        "qlles_step_record": (ci, [vp, vp, vp, vp, vp]),
        "qlles_step_vjp_record": (ci, [vp, vp, vp, vp, vp, vp, vp]),
        "qlles_stats_create": (ci, [ci, cd, ctypes.POINTER(vp)]),
        "qlles_stats_destroy": (ci, [vp]),
        "qlles_stats_moments": (ci, [vp, vp, vp, ctypes.c_longlong, ctypes.c_longlong, vp, vp, vp]),
        "qlles_stats_moments_noalloc": (ci, [ci, cd, vp, vp, ctypes.c_longlong, ctypes.c_longlong, vp, vp, vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    del pd
    _lib = lib
    return lib


def check(code):
    if code != QLLES_OK:
        msg = _lib.qlles_last_error().decode(errors="replace")
        raise QllesError(code, msg)


def ptr(t):
    """Device pointer of a tensor (None -> NULL)."""
    if t is None:
        return None
    if not t.is_cuda:
        raise ValueError("qlles entry points take device tensors")
    if not t.is_contiguous():
        raise ValueError("qlles entry points take contiguous tensors")
    return ctypes.c_void_p(t.data_ptr())


def stream_ptr(stream=None):
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


class Plan:
    """Owner of a qlles_plan (one solver configuration + device workspace)."""

    def __init__(self, dim, N, dtype, nu, dt, forcing="none", kf=0, e_kin_init=0.0, power=0.0,
                 kolm_k=0, kolm_amp=0.0, drag=0.0, viscous="explicit"):
        lib = load()
        self.desc = Desc(dim, N, QLLES_F32 if dtype == torch.float32 else QLLES_F64, float(nu), float(dt),
                         FORCING[forcing], int(kf or 0), float(e_kin_init or 0.0), float(power or 0.0),
                         int(kolm_k or 0), float(kolm_amp or 0.0), float(drag or 0.0), VISCOUS[viscous])
        h = ctypes.c_void_p()
        check(lib.qlles_plan_create(ctypes.byref(self.desc), ctypes.byref(h)))
        self.handle = h
        self.dim, self.N, self.dtype = dim, N, dtype
        self.cdtype = torch.complex64 if dtype == torch.float32 else torch.complex128

    @property
    def workspace_bytes(self):
        return int(_lib.qlles_plan_workspace_bytes(self.handle))

    def variant(self):
        """(xv, zv, number of choices per knob) of the kernel variants in use."""
        xv, zv, nv = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        check(_lib.qlles_plan_get_variant(self.handle, ctypes.byref(xv), ctypes.byref(zv), ctypes.byref(nv)))
        return xv.value, zv.value, nv.value

    def set_variant(self, xv, zv):
        check(_lib.qlles_plan_set_variant(self.handle, int(xv), int(zv)))

    def tune(self, u_hat, trial_steps=8):
        """Time every kernel variant on a scratch copy of u_hat and keep the fastest; returns us/step."""
        us = ctypes.c_double()
        check(_lib.qlles_plan_tune(self.handle, ptr(u_hat), int(trial_steps), ctypes.byref(us), stream_ptr()))
        return us.value

    def close(self):
        if getattr(self, "handle", None) is not None and _lib is not None:
            _lib.qlles_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Sph:
    """Owner of a qlles_sph (SPH density / relaxation context on an n^dim particle lattice)."""

    def __init__(self, dim, n, box, is_physical, u_ref):
        lib = load()
        h = ctypes.c_void_p()
        check(lib.qlles_sph_create(dim, n, float(box), int(bool(is_physical)), float(u_ref), ctypes.byref(h)))
        self.handle = h
        self.dim, self.n = dim, n

    def close(self):
        if getattr(self, "handle", None) is not None and _lib is not None:
            _lib.qlles_sph_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Stats:
    """Owner of a qlles_stats (trajectory moments for the dataset statistics)."""

    def __init__(self, dim, box):
        lib = load()
        h = ctypes.c_void_p()
        check(lib.qlles_stats_create(dim, float(box), ctypes.byref(h)))
        self.handle = h
        self.dim = dim

    def close(self):
        if getattr(self, "handle", None) is not None and _lib is not None:
            _lib.qlles_stats_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Interp:
    """Owner of a qlles_interp (particle interpolator on an n-grid)."""

    def __init__(self, dim, n, dtype, backend="nufft", tol=1e-6):
        lib = load()
        h = ctypes.c_void_p()
        check(lib.qlles_interp_create(dim, n, QLLES_F32 if dtype == torch.float32 else QLLES_F64,
                                      INTERP[backend], float(tol), ctypes.byref(h)))
        self.handle = h
        self.dim, self.n, self.dtype, self.backend = dim, n, dtype, backend

    def close(self):
        if getattr(self, "handle", None) is not None and _lib is not None:
            _lib.qlles_interp_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
