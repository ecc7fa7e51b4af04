"""Python host side of the MI355X sma-dft time stepper.

Thin ctypes binding over the C ABI in ``include/smadft/smadft.h``
(``libsmadft.so``, built in-tree by ``make -C sma-dft_amd``).  It mirrors the
reference programs' step loop (``while (L1 > limL1)``,
no-adv/cosNoAdvTrackMult.cpp:724-1147): build an initial condition, create a
:class:`Stepper`, advance it, check L1 every 100 steps, fetch fields.

There is no CPU fallback: if the shared library is missing, or no GPU is
visible when a context is created, the calls raise.

With torch in the same process (the autograd wrapper, device buffers),
import torch before the first context is created: torch's wheel carries its
own HIP runtime, which must load before the one the library links.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass, fields as dc_fields

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsmadft.so")

MODEL_PLAIN, MODEL_ADV, MODEL_QUASI, MODEL_PLAIN_CONSERVED, MODEL_ADV_CONSERVED = 0, 1, 2, 3, 4
MODEL_COMPRESSIBLE = 5
FP64, FP32_MIXED = 64, 32

FIELD_PSI, FIELD_PSI_HAT, FIELD_N_HAT = 0, 1, 2
FIELD_GRAD_X, FIELD_GRAD_Y, FIELD_GRAD_Z = 3, 4, 5
FIELD_VEL_X, FIELD_VEL_Y, FIELD_VEL_Z = 6, 7, 8
FIELD_PRESSURE, FIELD_RHO, FIELD_ENERGY, FIELD_DIVV_HAT = 9, 10, 11, 12
FIELD_PSI_HAT_OLD, FIELD_RHO_OLD = 13, 14
FIELD_GRAD_RHO_X, FIELD_GRAD_RHO_Y, FIELD_GRAD_RHO_Z = 15, 16, 17
FIELD_SPEC_XY = 18
FIELD_VSOL_X, FIELD_VSOL_Y, FIELD_VSOL_Z = 19, 20, 21
FIELD_FPROJ_X, FIELD_FPROJ_Y, FIELD_FPROJ_Z = 22, 23, 24
FIELD_FORCE_X, FIELD_FORCE_Y, FIELD_FORCE_Z = 25, 26, 27
FIELD_DIVV = 28
FIELD_DRHO_OLD = 29
FIELD_F_RHO, FIELD_AMP = 30, 31
FIELD_VIRR_X, FIELD_VIRR_Y, FIELD_VIRR_Z = 32, 33, 34
FIELD_CN_HIST = 35  # plain: the combined CN/AB2 history H = C1 psi_hat - (dt/2) s N_hat_past

RED_L1, RED_MASS, RED_ENERGY = 0, 1, 2

OK, EINVAL, EHIP, ENOMEM, ENONFINITE, ESTATE = 0, -1, -2, -3, -4, -5

# Symbols the C ABI exports (checked by the CPU tests against the header).
EXPORTS = (
    "smad_init", "smad_step", "smad_fetch", "smad_set_field", "smad_reduce",
    "smad_sync", "smad_stream", "smad_steps_taken", "smad_destroy",
    "smad_last_error", "smad_ic_focal_conic", "smad_ic_slab", "smad_ic_curved",
    "smad_ic_add_noise", "smad_default_params", "smad_r2r_3d", "smad_time_steps",
    "smad_profile_steps", "smad_kernel_class_name", "smad_forward_tape", "smad_adjoint", "smad_adjoint_timing", "smad_set_tape_budget", "smad_surface",
    "smad_surface_envelope", "smad_r2r_3d_noalloc", "smad_noalloc_last_error",
    "smad_state_fields", "smad_set_steps_taken", "smad_ic_tilted_slab", "smad_ic_two_bumps",
    "smad_config_params", "smad_config_ic", "smad_forward_tape_dev", "smad_adjoint_dev",
    "smad_set_params", "smad_ic_fc_torus", "smad_stream_copy",
)

REDFT10, REDFT01, RODFT10, RODFT01 = 0, 1, 2, 3


class CParams(ctypes.Structure):
    """ctypes mirror of ``smad_params``."""

    _fields_ = [
        ("model", ctypes.c_int), ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
        ("nw", ctypes.c_double), ("q0", ctypes.c_double), ("eps", ctypes.c_double),
        ("alpha", ctypes.c_double), ("beta", ctypes.c_double), ("gamma", ctypes.c_double),
        ("dt", ctypes.c_double), ("nu", ctypes.c_double), ("lambda_", ctypes.c_double),
        ("rho0", ctypes.c_double), ("kp", ctypes.c_double), ("rho_m", ctypes.c_double),
        ("filter_sigma", ctypes.c_double), ("div_switch", ctypes.c_int),
        ("precision", ctypes.c_int), ("load", ctypes.c_int),
        ("zeta", ctypes.c_double), ("rho_up", ctypes.c_int),
    ]


@dataclass
class Params:
    """Model parameters (meaning as in the reference programs' constants)."""

    model: int = MODEL_PLAIN
    nx: int = 128
    ny: int = 128
    nz: int = 128
    nw: float = 8.0
    q0: float = 1.0
    eps: float = -0.7
    alpha: float = 1.0
    beta: float = 2.0
    gamma: float = 1.0
    dt: float = 0.1
    nu: float = 1.0
    lambda_: float = -2.0 / 3.0
    rho0: float = 1.0
    kp: float = 0.0
    rho_m: float = 0.5
    filter_sigma: float = 1.57
    div_switch: int = 200
    precision: int = FP64
    load: int = 0
    zeta: float = 0.001
    rho_up: int = 16

    def c(self) -> CParams:
        return CParams(**{f.name: getattr(self, f.name) for f in dc_fields(self)})

    @property
    def shape(self):
        return (self.nx, self.ny, self.nz)

    @classmethod
    def default(cls, model: int, n: int) -> "Params":
        """The reference program's own constants for the model."""
        cp = CParams()
        lib().smad_default_params(model, n, ctypes.byref(cp))
        return cls(**{f.name: getattr(cp, f.name) for f in dc_fields(cls)})

    @classmethod
    def config(cls, config: int, n: int | None = None) -> "Params":
        """The benchmark configuration's parameters (0-4), optionally on an
        n^3 grid instead of its own."""
        cp = CParams()
        rc = lib().smad_config_params(config, ctypes.byref(cp))
        if rc != OK:
            raise SmadError(rc, f"unknown configuration {config}")
        p = cls(**{f.name: getattr(cp, f.name) for f in dc_fields(cls)})
        if n is not None:
            p.nx = p.ny = p.nz = n
        return p


_LIB = None


def lib() -> ctypes.CDLL:
    """Load libsmadft.so (raises if it has not been built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `make -C sma-dft_amd` "
                               "(or __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.POINTER
        dp = P(ctypes.c_double)
        vp = ctypes.c_void_p
        L.smad_init.argtypes = [P(CParams), dp, P(vp)]
        L.smad_step.argtypes = [vp, ctypes.c_int]
        L.smad_fetch.argtypes = [vp, ctypes.c_int, dp]
        L.smad_set_field.argtypes = [vp, ctypes.c_int, dp]
        L.smad_reduce.argtypes = [vp, ctypes.c_int, dp]
        L.smad_sync.argtypes = [vp]
        L.smad_stream.argtypes = [vp]
        L.smad_stream.restype = vp
        L.smad_steps_taken.argtypes = [vp]
        L.smad_steps_taken.restype = ctypes.c_long
        L.smad_destroy.argtypes = [vp]
        L.smad_destroy.restype = None
        L.smad_last_error.argtypes = [vp]
        L.smad_last_error.restype = ctypes.c_char_p
        L.smad_ic_focal_conic.argtypes = [P(CParams), ctypes.c_double, ctypes.c_double, ctypes.c_double, dp]
        L.smad_ic_fc_torus.argtypes = [P(CParams)] + [ctypes.c_double] * 4 + [dp]
        L.smad_stream_copy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p]
        L.smad_ic_slab.argtypes = [P(CParams), ctypes.c_double, ctypes.c_double, ctypes.c_ulonglong, dp]
        L.smad_ic_curved.argtypes = [P(CParams), ctypes.c_double, ctypes.c_double, ctypes.c_ulonglong, dp]
        L.smad_ic_add_noise.argtypes = [P(CParams), ctypes.c_double, ctypes.c_ulonglong, dp]
        L.smad_default_params.argtypes = [ctypes.c_int, ctypes.c_int, P(CParams)]
        L.smad_config_params.argtypes = [ctypes.c_int, P(CParams)]
        L.smad_config_ic.argtypes = [ctypes.c_int, P(CParams), dp]
        L.smad_ic_two_bumps.argtypes = [P(CParams), ctypes.c_double, ctypes.c_double, ctypes.c_ulonglong, dp]
        L.smad_ic_tilted_slab.argtypes = [P(CParams), ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                          ctypes.c_ulonglong, dp, dp]
        L.smad_time_steps.argtypes = [vp, ctypes.c_int, dp]
        L.smad_forward_tape.argtypes = [vp, ctypes.c_int, ctypes.c_int]
        L.smad_adjoint.argtypes = [vp, dp, dp, dp]
        L.smad_set_tape_budget.argtypes = [vp, ctypes.c_double]
        L.smad_surface.argtypes = [vp, dp, dp, dp, dp]
        L.smad_surface_envelope.argtypes = [vp, dp, dp, dp, dp]
        L.smad_adjoint_timing.argtypes = [vp, dp, dp]
        L.smad_profile_steps.argtypes = [vp, ctypes.c_int, dp, P(ctypes.c_int), ctypes.c_int,
                                         P(ctypes.c_int)]
        L.smad_kernel_class_name.argtypes = [vp, ctypes.c_int]
        L.smad_kernel_class_name.restype = ctypes.c_char_p
        L.smad_r2r_3d.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, P(ctypes.c_int),
                                  ctypes.c_int, dp, dp]
        L.smad_forward_tape_dev.argtypes = [vp, vp, ctypes.c_int, ctypes.c_int, vp]
        L.smad_adjoint_dev.argtypes = [vp, vp, vp, dp]
        L.smad_set_params.argtypes = [vp, P(CParams)]
        L.smad_state_fields.argtypes = [vp, P(ctypes.c_int), ctypes.c_int]
        L.smad_set_steps_taken.argtypes = [vp, ctypes.c_long]
        L.smad_r2r_3d_noalloc.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, P(ctypes.c_int),
                                          ctypes.c_int, vp, vp, vp]
        L.smad_noalloc_last_error.argtypes = []
        L.smad_noalloc_last_error.restype = ctypes.c_char_p
        for fn in ("smad_ic_focal_conic", "smad_ic_slab", "smad_ic_curved", "smad_ic_add_noise",
                   "smad_default_params", "smad_ic_two_bumps", "smad_ic_tilted_slab", "smad_ic_fc_torus"):
            getattr(L, fn).restype = None
        _LIB = L
    return _LIB


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class SmadError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"smad error {code}: {msg}")
        self.code = code


# ------------------------------------------------------------------ ICs --

def ic_slab(p: Params, amp=0.6, noise=1e-3, seed=0) -> np.ndarray:
    out = np.empty(p.shape, dtype=np.float64)
    lib().smad_ic_slab(ctypes.byref(p.c()), amp, noise, seed, _dptr(out))
    return out


def ic_curved(p: Params, amp=0.6, noise=1e-3, seed=1) -> np.ndarray:
    out = np.empty(p.shape, dtype=np.float64)
    lib().smad_ic_curved(ctypes.byref(p.c()), amp, noise, seed, _dptr(out))
    return out


def ic_two_bumps(p: Params, amp=1.328, noise=1e-3, seed=2) -> np.ndarray:
    out = np.empty(p.shape, dtype=np.float64)
    lib().smad_ic_two_bumps(ctypes.byref(p.c()), amp, noise, seed, _dptr(out))
    return out


def ic_tilted_slab(p: Params, amp=1.34164, max_tilt_deg=15.0, noise=1e-3, seed=3):
    """Returns (psi, tilt in degrees)."""
    out = np.empty(p.shape, dtype=np.float64)
    t = ctypes.c_double()
    lib().smad_ic_tilted_slab(ctypes.byref(p.c()), amp, max_tilt_deg, noise, seed, _dptr(out), ctypes.byref(t))
    return out, t.value


def config_ic(config: int, p: Params) -> np.ndarray:
    """The benchmark configuration's initial condition on the grid of p."""
    out = np.empty(p.shape, dtype=np.float64)
    rc = lib().smad_config_ic(config, ctypes.byref(p.c()), _dptr(out))
    if rc != OK:
        raise SmadError(rc, f"unknown configuration {config}")
    return out


def ic_focal_conic(p: Params, amp=1.328, a_e=None, b_e=None) -> np.ndarray:
    a_e = p.nx * 0.4 if a_e is None else a_e
    b_e = a_e if b_e is None else b_e
    out = np.empty(p.shape, dtype=np.float64)
    lib().smad_ic_focal_conic(ctypes.byref(p.c()), amp, a_e, b_e, _dptr(out))
    return out


def ic_fc_torus(p: Params, amp=None, a_e=None, b_e=None, x_scale=1.0) -> np.ndarray:
    """Focal-conic torus of the compressible program (compressible/compressible.cpp:
    703-727); amp defaults to the smectic amplitude of p's eps, beta, gamma,
    a_e = b_e to 124/130 of nz as in the shipped run."""
    if amp is None:
        amp = np.sqrt((3 * p.beta + np.sqrt(9 * p.beta ** 2 + 40 * p.eps * p.gamma)) / 5)
    a_e = p.nz * 124.0 / 130.0 if a_e is None else a_e
    b_e = a_e if b_e is None else b_e
    out = np.empty(p.shape, dtype=np.float64)
    lib().smad_ic_fc_torus(ctypes.byref(p.c()), amp, a_e, b_e, x_scale, _dptr(out))
    return out


def stream_copy(dst_ptr: int, src_ptr: int, nbytes: int, stream: int = 0) -> None:
    """Streaming device copy (smad_stream_copy): the copy-ceiling kernel."""
    rc = lib().smad_stream_copy(dst_ptr, src_ptr, nbytes, stream or None)
    if rc != OK:
        raise SmadError(rc, lib().smad_last_error(None).decode())


def add_noise(p: Params, psi: np.ndarray, amp=1e-3, seed=0) -> np.ndarray:
    out = np.ascontiguousarray(psi, dtype=np.float64).copy()
    lib().smad_ic_add_noise(ctypes.byref(p.c()), amp, seed, _dptr(out))
    return out


def r2r_3d_noalloc(in_ptr: int, out_ptr: int, shape, kinds, precision: int = FP64, stream: int = 0) -> None:
    """smad_r2r_3d_noalloc on device buffers given by address (FP64,
    nx*ny*nz values each): enqueues the transform on `stream` (a
    hipStream_t handle, 0 = default stream) and returns."""
    nx, ny, nz = shape
    k = (ctypes.c_int * 3)(*kinds)
    L = lib()
    rc = L.smad_r2r_3d_noalloc(nx, ny, nz, k, precision, ctypes.c_void_p(in_ptr), ctypes.c_void_p(out_ptr),
                               ctypes.c_void_p(stream))
    if rc != OK:
        raise SmadError(rc, L.smad_noalloc_last_error().decode())


def r2r_3d(x: np.ndarray, kinds, precision: int = FP64) -> np.ndarray:
    """Unnormalised separable 3D r2r transform on the GPU, kinds (x, y, z)
    as FFTW's REDFT10/REDFT01/RODFT10/RODFT01."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.empty_like(x)
    k = (ctypes.c_int * 3)(*kinds)
    L = lib()
    rc = L.smad_r2r_3d(x.shape[0], x.shape[1], x.shape[2], k, precision, _dptr(x), _dptr(out))
    if rc != OK:
        raise SmadError(rc, L.smad_last_error(None).decode())
    return out


# -------------------------------------------------------------- stepper --

class Stepper:
    """One device context: the reference's pre-loop at construction, then
    :meth:`step` for the loop body."""

    def __init__(self, params: Params, psi0: np.ndarray):
        self.params = params
        L = lib()
        psi0 = np.ascontiguousarray(psi0, dtype=np.float64)
        if psi0.shape != params.shape:
            raise ValueError(f"psi0 shape {psi0.shape} != grid {params.shape}")
        self._cp = params.c()
        h = ctypes.c_void_p()
        rc = L.smad_init(ctypes.byref(self._cp), _dptr(psi0), ctypes.byref(h))
        if rc != OK:
            raise SmadError(rc, L.smad_last_error(None).decode())
        self._h = h

    def _check(self, rc):
        if rc != OK:
            raise SmadError(rc, lib().smad_last_error(self._h).decode())

    def step(self, n: int = 1):
        self._check(lib().smad_step(self._h, int(n)))

    def fetch(self, field: int = FIELD_PSI, allow_nonfinite=False) -> np.ndarray:
        out = np.empty(self.params.shape, dtype=np.float64)
        rc = lib().smad_fetch(self._h, field, _dptr(out))
        if rc == ENONFINITE and allow_nonfinite:
            return out
        self._check(rc)
        return out

    def set_field(self, field: int, arr: np.ndarray):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        if a.shape != self.params.shape:
            raise ValueError("shape mismatch")
        self._check(lib().smad_set_field(self._h, field, _dptr(a)))

    def reduce(self, what: int = RED_L1) -> np.ndarray:
        out = np.zeros(4, dtype=np.float64)
        self._check(lib().smad_reduce(self._h, what, _dptr(out)))
        return out

    def l1(self) -> float:
        """L1 = sum|psi_new - psi_old| / (dt sum|psi_new|) over the last step."""
        return float(self.reduce(RED_L1)[2])

    def forward_tape(self, nsteps: int, ckpt_every: int):
        """nsteps steps from the pre-loop state, checkpointing every
        ckpt_every steps for a following :meth:`adjoint`."""
        self._check(lib().smad_forward_tape(self._h, int(nsteps), int(ckpt_every)))

    def surface(self, envelope: bool = False):
        """Interface tracker of the current psi (plain model): dict of
        (nx, ny) arrays surf (surface k, -1 if none), vel, curv_h, curv_k.
        envelope=False: the crossing scan (psi > 0.7 then psi < 0);
        envelope=True: the |psi|-envelope scan (|psi| > 0.6, walk up to
        |psi| <= 0.1)."""
        p = self.params
        out = [np.empty((p.nx, p.ny)) for _ in range(4)]
        f = lib().smad_surface_envelope if envelope else lib().smad_surface
        self._check(f(self._h, *[_dptr(a) for a in out]))
        return dict(zip(("surf", "vel", "curv_h", "curv_k"), out))

    def set_tape_budget(self, gib: float):
        """HBM (GiB) the next forward_tape may spend on tape beyond one
        segment; 0 = strict checkpointing.  Without a call: a quarter of
        the free device memory."""
        self._check(lib().smad_set_tape_budget(self._h, float(gib)))

    def adjoint(self, dl_dpsik: np.ndarray):
        """Discrete adjoint of the taped run: returns (dL/dpsi0,
        dL/d(eps, alpha, beta, gamma)) for the cotangent dL/dpsi_K."""
        g = np.ascontiguousarray(dl_dpsik, dtype=np.float64)
        if g.shape != self.params.shape:
            raise ValueError("shape mismatch")
        d0 = np.empty_like(g)
        dp = np.zeros(4)
        self._check(lib().smad_adjoint(self._h, _dptr(g), _dptr(d0), _dptr(dp)))
        return d0, dp

    def set_params(self, params: Params):
        """New coefficients for the same model, grid and precision."""
        cp = params.c()
        self._check(lib().smad_set_params(self._h, ctypes.byref(cp)))
        self.params = params
        self._cp = cp

    def forward_tape_dev(self, psi0_ptr: int, nsteps: int, ckpt_every: int, psik_ptr: int):
        """forward_tape on device FP64 buffers given by address: psi0 is the
        initial condition (the pre-loop runs on it), psi_K is written to
        psik.  Returns when done."""
        self._check(lib().smad_forward_tape_dev(self._h, ctypes.c_void_p(psi0_ptr), int(nsteps), int(ckpt_every),
                                                ctypes.c_void_p(psik_ptr)))

    def adjoint_dev(self, g_ptr: int, d0_ptr: int) -> np.ndarray:
        """adjoint on device FP64 buffers: cotangent at g, dL/dpsi0 written
        to d0; returns dL/d(eps, alpha, beta, gamma) (host)."""
        dp = np.zeros(4)
        self._check(lib().smad_adjoint_dev(self._h, ctypes.c_void_p(g_ptr), ctypes.c_void_p(d0_ptr), _dptr(dp)))
        return dp

    def adjoint_timing(self):
        """(device ms of the last forward_tape, device ms of the last adjoint)."""
        a, b = ctypes.c_double(), ctypes.c_double()
        self._check(lib().smad_adjoint_timing(self._h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def time_steps(self, n: int) -> float:
        """Run n steps between two HIP events on the context stream; returns ms."""
        ms = ctypes.c_double()
        self._check(lib().smad_time_steps(self._h, int(n), ctypes.byref(ms)))
        return ms.value

    def profile_steps(self, n: int):
        """Run n steps with an event after every launch.  Returns a list of
        (kernel class name, summed ms, launches)."""
        nmax = 32
        ms = (ctypes.c_double * nmax)()
        cnt = (ctypes.c_int * nmax)()
        nc = ctypes.c_int()
        self._check(lib().smad_profile_steps(self._h, int(n), ms, cnt, nmax, ctypes.byref(nc)))
        return [(lib().smad_kernel_class_name(self._h, i).decode(), ms[i], cnt[i])
                for i in range(nc.value)]

    def sync(self):
        self._check(lib().smad_sync(self._h))

    def state_fields(self):
        """The field ids of the model's full restart state, in restore order."""
        buf = (ctypes.c_int * 32)()
        nf = lib().smad_state_fields(self._h, buf, 32)
        if nf < 0:
            raise SmadError(nf, "smad_state_fields")
        return [buf[i] for i in range(nf)]

    def save_state(self):
        """Full restart state: (step count, {field id: array})."""
        return self.steps_taken, {f: self.fetch(f) for f in self.state_fields()}

    def load_state(self, state):
        """Restore a state from :meth:`save_state` (same parameters)."""
        steps, arrays = state
        for f in self.state_fields():
            self.set_field(f, arrays[f])
        self._check(lib().smad_set_steps_taken(self._h, int(steps)))

    @property
    def stream_ptr(self) -> int:
        return int(lib().smad_stream(self._h) or 0)

    @property
    def steps_taken(self) -> int:
        return int(lib().smad_steps_taken(self._h))

    def close(self):
        if getattr(self, "_h", None):
            lib().smad_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def run_to_l1(stepper: Stepper, lim_l1: float, every: int = 100, max_steps: int = 10**9):
    """The reference's loop control: step `every` steps, then compute L1;
    stop when L1 <= lim_l1 (no-adv/cosNoAdvTrackMult.cpp:724, :814-842).
    Non-finite psi raises instead of stopping silently."""
    hist = []
    taken = 0
    l1 = float("inf")
    while l1 > lim_l1 and taken < max_steps:
        stepper.step(every)
        taken += every
        l1 = stepper.l1()
        hist.append(l1)
    return hist
