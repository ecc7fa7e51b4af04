"""ctypes binding of the CPU FP64 oracle -- TEST INFRASTRUCTURE ONLY.

Imported by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg
as the checker.  The product (sma-dft_amd/) never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REDFT10, REDFT01, RODFT10, RODFT01 = 0, 1, 2, 3

_LIB = None


class OParams(ctypes.Structure):
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


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        pp = ctypes.POINTER(OParams)
        L.orc_r2r1d_naive.argtypes = [dp, dp, ctypes.c_int, ctypes.c_int]
        L.orc_r2r1d.argtypes = [dp, dp, ctypes.c_int, ctypes.c_int]
        L.orc_r2r3d.argtypes = [dp, dp] + [ctypes.c_int] * 6
        L.orc_plain_preloop.argtypes = [pp, dp, dp]
        L.orc_plain_step.argtypes = [pp, dp, dp, dp, dp, ctypes.c_int]
        L.orc_l1_sums.argtypes = [dp, dp, ctypes.c_long, dp]
        L.orc_adv_preloop.argtypes = [pp, dp, dp]
        L.orc_adv_step.argtypes = [pp, dp, dp, dp, dp, dp, ctypes.c_int]
        L.orc_quasi_preloop.argtypes = [pp, ctypes.POINTER(QState)]
        L.orc_quasi_step.argtypes = [pp, ctypes.POINTER(QState), ctypes.c_int, dp]
        L.orc_quasi_preloop.restype = None
        L.orc_quasi_step.restype = None
        L.orc_r2r3d_T.argtypes = [dp, dp] + [ctypes.c_int] * 6
        L.orc_r2r3d_T.restype = None
        L.orc_adv_stokes_T.argtypes = [pp, dp, dp, dp, dp, dp, dp]
        L.orc_adv_stokes_T.restype = None
        L.orc_adv_adjoint.argtypes = [pp, dp, ctypes.c_int, dp, dp, dp, dp]
        L.orc_adv_adjoint.restype = None
        L.orc_comp_preloop.argtypes = [pp, ctypes.POINTER(CState)]
        L.orc_comp_step.argtypes = [pp, ctypes.POINTER(CState), ctypes.c_int, dp]
        L.orc_comp_velocity.argtypes = [pp] + [dp] * 10
        for f in ("orc_comp_preloop", "orc_comp_step", "orc_comp_velocity"):
            getattr(L, f).restype = None
        L.orc_quasi_project.argtypes = [pp, dp, dp, dp, dp, dp, dp, dp, dp]
        L.orc_quasi_project.restype = None
        L.orc_adv_stokes.argtypes = [pp, dp, dp, dp, dp, dp, dp]
        L.orc_adv_stokes.restype = None
        L.orc_plain_adjoint.argtypes = [pp, dp, ctypes.c_int, dp, dp, dp, dp]
        L.orc_surface_aux.argtypes = [pp, dp, dp, dp, dp, dp]
        L.orc_surface_track.argtypes = [ctypes.c_int] * 3 + [ctypes.c_double] * 3 + [dp] * 9
        L.orc_surface_track_env.argtypes = [ctypes.c_int] * 3 + [ctypes.c_double] * 3 + [dp] * 9
        L.orc_surface_track_grad.argtypes = [ctypes.c_int] * 3 + [ctypes.c_double] * 3 + [dp] * 12
        L.orc_surface_track_grad.restype = None
        L.orc_adv_surface.argtypes = [pp, dp, dp, dp, dp, dp]
        L.orc_adv_surface.restype = None
        for f in ("orc_surface_aux", "orc_surface_track", "orc_surface_track_env", "orc_plain_adjoint",
                  "orc_r2r1d_naive",
                  "orc_r2r1d", "orc_r2r3d", "orc_plain_preloop",
                  "orc_plain_step", "orc_l1_sums", "orc_adv_preloop", "orc_adv_step"):
            getattr(L, f).restype = None
        _LIB = L
    return _LIB


def _p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def oparams(p):
    """Oracle parameter struct from any object with the smad_params fields."""
    return OParams(**{name: getattr(p, name) for name, _ in OParams._fields_})


def r2r1d(x, kind, naive=False):
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.empty_like(x)
    (lib().orc_r2r1d_naive if naive else lib().orc_r2r1d)(_p(x), _p(y), len(x), kind)
    return y


def r2r3d(x, kinds):
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.empty_like(x)
    lib().orc_r2r3d(_p(x), _p(y), *x.shape, *kinds)
    return y


class QState(ctypes.Structure):
    _fields_ = [(nm, ctypes.POINTER(ctypes.c_double)) for nm in
                ("psi", "psiq", "psiq_old", "nl", "rhoq", "rho", "rho_old", "divv", "vel", "p",
                 "energy", "grad_psi")] + [("nloop", ctypes.c_int),
                                           ("mu", ctypes.POINTER(ctypes.c_double))]

_QARRAYS = ("psi", "psiq", "psiq_old", "nl", "rhoq", "rho", "rho_old", "divv", "vel", "p", "energy",
            "grad_psi", "mu")


class QuasiOracle:
    """Reference-order quasi-incompressible stepper (quasi/quasi.cpp:990-2251)."""

    def __init__(self, p, psi0):
        self.op = oparams(p)
        sh = tuple(np.shape(psi0))
        self.arrays = {}
        for nm in _QARRAYS:
            shape = (3,) + sh if nm in ("vel", "grad_psi") else sh
            self.arrays[nm] = np.zeros(shape)
        self.arrays["psi"][...] = psi0
        self.st = QState(**{nm: _p(a) for nm, a in self.arrays.items()}, nloop=0)
        lib().orc_quasi_preloop(ctypes.byref(self.op), ctypes.byref(self.st))
        self.sums = np.zeros(4)

    def __getattr__(self, nm):
        if nm in ("arrays", "op", "st"):
            raise AttributeError(nm)
        return self.arrays[nm]

    def step(self, n=1):
        lib().orc_quasi_step(ctypes.byref(self.op), ctypes.byref(self.st), int(n), _p(self.sums))


class CState(ctypes.Structure):
    _fields_ = [(nm, ctypes.POINTER(ctypes.c_double)) for nm in
                ("psi", "psiq", "psiq_old", "nl", "rho", "frho", "amp", "f", "divv", "adg", "vel", "vsol",
                 "virr", "grad_psi", "grad_rho")] + [("nloop", ctypes.c_int),
                                                     ("mu", ctypes.POINTER(ctypes.c_double))]


_CARRAYS = ("psi", "psiq", "psiq_old", "nl", "rho", "frho", "amp", "f", "divv", "adg", "vel", "vsol", "virr",
            "grad_psi", "grad_rho", "mu")


class CompOracle:
    """Reference-order compressible stepper (compressible/compressible.cpp:
    1114-2700, one slab).  rho0_field: the starting density (zeros for a new
    start, as the reference, or the loaded density for load == 1)."""

    def __init__(self, p, psi0, rho0_field=None):
        self.op = oparams(p)
        sh = tuple(np.shape(psi0))
        self.arrays = {}
        for nm in _CARRAYS:
            shape = (3,) + sh if nm in ("vel", "vsol", "virr", "grad_psi", "grad_rho") else sh
            self.arrays[nm] = np.zeros(shape)
        self.arrays["psi"][...] = psi0
        if rho0_field is not None:
            self.arrays["rho"][...] = rho0_field
        self.st = CState(**{nm: _p(a) for nm, a in self.arrays.items()}, nloop=0)
        lib().orc_comp_preloop(ctypes.byref(self.op), ctypes.byref(self.st))
        self.sums = np.zeros(4)

    def __getattr__(self, nm):
        if nm in ("arrays", "op", "st"):
            raise AttributeError(nm)
        return self.arrays[nm]

    @property
    def nloop(self):
        return self.st.nloop

    def step(self, n=1):
        lib().orc_comp_step(ctypes.byref(self.op), ctypes.byref(self.st), int(n), _p(self.sums))


def comp_velocity(p, sx, sy, sz, g_hat):
    """(vsol_hat, virr_hat) of compressible/compressible.cpp:1826-2410, each (3,)+shape."""
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (sx, sy, sz, g_hat)]
    out = [np.empty_like(arrs[0]) for _ in range(6)]
    lib().orc_comp_velocity(ctypes.byref(oparams(p)), *[_p(a) for a in arrs], *[_p(a) for a in out])
    return np.stack(out[:3]), np.stack(out[3:])


class PlainOracle:
    """Reference-order plain stepper (no-adv/cosNoAdvTrackMult.cpp:630-805)."""

    def __init__(self, p, psi0):
        self.op = oparams(p)
        self.psi = np.ascontiguousarray(psi0, dtype=np.float64).copy()
        self.psiq = np.zeros_like(self.psi)
        self.nq = np.zeros_like(self.psi)
        self.nq_past = np.zeros_like(self.psi)
        lib().orc_plain_preloop(ctypes.byref(self.op), _p(self.psi), _p(self.nq))

    def step(self, n=1):
        lib().orc_plain_step(ctypes.byref(self.op), _p(self.psi), _p(self.psiq), _p(self.nq),
                             _p(self.nq_past), int(n))


class AdvOracle:
    """Reference-order advective stepper (adv/cosAdvTrackMult.cpp:760-1369)."""

    def __init__(self, p, psi0):
        self.op = oparams(p)
        self.psi = np.ascontiguousarray(psi0, dtype=np.float64).copy()
        self.nq = np.zeros_like(self.psi)
        self.nq_past = np.zeros_like(self.psi)
        n = self.psi.size
        self.grad = np.zeros((3,) + self.psi.shape)
        self.vel = np.zeros((3,) + self.psi.shape)
        lib().orc_adv_preloop(ctypes.byref(self.op), _p(self.psi), _p(self.nq))

    def step(self, n=1):
        lib().orc_adv_step(ctypes.byref(self.op), _p(self.psi), _p(self.nq), _p(self.nq_past),
                           _p(self.grad), _p(self.vel), int(n))


def l1_sums(new, old):
    out = np.zeros(2)
    a = np.ascontiguousarray(new, dtype=np.float64)
    b = np.ascontiguousarray(old, dtype=np.float64)
    lib().orc_l1_sums(_p(a), _p(b), a.size, _p(out))
    return out


def plain_adjoint(p, psi0, nsteps, g_psik):
    """Returns (psi_K, dJ/dpsi0, dJ/d(eps, alpha, beta, gamma)) for the
    cotangent g_psik = dJ/dpsi_K."""
    psi0 = np.ascontiguousarray(psi0, dtype=np.float64)
    g = np.ascontiguousarray(g_psik, dtype=np.float64)
    psik = np.empty_like(psi0)
    d0 = np.empty_like(psi0)
    dp = np.zeros(4)
    op = oparams(p)
    lib().orc_plain_adjoint(ctypes.byref(op), _p(psi0), int(nsteps), _p(g), _p(psik), _p(d0), _p(dp))
    return psik, d0, dp


def r2r3d_T(x, kinds):
    """Transpose of the separable transform r2r3d(., kinds)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.empty_like(x)
    lib().orc_r2r3d_T(_p(x), _p(y), *x.shape, *kinds)
    return y


def adv_stokes_T(p, ux, uy, uz):
    a = [np.ascontiguousarray(v, dtype=np.float64) for v in (ux, uy, uz)]
    out = [np.empty_like(a[0]) for _ in range(3)]
    lib().orc_adv_stokes_T(ctypes.byref(oparams(p)), *[_p(v) for v in a], *[_p(v) for v in out])
    return tuple(out)


def adv_adjoint(p, psi0, nsteps, g_psik):
    """Advective model: (psi_K, dJ/dpsi0, dJ/d(eps, alpha, beta, gamma))."""
    psi0 = np.ascontiguousarray(psi0, dtype=np.float64)
    g = np.ascontiguousarray(g_psik, dtype=np.float64)
    psik = np.empty_like(psi0)
    d0 = np.empty_like(psi0)
    dp = np.zeros(4)
    lib().orc_adv_adjoint(ctypes.byref(oparams(p)), _p(psi0), int(nsteps), _p(g), _p(psik), _p(d0), _p(dp))
    return psik, d0, dp


def adv_forward(p, psi0, nsteps):
    o = AdvOracle(p, psi0)
    o.step(nsteps)
    return o.psi


def plain_forward(p, psi0, nsteps):
    o = PlainOracle(p, psi0)
    o.step(nsteps)
    return o.psi


def adv_stokes(p, sx, sy, sz):
    """Velocity projection alone (spectral in, spectral out)."""
    a = [np.ascontiguousarray(v, dtype=np.float64) for v in (sx, sy, sz)]
    out = [np.empty_like(a[0]) for _ in range(3)]
    op = oparams(p)
    lib().orc_adv_stokes(ctypes.byref(op), *[_p(v) for v in a], *[_p(v) for v in out])
    return tuple(out)


def quasi_project(p, sx, sy, sz, divv_hat):
    """Quasi pressure and velocity projection alone (spectral in, spectral
    out): returns (p_hat, ux_hat, uy_hat, uz_hat)."""
    a = [np.ascontiguousarray(v, dtype=np.float64) for v in (sx, sy, sz, divv_hat)]
    out = [np.empty_like(a[0]) for _ in range(4)]
    lib().orc_quasi_project(ctypes.byref(oparams(p)), *[_p(v) for v in a], *[_p(v) for v in out])
    return tuple(out)


def surface_aux(p, psi):
    """Real-space psi_xx, psi_yy, psi_zz, dTpsi of psi (reference order)."""
    psi = np.ascontiguousarray(psi, dtype=np.float64)
    out = [np.empty_like(psi) for _ in range(4)]
    lib().orc_surface_aux(ctypes.byref(oparams(p)), _p(psi), *[_p(a) for a in out])
    return out


def surface_track(psi, dxx, dyy, dzz, dtp, h, envelope=False):
    """Column scan + velocity + curvatures with grid spacing h (scalar or
    3-tuple); envelope=True takes the |psi|-envelope scan."""
    hx, hy, hz = (h, h, h) if np.isscalar(h) else h
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (psi, dxx, dyy, dzz, dtp)]
    nx, ny, nz = arrs[0].shape
    out = [np.empty((nx, ny)) for _ in range(4)]
    f = lib().orc_surface_track_env if envelope else lib().orc_surface_track
    f(nx, ny, nz, hx, hy, hz, *[_p(a) for a in arrs], *[_p(a) for a in out])
    return dict(zip(("surf", "vel", "curv_h", "curv_k"), out))


def adv_surface(p, psi):
    """The advective save block's surface diagnostics of real psi."""
    psi = np.ascontiguousarray(psi, dtype=np.float64)
    out = [np.empty((p.nx, p.ny)) for _ in range(4)]
    lib().orc_adv_surface(ctypes.byref(oparams(p)), _p(psi), *[_p(a) for a in out])
    return dict(zip(("surf", "vel", "curv_h", "curv_k"), out))


def surface_track_grad(psi, grad, dxx, dyy, dzz, dtp, h):
    """Crossing scan with given gradient arrays (3, nx, ny, nz)."""
    hx, hy, hz = (h, h, h) if np.isscalar(h) else h
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (psi, grad[0], grad[1], grad[2], dxx, dyy, dzz, dtp)]
    nx, ny, nz = arrs[0].shape
    out = [np.empty((nx, ny)) for _ in range(4)]
    lib().orc_surface_track_grad(nx, ny, nz, hx, hy, hz, *[_p(a) for a in arrs], *[_p(a) for a in out])
    return dict(zip(("surf", "vel", "curv_h", "curv_k"), out))


def surface(p, psi, envelope=False):
    """The plain model's surface diagnostics of psi (spacing 2 pi/(q0 nw))."""
    dxx, dyy, dzz, dtp = surface_aux(p, psi)
    return surface_track(psi, dxx, dyy, dzz, dtp, 2 * np.pi / (p.q0 * p.nw), envelope)
