"""ctypes binding of libdnhip.so (include/denise_hip.h) for Python callers.

The hot path is the HIP library; this module only moves numpy arrays across
the C ABI.  Loading fails loudly when the library is missing: there is no
CPU fallback.

Array conventions are the C ABI's: models and gradients x-major (shape
(NX, NY)), seismograms trace-major (shape (ntr, NT)), padded fields in the
reference's NR layout (shape (NY+2nd, NX+2nd)).
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field, fields

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("DNHIP_LIB", os.path.join(_HERE, "libdnhip.so"))

DN_OK = 0
STATUS = {0: "DN_OK", -1: "DN_ERR_ARG", -2: "DN_ERR_HIP", -3: "DN_ERR_OOM",
          -4: "DN_ERR_UNSTABLE", -5: "DN_ERR_STATE", -6: "DN_ERR_CFL"}

F_VX, F_VY, F_SXX, F_SYY, F_SXY, F_VXP1, F_VYP1, F_UX, F_UY, F_UXY = range(10)
SEIS_VX, SEIS_VY, SEIS_P, SEIS_DIV, SEIS_CURL = range(5)


class DnError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{STATUS.get(code, code)}: {msg}")
        self.code = code


class _CParams(C.Structure):
    _fields_ = [(n, C.c_int) for n in (
        "nx", "ny", "nt", "fdorder", "maxrelerror", "fw", "free_surf", "boundary", "L",
        "seismo", "ndt", "invmat1", "grad_form", "dtinv", "quelltypb", "mode")] + \
        [(n, C.c_float) for n in ("dh", "dt", "damping", "fpml", "npower", "k_max_pml", "ts")] + \
        [("fl", C.c_float * 8), ("physics", C.c_int)]


@dataclass
class Params:
    """Run parameters (the reference globals the P/SV path reads)."""
    nx: int
    ny: int
    nt: int
    dh: float
    dt: float
    fdorder: int = 8
    maxrelerror: int = 1
    fw: int = 20
    free_surf: int = 1
    boundary: int = 0
    L: int = 0
    seismo: int = 1
    ndt: int = 1
    invmat1: int = 1
    grad_form: int = 1
    dtinv: int = 1
    quelltypb: int = 1
    mode: int = 0
    damping: float = 1500.0
    fpml: float = 10.0
    npower: float = 4.0
    k_max_pml: float = 1.0
    ts: float = 0.1
    fl: list = field(default_factory=list)
    physics: int = 1

    def to_c(self, cls=_CParams):
        c = cls()
        for f in fields(self):
            if f.name == "fl":
                for k, v in enumerate(self.fl[:8]):
                    c.fl[k] = v
            else:
                setattr(c, f.name, getattr(self, f.name))
        return c


_lib = None


def load():
    """Load libdnhip.so (raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"libdnhip.so not built at {LIB_PATH}; run __graft_entry__.build()")
    lib = C.CDLL(LIB_PATH)
    fp = C.POINTER(C.c_float)
    ip = C.POINTER(C.c_int)
    vp = C.c_void_p
    sig = {
        "dn_abi_version": ([], C.c_int),
        "dn_last_error": ([], C.c_char_p),
        "dn_device_count": ([ip], C.c_int),
        "dn_ctx_create": ([C.POINTER(_CParams), C.c_int, C.POINTER(vp)], C.c_int),
        "dn_ctx_destroy": ([vp], C.c_int),
        "dn_ctx_create_in": ([C.POINTER(_CParams), vp, C.POINTER(vp)], C.c_int),
        "dn_arena_create": ([C.c_int, C.c_size_t, C.POINTER(vp)], C.c_int),
        "dn_arena_destroy": ([vp], C.c_int),
        "dn_arena_info": ([vp, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)], C.c_int),
        "dn_arena_fits": ([vp, C.POINTER(_CParams), C.c_int, C.c_int, ip], C.c_int),
        "dn_set_model": ([vp, fp, fp, fp, fp, fp], C.c_int),
        "dn_set_model_aniso": ([vp, fp, fp, fp, fp, fp, fp], C.c_int),
        "dn_set_sources": ([vp, C.c_int, fp, fp], C.c_int),
        "dn_set_receivers": ([vp, C.c_int, ip], C.c_int),
        "dn_run_forward": ([vp, C.c_int], C.c_int),
        "dn_get_seismogram": ([vp, C.c_int, fp], C.c_int),
        "dn_calc_residuals": ([vp, fp, fp, fp, C.c_int, C.POINTER(C.c_double)], C.c_int),
        "dn_calc_residuals_noalloc": ([vp, fp, fp, fp, C.c_int, C.POINTER(C.c_double)], C.c_int),
        "dn_set_residual_weights": ([vp, C.c_float, C.c_float], C.c_int),
        "dn_set_adjoint_sources": ([vp, fp, fp, fp], C.c_int),
        "dn_get_adjoint_sources": ([vp, C.c_int, fp], C.c_int),
        "dn_run_adjoint": ([vp, C.c_int], C.c_int),
        "dn_assemble_gradient": ([vp], C.c_int),
        "dn_get_gradient": ([vp, C.c_int, fp], C.c_int),
        "dn_get_dims": ([vp, ip, ip], C.c_int),
        "dn_get_field": ([vp, C.c_int, fp], C.c_int),
        "dn_get_psi": ([vp, C.c_int, fp], C.c_int),
        "dn_get_memvar": ([vp, C.c_int, C.c_int, fp], C.c_int),
        "dn_get_pml": ([vp, C.c_int, fp], C.c_int),
        "dn_get_hc": ([vp, fp], C.c_int),
        "dn_set_kernel_sampling": ([vp, C.c_int], C.c_int),
        "dn_set_energy_weights": ([vp, C.c_int], C.c_int),
        "dn_get_energy": ([vp, C.c_int, fp], C.c_int),
        "dn_set_snapshots": ([vp, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p], C.c_int),
        "dn_set_tile": ([vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int], C.c_int),
        "dn_get_tile": ([vp, C.c_int, ip, ip, ip, ip], C.c_int),
        "dn_autotune_tile": ([vp, C.c_int, ip, ip], C.c_int),
        "dn_set_store_hint": ([vp, C.c_int, C.c_int], C.c_int),
        "dn_get_store_hint": ([vp, C.c_int, ip], C.c_int),
        "dn_get_timing": ([vp, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double), ip], C.c_int),
        "dn_get_footprint": ([vp, C.POINTER(C.c_size_t)], C.c_int),
        "dn_estimate_footprint": ([C.POINTER(_CParams), C.c_int, C.c_int, C.POINTER(C.c_size_t)], C.c_int),
        "dn_device_memory": ([C.c_int, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)], C.c_int),
        "dn_engine_create": ([C.c_int, C.POINTER(vp)], C.c_int),
        "dn_engine_destroy": ([vp], C.c_int),
        "dn_engine_attach": ([vp, vp, C.c_int], C.c_int),
        "dn_engine_step": ([vp, C.c_int, ip], C.c_int),
        "dn_engine_status": ([vp, vp, ip], C.c_int),
        "dn_engine_forget": ([vp, vp], C.c_int),
        "dn_engine_set_graph": ([vp, C.c_int], C.c_int),
        "dn_engine_launch": ([vp, C.c_int], C.c_int),
        "dn_engine_wait": ([vp, ip], C.c_int),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _lib = lib
    return lib


def _f32(a):
    if a is None:
        return None
    return np.ascontiguousarray(a, dtype=np.float32)


def _ptr(a, t=C.c_float):
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(t))


def _check(rc):
    if rc != DN_OK:
        raise DnError(rc, load().dn_last_error().decode(errors="replace"))


def estimate_footprint(params: "Params", nsrc: int, ntr: int) -> int:
    """Device bytes a context of these parameters will hold (no device
    needed)."""
    b = C.c_size_t()
    cp = params.to_c()
    _check(load().dn_estimate_footprint(C.byref(cp), nsrc, ntr, C.byref(b)))
    return b.value


def device_memory(device: int = 0):
    """(free, total) device bytes."""
    f, t = C.c_size_t(), C.c_size_t()
    _check(load().dn_device_memory(device, C.byref(f), C.byref(t)))
    return f.value, t.value


def device_count():
    n = C.c_int(0)
    load().dn_device_count(C.byref(n))
    return n.value


class Arena:
    """Device-memory arena (dn_arena): one allocation carved into the
    buffers of the contexts created in it; creating and destroying those
    contexts allocates nothing on the device."""

    def __init__(self, nbytes: int, device: int = 0):
        h = C.c_void_p()
        _check(load().dn_arena_create(device, nbytes, C.byref(h)))
        self.h = h
        self.device = device

    def info(self):
        """(free bytes, largest free block, total bytes)."""
        f, l, t = C.c_size_t(), C.c_size_t(), C.c_size_t()
        _check(load().dn_arena_info(self.h, C.byref(f), C.byref(l), C.byref(t)))
        return f.value, l.value, t.value

    def fits(self, params: "Params", nsrc: int, ntr: int) -> bool:
        """Would a context of these parameters fit in the free blocks now."""
        cp = params.to_c()
        r = C.c_int()
        _check(load().dn_arena_fits(self.h, C.byref(cp), nsrc, ntr, C.byref(r)))
        return bool(r.value)

    def close(self):
        if getattr(self, "h", None):
            _check(load().dn_arena_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Context:
    """One shot's device state (dn_ctx); with `arena`, its device buffers
    come from that arena (on the arena's device)."""

    def __init__(self, params: Params, device: int = 0, arena: Arena | None = None):
        lib = load()
        self.p = params
        self._cp = params.to_c()
        h = C.c_void_p()
        if arena is not None:
            _check(lib.dn_ctx_create_in(C.byref(self._cp), arena.h, C.byref(h)))
            self._arena = arena   # the arena outlives its contexts
        else:
            _check(lib.dn_ctx_create(C.byref(self._cp), device, C.byref(h)))
        self.h = h
        nd = C.c_int()
        ntd = C.c_int()
        _check(lib.dn_get_dims(self.h, C.byref(nd), C.byref(ntd)))
        self.nd, self.ntdtinv = nd.value, ntd.value
        self.ntr = 0
        self._keep = []

    def close(self):
        if getattr(self, "h", None):
            load().dn_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_model(self, vp, vs, rho, taus=None, taup=None):
        arrs = [_f32(a) for a in (vp, vs, rho, taus, taup)]
        _check(load().dn_set_model(self.h, *[_ptr(a) for a in arrs]))

    def set_model_aniso(self, rho, c11, c13, c33, c44, theta=None):
        """VTI (params.physics 3) or TTI (4, theta: tilt in degrees) model,
        each NX x NY x-major."""
        arrs = [_f32(a) for a in (rho, c11, c13, c33, c44, theta)]
        _check(load().dn_set_model_aniso(self.h, *[_ptr(a) for a in arrs]))

    def set_sources(self, srcpos, signals):
        srcpos = _f32(srcpos)
        signals = _f32(signals)
        nsrc = srcpos.shape[1] if srcpos.ndim == 2 else 0
        _check(load().dn_set_sources(self.h, nsrc, _ptr(srcpos), _ptr(signals)))

    def set_receivers(self, recpos):
        recpos = np.ascontiguousarray(recpos, dtype=np.int32)
        self.ntr = recpos.shape[1]
        _check(load().dn_set_receivers(self.h, self.ntr, _ptr(recpos, C.c_int)))

    def forward(self, nsteps=0):
        _check(load().dn_run_forward(self.h, nsteps))

    def seismogram(self, comp):
        out = np.zeros((self.ntr, self.p.nt), np.float32)
        _check(load().dn_get_seismogram(self.h, comp, _ptr(out)))
        return out

    def residuals(self, obs_vx=None, obs_vy=None, obs_p=None, testshot=1):
        l2 = C.c_double(0.0)
        arrs = [_f32(a) for a in (obs_vx, obs_vy, obs_p)]
        _check(load().dn_calc_residuals(self.h, *[_ptr(a) for a in arrs], testshot, C.byref(l2)))
        return l2.value

    def residuals_noalloc(self, obs_vx=None, obs_vy=None, obs_p=None, testshot=1):
        """residuals() through dn_calc_residuals_noalloc (no device
        allocation: static device scratch, ntr * NT <= DN_NOALLOC_MAX_SAMPLES)."""
        l2 = C.c_double(0.0)
        arrs = [_f32(a) for a in (obs_vx, obs_vy, obs_p)]
        _check(load().dn_calc_residuals_noalloc(self.h, *[_ptr(a) for a in arrs], testshot, C.byref(l2)))
        return l2.value

    def set_residual_weights(self, w_vx=1.0, w_vy=1.0):
        """COMP_WEIGHT factors of the vx / vy residuals (calc_res.c:100-123)."""
        _check(load().dn_set_residual_weights(self.h, w_vx, w_vy))

    def set_adjoint_sources(self, dvx=None, dvy=None, dp=None):
        arrs = [_f32(a) for a in (dvx, dvy, dp)]
        _check(load().dn_set_adjoint_sources(self.h, *[_ptr(a) for a in arrs]))

    def adjoint_sources(self, comp):
        out = np.zeros((self.ntr, self.p.nt), np.float32)
        _check(load().dn_get_adjoint_sources(self.h, comp, _ptr(out)))
        return out

    def adjoint(self, nsteps=0):
        _check(load().dn_run_adjoint(self.h, nsteps))

    def assemble(self):
        _check(load().dn_assemble_gradient(self.h))

    def gradient(self, which):
        out = np.zeros((self.p.nx, self.p.ny), np.float32)
        _check(load().dn_get_gradient(self.h, which, _ptr(out)))
        return out

    def field(self, fid):
        nd = self.nd
        out = np.zeros((self.p.ny + 2 * nd, self.p.nx + 2 * nd), np.float32)
        _check(load().dn_get_field(self.h, fid, _ptr(out)))
        return out

    def psi(self, pid):
        fw2 = 2 * self.p.fw
        shape = (self.p.ny, fw2) if pid < 5 else (fw2, self.p.nx)
        out = np.zeros(shape, np.float32)
        _check(load().dn_get_psi(self.h, pid, _ptr(out)))
        return out

    def memvar(self, which, l):
        nd = self.nd
        out = np.zeros((self.p.ny + 2 * nd, self.p.nx + 2 * nd), np.float32)
        _check(load().dn_get_memvar(self.h, which, l, _ptr(out)))
        return out

    def pml(self, which):
        out = np.zeros(2 * self.p.fw, np.float32)
        _check(load().dn_get_pml(self.h, which, _ptr(out)))
        return out

    def hc(self):
        out = np.zeros(7, np.float32)
        _check(load().dn_get_hc(self.h, _ptr(out)))
        return out

    def set_tile(self, tbx, by, ty, vw=1, sweep=-1):
        """Shape of one sweep (0 velocity, 1 stress) or of both (-1)."""
        _check(load().dn_set_tile(self.h, sweep, tbx, by, ty, vw))

    def tile(self, sweep=0):
        a, b, t, w = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        _check(load().dn_get_tile(self.h, sweep, C.byref(a), C.byref(b), C.byref(t), C.byref(w)))
        return a.value, b.value, t.value, w.value

    def autotune(self, nsteps=50):
        """Pick the fastest compiled shape (tbx, by, ty, vw) of each sweep for
        this configuration; returns (velocity shape, stress shape)."""
        v, s = (C.c_int * 4)(), (C.c_int * 4)()
        _check(load().dn_autotune_tile(self.h, nsteps, v, s))
        return tuple(v), tuple(s)

    def set_store_hint(self, streaming, sweep=-1):
        """Cache policy of a sweep: 0 plain, 1 streaming (nontemporal)
        access to the planes a step touches once (material and coefficient
        loads, memory-variable stores)."""
        _check(load().dn_set_store_hint(self.h, sweep, streaming))

    def store_hint(self, sweep=0):
        v = C.c_int()
        _check(load().dn_get_store_hint(self.h, sweep, C.byref(v)))
        return v.value

    def set_energy_weights(self, eprecond):
        """EPRECOND 1 or 3: sum vx^2+vy^2 of the forward (and adjoint) field."""
        _check(load().dn_set_energy_weights(self.h, eprecond))

    def energy(self, which=0):
        out = np.zeros((self.p.nx, self.p.ny), np.float32)
        _check(load().dn_get_energy(self.h, which, _ptr(out)))
        return out

    def set_kernel_sampling(self, every):
        _check(load().dn_set_kernel_sampling(self.h, every))

    def timing(self):
        r, v, s = C.c_double(), C.c_double(), C.c_double()
        n = C.c_int()
        _check(load().dn_get_timing(self.h, C.byref(r), C.byref(v), C.byref(s), C.byref(n)))
        return {"run_ms": r.value, "v_kernel_ms": v.value, "s_kernel_ms": s.value, "samples": n.value}

    def footprint(self):
        b = C.c_size_t()
        _check(load().dn_get_footprint(self.h, C.byref(b)))
        return b.value


class Engine:
    """Multi-shot engine (dn_engine): advances many prepared contexts
    together, one launch per sweep and shot group per time step."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        _check(load().dn_engine_create(device, C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            load().dn_engine_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def attach(self, ctx: Context, phase: int = 0):
        _check(load().dn_engine_attach(self.h, ctx.h, phase))

    def step(self, nsteps: int) -> int:
        n = C.c_int()
        _check(load().dn_engine_step(self.h, nsteps, C.byref(n)))
        return n.value

    def launch(self, nsteps: int) -> None:
        """Enqueue up to nsteps steps and return at once (see wait)."""
        _check(load().dn_engine_launch(self.h, nsteps))

    def wait(self) -> int:
        """Wait for the launched run; returns the shots still attached."""
        n = C.c_int()
        _check(load().dn_engine_wait(self.h, C.byref(n)))
        return n.value

    def run(self, chunk: int = 200) -> None:
        """Step until no shot is attached."""
        while self.step(chunk) > 0:
            pass

    def status(self, ctx: Context) -> int:
        st = C.c_int()
        _check(load().dn_engine_status(self.h, ctx.h, C.byref(st)))
        return st.value

    def forget(self, ctx: Context) -> None:
        _check(load().dn_engine_forget(self.h, ctx.h))

    def set_graph(self, steps: int) -> None:
        _check(load().dn_engine_set_graph(self.h, steps))
