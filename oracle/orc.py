"""ctypes binding of the CPU oracle (liboracle_f32.so / liboracle_f64.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product never uses it.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))

ORC_VX, ORC_VY, ORC_SXX, ORC_SYY, ORC_SXY, ORC_VXP1, ORC_VYP1, ORC_UX, ORC_UY, ORC_UXY = range(10)
ORC_RIP, ORC_RJP, ORC_UIPJP, ORC_PRHO, ORC_PPI, ORC_PU, ORC_FIPJP, ORC_F, ORC_G = range(10, 19)
ORC_C11, ORC_C13, ORC_C33, ORC_C44H = range(19, 23)
ORC_D11, ORC_D13, ORC_D15, ORC_D15H, ORC_D33, ORC_D35, ORC_D35H, ORC_D55H = range(23, 31)
ORC_UIP, ORC_UJP = range(31, 33)


class _OParams(C.Structure):
    _fields_ = [(n, C.c_int) for n in (
        "nx", "ny", "nt", "fdorder", "maxrelerror", "fw", "free_surf", "boundary", "L",
        "seismo", "ndt", "invmat1", "grad_form", "dtinv", "quelltypb", "mode")] + \
        [(n, C.c_float) for n in ("dh", "dt", "damping", "fpml", "npower", "k_max_pml", "ts")] + \
        [("fl", C.c_float * 8)]


_libs = {}


def load(prec="f32"):
    if prec in _libs:
        return _libs[prec]
    path = os.path.join(_HERE, f"liboracle_{prec}.so")
    if not os.path.exists(path):
        raise RuntimeError(f"oracle not built: {path} (run make -C oracle)")
    lib = C.CDLL(path)
    vp = C.c_void_p
    fp = C.POINTER(C.c_float)
    ip = C.POINTER(C.c_int)
    sig = {
        "orc_create": ([C.POINTER(_OParams), fp, fp, fp, fp, fp], vp),
        "orc_create_aniso": ([C.POINTER(_OParams), C.c_int, fp, fp, fp, fp, fp, fp], vp),
        "orc_create_ac": ([C.POINTER(_OParams), fp, fp], vp),
        "orc_create_sh": ([C.POINTER(_OParams), fp, fp], vp),
        "orc_create_sh_visc": ([C.POINTER(_OParams), fp, fp, fp], vp),
        "orc_destroy": ([vp], None),
        "orc_error": ([], C.c_char_p),
        "orc_set_threads": ([C.c_int], None),
        "orc_set_sources": ([vp, C.c_int, fp, fp], C.c_int),
        "orc_set_receivers": ([vp, C.c_int, ip], C.c_int),
        "orc_forward": ([vp, C.c_int], C.c_int),
        "orc_step": ([vp, C.c_int, C.c_int], C.c_int),
        "orc_zero": ([vp], C.c_int),
        "orc_set_field": ([vp, C.c_int, vp], C.c_int),
        "orc_calc_res": ([vp, fp, fp, fp, C.c_int], C.c_double),
        "orc_set_res_weights": ([vp, C.c_float, C.c_float], C.c_int),
        "orc_set_adjoint": ([vp, fp, fp, fp], C.c_int),
        "orc_adjoint": ([vp, C.c_int], C.c_int),
        "orc_assemble": ([vp], C.c_int),
        "orc_dims": ([vp, ip, ip], C.c_int),
        "orc_get_field": ([vp, C.c_int, vp], C.c_int),
        "orc_get_psi": ([vp, C.c_int, vp], C.c_int),
        "orc_get_mem": ([vp, C.c_int, C.c_int, vp], C.c_int),
        "orc_get_seis": ([vp, C.c_int, vp], C.c_int),
        "orc_get_adjsrc": ([vp, C.c_int, vp], C.c_int),
        "orc_get_grad": ([vp, C.c_int, vp], C.c_int),
        "orc_get_pml": ([vp, C.c_int, vp], C.c_int),
        "orc_get_hc": ([vp, vp], C.c_int),
        "orc_real_size": ([], C.c_int),
        "orc_holberg": ([C.c_int, C.c_int, fp], C.c_int),
        "orc_ricker": ([C.c_int, C.c_float, C.c_float, C.c_float, C.c_float, fp], None),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _libs[prec] = lib
    return lib


def _f32(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float32)


def _p(a, t=C.c_float):
    return None if a is None else a.ctypes.data_as(C.POINTER(t))


def holberg(fdorder, maxrelerror):
    out = np.zeros(7, np.float32)
    if load().orc_holberg(fdorder, maxrelerror, _p(out)):
        raise ValueError(load().orc_error().decode())
    return out


def ricker(nt, dt, tshift, fc, amp):
    out = np.zeros(nt, np.float32)
    load().orc_ricker(nt, dt, tshift, fc, amp, _p(out))
    return out


def set_threads(n):
    for prec in ("f32", "f64"):
        try:
            load(prec).orc_set_threads(int(n))
        except RuntimeError:
            pass


class Oracle:
    """psv() restated on the CPU; prec 'f32' (reference FP32 path) or 'f64'."""

    def __init__(self, params, vp, vs, rho, taus=None, taup=None, prec="f32", aniso=None):
        """aniso: None (isotropic vp, vs, rho), "ac" (acoustic: vp and rho,
        vs ignored; the pressure is field ORC_SXX), "sh" (SH: vs and rho, vp
        ignored; vz is field ORC_VX, sxz ORC_SXX, syz ORC_SYY) or (physics, c11, c13, c33,
        c44, theta) for VTI (physics 3, theta None) / TTI (physics 4, theta in
        degrees); then vp and vs are ignored."""
        self.lib = load(prec)
        self.prec = prec
        self.dtype = np.float32 if self.lib.orc_real_size() == 4 else np.float64
        self.p = params
        cp = _OParams()
        for name, _ in _OParams._fields_:
            if name == "fl":
                for k, v in enumerate(list(params.fl)[:8]):
                    cp.fl[k] = v
            else:
                setattr(cp, name, getattr(params, name))
        self._cp = cp
        if aniso == "ac":
            arrs = [_f32(a) for a in (vp, rho)]
            self.h = self.lib.orc_create_ac(C.byref(cp), *[_p(a) for a in arrs])
        elif aniso == "sh":
            arrs = [_f32(a) for a in (vs, rho, taus)]
            self.h = self.lib.orc_create_sh_visc(C.byref(cp), *[_p(a) for a in arrs])
        elif aniso is None:
            arrs = [_f32(a) for a in (vp, vs, rho, taus, taup)]
            self.h = self.lib.orc_create(C.byref(cp), *[_p(a) for a in arrs])
        else:
            phys, *mods = aniso
            arrs = [_f32(a) for a in [rho] + list(mods)]
            self.h = self.lib.orc_create_aniso(C.byref(cp), int(phys), *[_p(a) for a in arrs])
        if not self.h:
            raise ValueError(self.lib.orc_error().decode())
        nd, ntd = C.c_int(), C.c_int()
        self.lib.orc_dims(self.h, C.byref(nd), C.byref(ntd))
        self.nd, self.ntdtinv = nd.value, ntd.value
        self.ntr = 0

    def close(self):
        if getattr(self, "h", None):
            self.lib.orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _rc(self, rc):
        if rc != 0:
            raise RuntimeError(f"oracle error {rc}: {self.lib.orc_error().decode()}")

    def set_sources(self, srcpos, signals):
        srcpos, signals = _f32(srcpos), _f32(signals)
        self._rc(self.lib.orc_set_sources(self.h, srcpos.shape[1], _p(srcpos), _p(signals)))

    def set_receivers(self, recpos):
        recpos = np.ascontiguousarray(recpos, dtype=np.int32)
        self.ntr = recpos.shape[1]
        self._rc(self.lib.orc_set_receivers(self.h, self.ntr, _p(recpos, C.c_int)))

    def forward(self, nsteps=0):
        self._rc(self.lib.orc_forward(self.h, nsteps))

    def step(self, nt, mode=0):
        self._rc(self.lib.orc_step(self.h, nt, mode))

    def zero(self):
        self._rc(self.lib.orc_zero(self.h))

    def set_field(self, fid, arr):
        a = np.ascontiguousarray(arr, dtype=self.dtype)
        nd = self.nd
        assert a.shape == (self.p.ny + 2 * nd, self.p.nx + 2 * nd)
        self._rc(self.lib.orc_set_field(self.h, fid, a.ctypes.data))

    def residuals(self, obs_vx=None, obs_vy=None, obs_p=None, testshot=1):
        arrs = [_f32(a) for a in (obs_vx, obs_vy, obs_p)]
        return self.lib.orc_calc_res(self.h, *[_p(a) for a in arrs], testshot)

    def set_residual_weights(self, w_vx=1.0, w_vy=1.0):
        self._rc(self.lib.orc_set_res_weights(self.h, w_vx, w_vy))

    def set_adjoint_sources(self, dvx=None, dvy=None, dp=None):
        arrs = [_f32(a) for a in (dvx, dvy, dp)]
        self._rc(self.lib.orc_set_adjoint(self.h, *[_p(a) for a in arrs]))

    def adjoint(self, nsteps=0):
        self._rc(self.lib.orc_adjoint(self.h, nsteps))

    def assemble(self):
        self._rc(self.lib.orc_assemble(self.h))

    def _out(self, shape):
        return np.zeros(shape, self.dtype)

    def field(self, fid):
        nd = self.nd
        out = self._out((self.p.ny + 2 * nd, self.p.nx + 2 * nd))
        self._rc(self.lib.orc_get_field(self.h, fid, out.ctypes.data))
        return out

    def psi(self, pid):
        fw2 = 2 * self.p.fw
        out = self._out((self.p.ny, fw2) if pid < 5 else (fw2, self.p.nx))
        self._rc(self.lib.orc_get_psi(self.h, pid, out.ctypes.data))
        return out

    def memvar(self, which, l):
        nd = self.nd
        out = self._out((self.p.ny + 2 * nd, self.p.nx + 2 * nd))
        self._rc(self.lib.orc_get_mem(self.h, which, l, out.ctypes.data))
        return out

    def seismogram(self, comp):
        out = self._out((self.ntr, self.p.nt))
        self._rc(self.lib.orc_get_seis(self.h, comp, out.ctypes.data))
        return out

    def adjoint_sources(self, comp):
        out = self._out((self.ntr, self.p.nt))
        self._rc(self.lib.orc_get_adjsrc(self.h, comp, out.ctypes.data))
        return out

    def gradient(self, which):
        out = self._out((self.p.nx, self.p.ny))
        self._rc(self.lib.orc_get_grad(self.h, which, out.ctypes.data))
        return out

    def pml(self, which):
        out = self._out(2 * self.p.fw)
        self._rc(self.lib.orc_get_pml(self.h, which, out.ctypes.data))
        return out

    def hc(self):
        out = self._out(7)
        self._rc(self.lib.orc_get_hc(self.h, out.ctypes.data))
        return out
