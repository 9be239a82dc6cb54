"""ctypes binding of the host FWI functions of libdnio.so
(include/denise_fwi.h): L-BFGS search direction, parabolic step-length
search, bounded model update -- the FWI iteration of src/PSV/FWI_PSV.c
around the device gradients -- and the adjoint source of reverse-time
migration (MODE 2).

Arrays are float32 of shape (nx, ny) (model-file order, x-major)."""
from __future__ import annotations

import ctypes as C

import numpy as np

import dnio

_MISFIT = C.CFUNCTYPE(C.c_double, C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_float),
                      C.POINTER(C.c_float), C.c_int)


class UpdatePar(C.Structure):
    """dnfwi_update_par: limits and switches of calc_mat_change_test_PSV."""
    _fields_ = [(n, C.c_float) for n in ("vpupperlim", "vplowerlim", "vsupperlim", "vslowerlim",
                                         "rhoupperlim", "rholowerlim", "scalerho")] + \
               [(n, C.c_int) for n in ("inv_vp_iter", "inv_vs_iter", "inv_rho_iter", "invmat1")]


class LineSearch(C.Structure):
    """dnfwi_linesearch (also the state carried between iterations)."""
    _fields_ = [("eps_scale", C.c_float), ("epst", C.c_float * 3), ("L2t", C.c_double * 3),
                ("step1", C.c_int), ("step3", C.c_int), ("nevals", C.c_int)]


_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    lib = dnio.load()
    fp, vp_ = C.POINTER(C.c_float), C.c_void_p
    sig = {
        "dnfwi_lbfgs_create": ([C.c_int, C.c_int, C.c_int, C.POINTER(vp_)], C.c_int),
        "dnfwi_lbfgs_create_ac": ([C.c_int, C.c_int, C.c_int, C.POINTER(vp_)], C.c_int),
        "dnfwi_lbfgs_create_sh": ([C.c_int, C.c_int, C.c_int, C.POINTER(vp_)], C.c_int),
        "dnfwi_assemble_sh": ([C.c_int, C.c_int, C.c_float, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, fp, fp, fp,
                               fp], C.c_int),
        "dnfwi_lbfgs_destroy": ([vp_], None),
        "dnfwi_lbfgs_reset": ([vp_], None),
        "dnfwi_lbfgs_pointer": ([vp_], C.c_int),
        "dnfwi_lbfgs_save": ([vp_, C.c_char_p, fp, fp, fp], C.c_int),
        "dnfwi_lbfgs_direction": ([vp_, C.c_int, C.c_float, C.c_float, C.c_float, fp, fp, fp, fp, fp, fp],
                                  C.c_int),
        "dnfwi_model_mean": ([fp, C.c_int, C.c_int], C.c_float),
        "dnfwi_model_update": ([C.POINTER(UpdatePar), C.c_int, C.c_float, C.c_int, C.c_int, fp, fp, fp, fp, fp, fp,
                                fp, fp, fp], C.c_float),
        "dnfwi_opt_step": ([C.POINTER(C.c_double), fp], C.c_float),
        "dnfwi_step_length": ([C.POINTER(UpdatePar), C.c_int, C.c_float, C.c_float, C.c_int, C.c_double, C.c_int,
                               C.c_int, fp, fp, fp, fp, fp, fp, _MISFIT, vp_, C.POINTER(LineSearch)], C.c_int),
        "dnfwi_energy_precond": ([C.c_int, C.c_int, C.c_int, C.c_float, fp, fp, C.c_float, C.c_float, C.c_float,
                                  fp, fp, fp, fp], C.c_int),
        "dnfwi_taper_hor": ([C.c_int, C.c_int, C.c_float, C.c_int, C.c_int, C.c_float, fp, fp], C.c_int),
        "dnfwi_median_filter": ([C.c_int, C.c_int, C.c_int, fp], C.c_int),
        "dnfwi_gauss_filter_var": ([C.c_int, C.c_int, C.c_float, C.c_float, C.c_float, C.c_float, fp, fp],
                                   C.c_int),
        "dnfwi_residual_gc": ([fp, fp, C.c_int, C.c_int, C.c_float, C.c_int, C.c_int, C.POINTER(C.c_int), fp,
                               C.POINTER(C.c_double)], C.c_int),
        "dnfwi_time_window": ([C.c_int, C.c_int, C.c_int, C.c_float, C.c_float, C.c_float, C.c_float, fp, fp, fp],
                              C.c_int),
        "dnfwi_gauss_filter": ([C.c_int, C.c_int, C.c_float, C.c_float, C.c_float, C.c_float, fp,
                                C.POINTER(C.c_int)], C.c_int),
        "dnfwi_pcg_create": ([C.c_int, C.c_int, C.POINTER(vp_)], C.c_int),
        "dnfwi_pcg_destroy": ([vp_], None),
        "dnfwi_pcg_reset": ([vp_], None),
        "dnfwi_pcg_direction": ([vp_, C.c_int, fp, fp, fp, fp], C.c_int),
        "dnfwi_taper_vert": ([C.c_int, C.c_int, C.c_float, C.POINTER(C.c_int), C.c_int, fp, C.c_int,
                              C.POINTER(C.c_int), fp, fp], C.c_int),
        "dnfwi_taper_sources": ([C.c_int, C.c_int, C.c_float, C.c_float, C.c_int, C.c_int, C.c_int, fp, fp, fp],
                                C.c_int),
        "dnfwi_rtm_adjoint_source": ([fp, fp, C.c_int, C.c_int, C.c_float, C.c_int, C.c_int, fp,
                                      C.POINTER(C.c_double)], C.c_int),
        "dnfwi_stf": ([fp, fp, C.c_int, C.c_int, C.POINTER(C.c_int), C.c_int, C.c_float, fp, fp], C.c_int),
        "dnfwi_stf_npad": ([C.c_int], C.c_int),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _lib = lib
    return lib


def _f(a):
    if a is None:
        return None
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a


def _c(a):
    """a float32 copy (None stays None: no vs in the acoustic problem)"""
    return None if a is None else _f(a).copy()


def _p(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_float))


def update_par(vpupperlim=1e9, vplowerlim=0.0, vsupperlim=1e9, vslowerlim=0.0, rhoupperlim=1e9, rholowerlim=0.0,
               scalerho=1.0, inv_vp_iter=0, inv_vs_iter=0, inv_rho_iter=0, invmat1=1):
    return UpdatePar(vpupperlim, vplowerlim, vsupperlim, vslowerlim, rhoupperlim, rholowerlim, scalerho,
                     inv_vp_iter, inv_vs_iter, inv_rho_iter, invmat1)


class LBFGS:
    """dnfwi_lbfgs: the L-BFGS memory of one FWI run (ac: the acoustic
    problem's two classes, no vs; sh: the SH problem's, no vp; pass the
    missing parameter's arguments as None)."""

    def __init__(self, nx, ny, nlbfgs, ac=False, sh=False):
        self.nx, self.ny = nx, ny
        h = C.c_void_p()
        create = (load().dnfwi_lbfgs_create_sh if sh else
                  load().dnfwi_lbfgs_create_ac if ac else load().dnfwi_lbfgs_create)
        dnio._check(create(nx, ny, nlbfgs, C.byref(h)))
        self.h = h

    def direction(self, it, c_vp, c_vs, c_rho, g_vp, g_vs, g_rho, vp, vs, rho):
        """The search direction (new arrays); the gradients are not modified."""
        d = [_c(g) for g in (g_vp, g_vs, g_rho)]
        m = [_f(a) for a in (vp, vs, rho)]
        for a in d + m:
            assert a is None or a.shape == (self.nx, self.ny)
        dnio._check(load().dnfwi_lbfgs_direction(self.h, it, c_vp, c_vs, c_rho, *[_p(a) for a in d],
                                                 *[_p(a) for a in m]))
        return tuple(d)

    def save(self, jacobian, d_vp, d_vs, d_rho):
        """Write the reference's JACOBIAN_*.old files of the last direction."""
        d = [_f(a) for a in (d_vp, d_vs, d_rho)]
        dnio._check(load().dnfwi_lbfgs_save(self.h, jacobian.encode(), *[_p(a) for a in d]))

    def pointer(self):
        return load().dnfwi_lbfgs_pointer(self.h)

    def reset(self):
        load().dnfwi_lbfgs_reset(self.h)

    def close(self):
        if self.h:
            load().dnfwi_lbfgs_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def model_mean(a):
    a = _f(a)
    return float(load().dnfwi_model_mean(_p(a), a.shape[0], a.shape[1]))


def model_update(par, it, eps_scale, vp, vs, rho, d_vp, d_vs, d_rho):
    """(vp_new, vs_new, rho_new, eps_true)."""
    m = [_f(a) for a in (vp, vs, rho)]
    d = [_f(a) for a in (d_vp, d_vs, d_rho)]
    out = [None if m[k] is None else np.empty_like(m[2]) for k in range(3)]
    nx, ny = m[2].shape
    e = load().dnfwi_model_update(C.byref(par), it, eps_scale, nx, ny, *[_p(a) for a in m + d + out])
    return out[0], out[1], out[2], float(e)


def assemble_sh(dt, invmat1, grad_form, it, inv_vs_iter, inv_rho_iter, vs, rho, g_u, g_rho):
    """dnfwi_assemble_sh on copies of the stacked sums: (g_vs, g_rho)."""
    v, r = _f(vs), _f(rho)
    gu, gr = _c(g_u), _c(g_rho)
    nx, ny = v.shape
    dnio._check(load().dnfwi_assemble_sh(nx, ny, dt, invmat1, grad_form, it, inv_vs_iter, inv_rho_iter, _p(v), _p(r),
                                         _p(gu), _p(gr)))
    return gu, gr


def opt_step(L2t, epst):
    L = (C.c_double * 3)(*[float(v) for v in L2t])
    e = np.ascontiguousarray(epst, dtype=np.float32)
    return float(load().dnfwi_opt_step(L, _p(e)))


def step_length(par, it, eps_scale0, scalefac, stepmax, L2t0, vp, vs, rho, d_vp, d_vs, d_rho, misfit, state=None):
    """step_length_est_psv with misfit(vp, vs, rho, itest) -> float;
    `state` (a LineSearch) carries epst/L2t between iterations."""
    m = [_f(a) for a in (vp, vs, rho)]
    d = [_f(a) for a in (d_vp, d_vs, d_rho)]
    nx, ny = m[2].shape
    err = []

    def cb(_user, pv, ps, pr, itest):
        try:
            arrs = [np.ctypeslib.as_array(p, shape=(nx, ny)).copy() if p else None for p in (pv, ps, pr)]
            return float(misfit(*arrs, itest))
        except Exception as e:  # an exception cannot cross the C frame
            err.append(e)
            return float("nan")

    fn = _MISFIT(cb)
    st = state if state is not None else LineSearch()
    dnio._check(load().dnfwi_step_length(C.byref(par), it, eps_scale0, scalefac, stepmax, L2t0, nx, ny,
                                         *[_p(a) for a in m + d], fn, None, C.byref(st)))
    if err:
        raise err[0]
    return st


def rtm_adjoint_source(data, syn, dt, grad_form=1, testshot=1):
    """MODE 2 adjoint source (ntr x ns, time-reversed) and its L2."""
    d = _f(data)
    sy = _f(syn)
    assert d.shape == sy.shape
    out = np.empty_like(d)
    l2 = C.c_double()
    dnio._check(load().dnfwi_rtm_adjoint_source(_p(d), _p(sy), d.shape[0], d.shape[1], dt, grad_form, testshot,
                                                 _p(out), C.byref(l2)))
    return out, l2.value


def energy_precond(eprecond, dh, Ws, Wr, c_vp, c_vs, c_rho, g_vp, g_vs, g_rho):
    """(We, g_vp, g_vs, g_rho) after dnfwi_energy_precond (new arrays)."""
    ws = _f(Ws)
    wr = _f(Wr) if Wr is not None else np.zeros_like(ws)
    g = [_c(a) for a in (g_vp, g_vs, g_rho)]
    We = np.empty_like(ws)
    nx, ny = ws.shape
    dnio._check(load().dnfwi_energy_precond(eprecond, nx, ny, dh, _p(ws), _p(wr), c_vp, c_vs, c_rho,
                                            *[_p(a) for a in g], _p(We)))
    return We, g[0], g[1], g[2]


def taper_hor(nx, ny, dh, gradt1, gradt2, exp_taper):
    coeff = np.empty((nx, ny), np.float32)
    dnio._check(load().dnfwi_taper_hor(nx, ny, dh, gradt1, gradt2, exp_taper, _p(coeff), None))
    return coeff


def gauss_filter_var(g, vel, dh, wd_damp, wd_damp1, fc_end):
    """dnfwi_gauss_filter_var on a copy of the (nx, ny) gradient g (SPATFILTER 4)."""
    out = np.array(g, np.float32, copy=True, order="C")
    v = np.ascontiguousarray(vel, np.float32)
    dnio._check(load().dnfwi_gauss_filter_var(out.shape[0], out.shape[1], dh, wd_damp, wd_damp1, fc_end, _p(v),
                                              _p(out)))
    return out


def median_filter(a, filt_size):
    """dnfwi_median_filter on a copy of the (nx, ny) array a (MODEL_FILTER 1)."""
    out = np.array(a, np.float32, copy=True, order="C")
    dnio._check(load().dnfwi_median_filter(out.shape[0], out.shape[1], filt_size, _p(out)))
    return out


def gauss_filter(g, dh, wd_damp, c_vs, fc_end):
    """dnfwi_gauss_filter on a copy of the (nx, ny) gradient g (SPATFILTER 3);
    returns (filtered, width)."""
    out = np.array(g, np.float32, copy=True, order="C")
    w = C.c_int(0)
    dnio._check(load().dnfwi_gauss_filter(out.shape[0], out.shape[1], dh, wd_damp, c_vs, fc_end, _p(out),
                                          C.byref(w)))
    return out, w.value


def residual_gc(data, syn, dt, grad_form, testshot, skip=None, l2=0.0):
    """dnfwi_residual_gc: (time-reversed adjoint sources (ntr, ns), l2)."""
    data = np.ascontiguousarray(data, np.float32)
    syn = np.ascontiguousarray(syn, np.float32)
    out = np.empty_like(syn)
    sk = None if skip is None else np.ascontiguousarray(skip, np.int32)
    v = C.c_double(l2)
    dnio._check(load().dnfwi_residual_gc(_p(data), _p(syn), syn.shape[0], syn.shape[1], dt, grad_form, testshot,
                                         None if sk is None else sk.ctypes.data_as(C.POINTER(C.c_int)), _p(out),
                                         C.byref(v)))
    return out, v.value


def time_window(timewin, diff, dt, gamma, twlength_minus, twlength_plus, picks=None, offset=None):
    """dnfwi_time_window on a copy of the (ntr, ns) time-reversed residuals."""
    out = np.array(diff, np.float32, copy=True, order="C")
    pk = None if picks is None else np.ascontiguousarray(picks, np.float32)
    of = None if offset is None else np.ascontiguousarray(offset, np.float32)
    dnio._check(load().dnfwi_time_window(timewin, out.shape[0], out.shape[1], dt, gamma, twlength_minus,
                                         twlength_plus, _p(pk), _p(of), _p(out)))
    return out


class PCG:
    """dnfwi_pcg: the conjugate-gradient memory of one FWI run (GRAD_METHOD 1)."""

    def __init__(self, nx, ny):
        h = C.c_void_p()
        dnio._check(load().dnfwi_pcg_create(nx, ny, C.byref(h)))
        self.h = h

    def direction(self, pcg_beta, g_vp, g_vs, g_rho):
        """(d_vp, d_vs, d_rho, betas): new arrays."""
        d = [_c(a) for a in (g_vp, g_vs, g_rho)]
        b = np.zeros(3, np.float32)
        dnio._check(load().dnfwi_pcg_direction(self.h, pcg_beta, *[_p(a) for a in d], _p(b)))
        return d[0], d[1], d[2], b

    def reset(self):
        load().dnfwi_pcg_reset(self.h)

    def close(self):
        if self.h:
            load().dnfwi_pcg_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def taper_vert(nx, ny, dh, gradt, srcpos, recpos, g=None):
    """(coeff, g tapered) of dnfwi_taper_vert; srcpos 8 x nsrc (metres),
    recpos 2 x ntr (grid indices)."""
    gt = (C.c_int * 4)(*[int(v) for v in gradt])
    sp = np.ascontiguousarray(srcpos, np.float32)
    rp = np.ascontiguousarray(recpos[:2], np.int32)
    coeff = np.empty((nx, ny), np.float32)
    gg = _f(g).copy() if g is not None else None
    dnio._check(load().dnfwi_taper_vert(nx, ny, dh, gt, sp.shape[1], _p(sp), rp.shape[1],
                                        rp.ctypes.data_as(C.POINTER(C.c_int)), _p(coeff),
                                        _p(gg) if gg is not None else None))
    return coeff, gg


def taper_sources(nx, ny, dh, srtradius, srtshape, filtsize, srcpos, g=None):
    """(coeff, g tapered) of dnfwi_taper_sources; srcpos 8 x nsrc (metres)."""
    sp = np.ascontiguousarray(srcpos, np.float32)
    coeff = np.empty((nx, ny), np.float32)
    gg = _f(g).copy() if g is not None else None
    dnio._check(load().dnfwi_taper_sources(nx, ny, dh, srtradius, srtshape, filtsize, sp.shape[1], _p(sp), _p(coeff),
                                           _p(gg) if gg is not None else None))
    return coeff, gg


def stf(obs, syn, mute, normalize, eps_stf, spike):
    """dnfwi_stf: the source wavelet (ns floats) of one shot from its observed
    and synthetic vy sections (ntr, ns).  Returns (wavelet, obs, syn) with
    the sections as the function left them (muted, normalised)."""
    obs = _f(obs).copy()
    syn = _f(syn).copy()
    ntr, ns = obs.shape
    m = None if mute is None else np.ascontiguousarray(mute, dtype=np.int32)
    out = np.zeros(ns, np.float32)
    sp = _f(spike)
    assert sp.size >= ns
    dnio._check(load().dnfwi_stf(_p(obs), _p(syn), ntr, ns,
                                 None if m is None else m.ctypes.data_as(C.POINTER(C.c_int)), normalize, eps_stf,
                                 _p(sp), _p(out)))
    return out, obs, syn
