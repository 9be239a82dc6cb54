"""ctypes binding of libdnio.so (include/denise_io.h): the reference's file
formats (.inp, workflow, sources, receivers, models, SU) for Python callers
and tests.  Host code only; no GPU is touched."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("DNIO_LIB", os.path.join(_HERE, "libdnio.so"))
STR = 256
MAXL = 8


def _s():
    return C.c_char * STR


class Par(C.Structure):
    _fields_ = (
        [(n, C.c_int) for n in ("mode", "physics", "nprocx", "nprocy", "fdorder", "maxrelerror", "nx", "ny")]
        + [(n, C.c_float) for n in ("dh", "time", "dt")]
        + [("quellart", C.c_int), ("signal_file", _s()), ("ts", C.c_float), ("srcrec", C.c_int),
           ("source_file", _s()), ("run_multiple_shots", C.c_int), ("fc_spike_1", C.c_float),
           ("fc_spike_2", C.c_float), ("order_spike", C.c_int), ("write_stf", C.c_int), ("readmod", C.c_int),
           ("mfile", _s()), ("writemod", C.c_int), ("L", C.c_int), ("fl", C.c_float * MAXL), ("tau", C.c_float),
           ("free_surf", C.c_int), ("fw", C.c_int)]
        + [(n, C.c_float) for n in ("damping", "fpml", "npower", "k_max_pml")]
        + [(n, C.c_int) for n in ("boundary", "snap", "snap_shot")]
        + [(n, C.c_float) for n in ("tsnap1", "tsnap2", "tsnapinc")]
        + [(n, C.c_int) for n in ("idx", "idy", "snap_format")]
        + [("snap_file", _s()), ("seismo", C.c_int), ("readrec", C.c_int), ("rec_file", _s()),
           ("refrec", C.c_float * 4), ("n_streamer", C.c_int), ("rec_incr_x", C.c_float),
           ("rec_incr_y", C.c_float), ("ndt", C.c_int), ("seis_format", C.c_int)]
        + [(n, _s()) for n in ("seis_file_vx", "seis_file_vy", "seis_file_curl", "seis_file_div", "seis_file_p",
                               "log_file")]
        + [("log", C.c_int), ("itermax", C.c_int), ("jacobian", _s()), ("data_dir", _s()),
           ("taper", C.c_int), ("taperlength", C.c_int), ("gradt", C.c_int * 4), ("invmat1", C.c_int),
           ("grad_form", C.c_int), ("quelltypb", C.c_int), ("testshot_start", C.c_int),
           ("testshot_end", C.c_int), ("testshot_incr", C.c_int), ("sws_taper_grad_vert", C.c_int),
           ("sws_taper_grad_hor", C.c_int), ("exp_taper_grad_hor", C.c_float),
           ("sws_taper_grad_sources", C.c_int), ("sws_taper_circular_per_shot", C.c_int), ("srtshape", C.c_int),
           ("srtradius", C.c_float), ("filtsize", C.c_int), ("sws_taper_file", C.c_int), ("tfile", _s()),
           ("inv_mod_out", C.c_int), ("inv_modelfile", _s())]
        + [(n, C.c_float) for n in ("vpupperlim", "vplowerlim", "vsupperlim", "vslowerlim", "rhoupperlim",
                                    "rholowerlim", "qsupperlim", "qslowerlim")]
        + [(n, C.c_int) for n in ("grad_method", "pcg_beta", "nlbfgs", "model_filter", "filt_size", "dtinv")]
        + [("eps_scale", C.c_float), ("stepmax", C.c_int), ("scalefac", C.c_float), ("trkill", C.c_int),
           ("trkill_file", _s()), ("picks_file", _s()), ("misfit_log_file", _s())]
        + [(n, C.c_int) for n in ("min_iter", "grad_filter", "filt_size_grad", "timelapse")]
        + [("data_dir_t0", _s())]
        + [(n, C.c_int) for n in ("rtmod", "gravity", "ngravb", "nzgrav", "grav_type", "back_density")]
        + [("dfile", _s()), ("rtm_shot", C.c_int), ("nt", C.c_int), ("lines", C.c_int)]
    )

    def get(self, name):
        v = getattr(self, name)
        return v.decode() if isinstance(v, bytes) else v


class Stage(C.Structure):
    _fields_ = (
        [("pro", C.c_float), ("time_filt", C.c_int), ("fc_start", C.c_float), ("fc_end", C.c_float),
         ("order", C.c_int), ("timewin", C.c_int), ("gamma", C.c_float), ("twlength_minus", C.c_float),
         ("twlength_plus", C.c_float)]
        + [(n, C.c_int) for n in ("inv_vp_iter", "inv_vs_iter", "inv_rho_iter", "inv_qs_iter", "spatfilter")]
        + [("wd_damp", C.c_float), ("wd_damp1", C.c_float)]
        + [(n, C.c_int) for n in ("eprecond", "lnorm", "rowi", "inv_stf")]
        + [("offsetc_stf", C.c_float), ("eps_stf", C.c_float), ("normalize", C.c_int), ("offset_mute", C.c_int),
           ("offsetc", C.c_float), ("scalerho", C.c_float), ("scaleqs", C.c_float), ("env", C.c_int),
           ("gamma_grav", C.c_float), ("n_order", C.c_int)]
    )


_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"libdnio.so not built at {LIB_PATH}; run __graft_entry__.build()")
    lib = C.CDLL(LIB_PATH)
    fp, ip, cp = C.POINTER(C.c_float), C.POINTER(C.c_int), C.c_char_p
    sig = {
        "dnio_last_error": ([], C.c_char_p),
        "dnio_free": ([C.c_void_p], None),
        "dnio_read_par": ([cp, C.POINTER(Par)], C.c_int),
        "dnio_read_stage": ([cp, C.c_int, C.POINTER(Stage)], C.c_int),
        "dnio_count_stages": ([cp, ip], C.c_int),
        "dnio_read_sources": ([cp, ip, C.POINTER(fp), fp], C.c_int),
        "dnio_split_sources": ([C.POINTER(Par), fp, C.c_int, C.c_int, C.c_int, fp, ip], C.c_int),
        "dnio_read_receivers": ([C.POINTER(Par), C.c_int, ip, C.POINTER(ip)], C.c_int),
        "dnio_wavelet": ([C.POINTER(Par), fp, C.c_int, C.c_int, C.c_float, C.c_int, fp], C.c_int),
        "dnio_read_model": ([C.POINTER(Par), fp, fp, fp, fp, fp], C.c_int),
        "dnio_write_model": ([cp, fp, C.c_int, C.c_int], C.c_int),
        "dnio_write_seis": ([cp, C.POINTER(Par), fp, C.c_int, C.c_int, ip, C.c_float, C.c_float], C.c_int),
        "dnio_read_su": ([cp, C.c_int, C.c_int, fp], C.c_int),
        "dnio_tdfilt": ([fp, C.c_int, C.c_int, C.c_float, C.c_int, C.c_float, C.c_float], C.c_int),
        "dnio_write_snap": ([C.POINTER(Par), C.c_int, fp, fp, fp, fp, fp, fp, fp], C.c_int),
        "dnio_seis_name": ([C.POINTER(Par), C.c_int, C.c_int, C.c_int, C.c_char_p, C.c_int], C.c_int),
        "dnio_data_name": ([C.POINTER(Par), C.c_int, C.c_int, C.c_char_p, C.c_int], C.c_int),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _lib = lib
    return lib


class IOError_(RuntimeError):
    pass


def _check(rc):
    if rc != 0:
        raise IOError_(load().dnio_last_error().decode(errors="replace"))


def _b(path):
    return os.fsencode(path)


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def read_par(path) -> Par:
    p = Par()
    _check(load().dnio_read_par(_b(path), C.byref(p)))
    return p


def read_stage(path, stage=1) -> Stage:
    s = Stage()
    _check(load().dnio_read_stage(_b(path), stage, C.byref(s)))
    return s


def count_stages(path):
    n = C.c_int()
    _check(load().dnio_count_stages(_b(path), C.byref(n)))
    return n.value


def read_sources(path):
    """Returns (srcpos (8, nsrc) float32, ts)."""
    lib = load()
    n = C.c_int()
    ptr = C.POINTER(C.c_float)()
    ts = C.c_float()
    _check(lib.dnio_read_sources(_b(path), C.byref(n), C.byref(ptr), C.byref(ts)))
    a = np.ctypeslib.as_array(ptr, shape=(8 * n.value,)).reshape(8, n.value).copy()
    lib.dnio_free(ptr)
    return a, ts.value


def split_sources(par, srcpos, first, count):
    srcpos = np.ascontiguousarray(srcpos, np.float32)
    loc = np.zeros((8, count), np.float32)
    nloc = C.c_int()
    _check(load().dnio_split_sources(C.byref(par), _fp(srcpos), srcpos.shape[1], first, count, _fp(loc),
                                     C.byref(nloc)))
    return np.ascontiguousarray(loc[:, :nloc.value])


def read_receivers(par, ishot=1):
    lib = load()
    n = C.c_int()
    ptr = C.POINTER(C.c_int)()
    _check(lib.dnio_read_receivers(C.byref(par), ishot, C.byref(n), C.byref(ptr)))
    a = np.ctypeslib.as_array(ptr, shape=(3 * max(n.value, 1),))[:3 * n.value].reshape(3, n.value).copy()
    lib.dnio_free(ptr)
    return a


def wavelet(par, loc, ishot, ts, n_order=0):
    loc = np.ascontiguousarray(loc, np.float32)
    nsrc = loc.shape[1]
    out = np.zeros((nsrc, par.nt), np.float32)
    _check(load().dnio_wavelet(C.byref(par), _fp(loc), nsrc, ishot, ts, n_order, _fp(out)))
    return out


def read_model(par):
    n = (par.nx, par.ny)
    vp, vs, rho = (np.zeros(n, np.float32) for _ in range(3))
    taus = np.zeros(n, np.float32) if par.L else None
    taup = np.zeros(n, np.float32) if par.L else None
    _check(load().dnio_read_model(C.byref(par), _fp(vp), _fp(vs), _fp(rho),
                                  _fp(taus) if taus is not None else None,
                                  _fp(taup) if taup is not None else None))
    return vp, vs, rho, taus, taup


def tdfilt(data, dt, order, fc2, fc1=0.0):
    """Butterworth low-pass at fc2 (and high-pass at fc1 > 0) of each row
    of data (ntr x ns); returns the filtered float32 copy."""
    d = np.array(data, np.float32, copy=True, order="C")
    if d.ndim == 1:
        d = d[None, :]
    _check(load().dnio_tdfilt(_fp(d), d.shape[0], d.shape[1], dt, order, fc2, fc1))
    return d


def write_snap(par, nsnap, vx, vy, sxx, syy, ppi, pu, hc):
    """One snapshot (snap_ssg): padded fields as dn_get_field gives them,
    ppi/pu x-major (nx, ny), hc[0..6]."""
    arrs = [np.ascontiguousarray(a, np.float32) for a in (vx, vy, sxx, syy, ppi, pu, hc)]
    _check(load().dnio_write_snap(C.byref(par), nsnap, *[_fp(a) for a in arrs]))


def write_model(path, a):
    a = np.ascontiguousarray(a, np.float32)
    _check(load().dnio_write_model(_b(path), _fp(a), a.shape[0], a.shape[1]))


def write_seis(path, par, section, recpos, xs, ys):
    section = np.ascontiguousarray(section, np.float32)
    recpos = np.ascontiguousarray(recpos, np.int32)
    _check(load().dnio_write_seis(_b(path), C.byref(par), _fp(section), section.shape[0], section.shape[1],
                                  recpos.ctypes.data_as(C.POINTER(C.c_int)), xs, ys))


def read_su(path, ntr, ns):
    out = np.zeros((ntr, ns), np.float32)
    _check(load().dnio_read_su(_b(path), ntr, ns, _fp(out)))
    return out


def seis_name(par, comp, ishot, it=1):
    buf = C.create_string_buffer(1024)
    _check(load().dnio_seis_name(C.byref(par), comp, ishot, it, buf, 1024))
    return buf.value.decode()


def data_name(par, comp, ishot):
    buf = C.create_string_buffer(1024)
    _check(load().dnio_data_name(C.byref(par), comp, ishot, buf, 1024))
    return buf.value.decode()


SU_HEADER = np.dtype([
    ("tracl", "<i4"), ("tracr", "<i4"), ("fldr", "<i4"), ("tracf", "<i4"), ("ep", "<i4"), ("cdp", "<i4"),
    ("cdpt", "<i4"), ("trid", "<i2"), ("nvs", "<i2"), ("nhs", "<i2"), ("duse", "<i2"), ("offset", "<i4"),
    ("gelev", "<i4"), ("selev", "<i4"), ("sdepth", "<i4"), ("gdel", "<i4"), ("sdel", "<i4"), ("swdep", "<i4"),
    ("gwdep", "<i4"), ("scalel", "<i2"), ("scalco", "<i2"), ("sx", "<i4"), ("sy", "<i4"), ("gx", "<i4"),
    ("gy", "<i4"), ("counit", "<i2"), ("pad1", "<i2", (12,)), ("ns", "<u2"), ("dt", "<u2"),
    ("pad2", "<i2", (31,)), ("d1", "<f4"), ("f1", "<f4"), ("d2", "<f4"), ("f2", "<f4"), ("ungpow", "<f4"),
    ("unscale", "<f4"), ("ntr", "<i4"), ("mark", "<i2"), ("unass", "<i2", (15,)),
])
assert SU_HEADER.itemsize == 240


def parse_su(path):
    """Headers and samples of an SU file (all traces the same length)."""
    raw = np.fromfile(path, np.uint8)
    ns = int(np.frombuffer(raw[114:116].tobytes(), "<u2")[0])
    rec = 240 + 4 * ns
    ntr = raw.size // rec
    assert ntr * rec == raw.size, "ragged SU file"
    blocks = raw.reshape(ntr, rec)
    hdr = np.frombuffer(blocks[:, :240].copy().tobytes(), SU_HEADER)
    data = np.frombuffer(blocks[:, 240:].copy().tobytes(), "<f4").reshape(ntr, ns)
    return hdr, data
