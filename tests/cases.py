"""Small seeded P/SV cases shared by the tests (oracle pins and GPU parity)."""
from __future__ import annotations

import numpy as np

import synth
from dnhip import Params


def visco_tau(nx, ny, q=50.0):
    t = np.full((nx, ny), 2.0 / q, np.float32)
    return t, t.copy()


def make_case(nx=96, ny=64, nt=300, dh=10.0, fdorder=8, L=0, free_surf=1, fw=10, seismo=1,
              stype=1, mode=0, seed=3, maxrelerror=1, invmat1=1, grad_form=1, dtinv=2,
              quelltypb=1, fc=20.0, src_xz=None, nrec=None, rec_z=None, homogeneous=False,
              safety=0.7):
    """Returns (params, model dict, srcpos, signals, recpos)."""
    if homogeneous:
        vp = np.full((nx, ny), 3000.0, np.float32)
        vs = (vp / np.sqrt(3.0)).astype(np.float32)
        rho = np.full((nx, ny), 2000.0, np.float32)
    else:
        vp, vs, rho = synth.layered_model(nx, ny, dh, seed, sigma=3.0)
    taus = taup = None
    if L > 0:
        taus, taup = visco_tau(nx, ny)
    vmax = float(vp.max()) * (1.3 if L > 0 else 1.0)
    dt = float(np.float32(synth.stable_dt(vmax, dh, fdorder, safety=safety)))
    if invmat1 == 3:
        mu = (rho * vs * vs).astype(np.float32)
        lam = (rho * vp * vp - 2 * rho * vs * vs).astype(np.float32)
        vp, vs = lam, mu
    p = Params(nx=nx, ny=ny, nt=nt, dh=dh, dt=dt, fdorder=fdorder, maxrelerror=maxrelerror, fw=fw,
               free_surf=free_surf, L=L, seismo=seismo, invmat1=invmat1, grad_form=grad_form,
               dtinv=dtinv, quelltypb=quelltypb, mode=mode, ts=1.0 / fc,
               fl=[5.0, 40.0, 15.0, 60.0, 2.0, 80.0, 10.0, 25.0][:L] if L > 0 else [])
    if src_xz is None:
        src_xz = (dh * (nx // 2), dh * 3)
    srcpos = synth.source_table(src_xz[0], src_xz[1], dh, fc=fc, stype=stype, azimuth=30.0)
    signals = synth.ricker(nt, dt, fc, amp=1.0e3)[None, :]
    if nrec is None:
        nrec = nx // 2 - 4
    rz = rec_z if rec_z is not None else 2 * dh
    recpos = synth.receiver_line(dh * 3, 2 * dh, nrec, rz, dh)
    model = dict(vp=vp, vs=vs, rho=rho, taus=taus, taup=taup)
    return p, model, srcpos, signals, recpos


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / nb) if nb > 0 else float(np.linalg.norm(a - b))
