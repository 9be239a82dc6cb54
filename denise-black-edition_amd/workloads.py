"""Synthetic workloads of the multi-shot forms, shared by the benchmark and
the tests: the batched mix of independent uneven shots and the served
Poisson stream of forward and gradient requests."""
import math

import numpy as np


def batch_workload(nshots=64, seed=2024):
    """The multi-shot workload: nshots independent elastic P/SV shots with
    NX ~ U[500,3000], NY ~ U[200,800], DH = 10 m, FDORDER in {4, 8},
    NT ~ U[1500,5000], DT = 1 ms, each with its own seeded smooth-random
    model and a Ricker source (fc ~ U[5,20] Hz) at the centre 20 m deep,
    receivers every 20 m at 20 m depth."""
    import dnhip
    import synth
    rng = np.random.default_rng(seed)
    shots = []
    for s in range(nshots):
        nx, ny = int(rng.integers(500, 3001)), int(rng.integers(200, 801))
        fd = int(rng.choice([4, 8]))
        nt = int(rng.integers(1500, 5001))
        fc = float(rng.uniform(5.0, 20.0))
        dh, dt = 10.0, 1e-3
        vp, vs, rho = synth.quick_layered_model(nx, ny, dh, seed=s + 1)
        p = dnhip.Params(nx=nx, ny=ny, nt=nt, dh=dh, dt=dt, fdorder=fd, maxrelerror=1, fw=20, free_surf=1,
                         L=0, seismo=1, ts=float(np.float32(1.0 / fc)), damping=1500.0, fpml=10.0)
        src = synth.source_table(nx * dh / 2, 20.0, dh, fc=fc)
        sig = synth.ricker(nt, dt, fc)[None, :]
        nrec = (nx - 2) * int(dh) // 20
        rec = synth.receiver_line(20.0, 20.0, nrec, 20.0, dh)
        shots.append((p, (vp, vs, rho), src, sig, rec))
    return shots


def serving_workload(nreq, qps, seed=77):
    """Poisson arrivals at `qps`: 70% elastic forward modelling, 30% FWI
    gradients (forward with snapshots every 2nd step + adjoint); grids
    log-uniform from 256x256 to 4096x1024 (NX, NY drawn independently), NT
    uniform in [1000, 8000], DH = 10 m, DT = 1 ms, FD8, seeded smooth-random
    models, Ricker source (fc 10 Hz) at the centre, receivers every 20 m.
    Observed data of a gradient job are zero traces (the work of a gradient
    does not depend on them)."""
    import dnhip
    import serve
    import synth
    rng = np.random.default_rng(seed)
    t = 0.0
    reqs = []
    for r in range(nreq):
        t += float(rng.exponential(1.0 / qps))
        nx = int(round(math.exp(rng.uniform(math.log(256), math.log(4096)))))
        ny = int(round(math.exp(rng.uniform(math.log(256), math.log(1024)))))
        nt = int(rng.integers(1000, 8001))
        kind = "gradient" if rng.random() < 0.3 else "forward"
        dh, dt, fc = 10.0, 1e-3, 10.0
        vp, vs, rho = synth.quick_layered_model(nx, ny, dh, seed=1000 + r)
        p = dnhip.Params(nx=nx, ny=ny, nt=nt, dh=dh, dt=dt, fdorder=8, fw=20, free_surf=1, seismo=1,
                         mode=1 if kind == "gradient" else 0, dtinv=2, ts=float(np.float32(1.0 / fc)))
        src = synth.source_table(nx * dh / 2, 20.0, dh, fc=fc)
        sig = synth.ricker(nt, dt, fc)[None, :]
        nrec = (nx - 2) * int(dh) // 20
        rec = synth.receiver_line(20.0, 20.0, nrec, 20.0, dh)
        obs = ()
        if kind == "gradient":
            z = np.zeros((nrec, nt), np.float32)
            obs = (z, z, None)
        reqs.append(serve.Request(rid=r, kind=kind, params=p, model=(vp, vs, rho), srcpos=src, signals=sig,
                                  recpos=rec, obs=obs, arrival=t))
    return reqs
