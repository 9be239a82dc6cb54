"""Elastic FWI gradient of one shot at full size (2000 x 600, DH = 10 m,
FD8, NT = 4000, forward wavefield stored every 2nd step): the vp gradient
from the adjoint correlation against central finite differences of the L2
misfit at 16 seeded grid cells, and the vs and rho gradients along smooth
directions, all at the one ratio DTINV / rho_rec the discretisation fixes.  Observed data come from the true model
(smooth background plus a +10% Gaussian anomaly of radius 300 m); the
starting model is the background."""
import numpy as np
import pytest

import dnhip
import synth

pytestmark = pytest.mark.gpu


def test_config3_gradient_vs_finite_differences():
    nx, ny, nt, dh, dt, fc = 2000, 600, 4000, 10.0, 1e-3, 10.0
    bg = synth.layered_model(nx, ny, dh, seed=3)
    rng = np.random.default_rng(33)
    cx, cz = float(rng.uniform(8000.0, 12000.0)), float(rng.uniform(2000.0, 3500.0))
    true = synth.anomaly(*bg, dh, cx, cz, radius=300.0, amp=0.10)
    common = dict(nx=nx, ny=ny, nt=nt, dh=dh, dt=dt, fdorder=8, fw=20, free_surf=1, seismo=1, dtinv=2,
                  ts=float(np.float32(1.0 / fc)))
    src = synth.source_table(nx * dh / 2, 20.0, dh, fc=fc)
    sig = synth.ricker(nt, dt, fc)[None, :]
    rec = synth.receiver_line(20.0, 20.0, 999, 20.0, dh)
    print(f"[config3] anomaly at ({cx:.0f} m, {cz:.0f} m)", flush=True)

    fwd = dnhip.Context(dnhip.Params(mode=0, **common))
    fwd.set_sources(src, sig)
    fwd.set_receivers(rec)
    fwd.set_model(*true)
    fwd.forward()
    obs = (fwd.seismogram(0), fwd.seismogram(1))

    g = dnhip.Context(dnhip.Params(mode=1, **common))
    g.set_model(*bg)
    g.set_sources(src, sig)
    g.set_receivers(rec)
    g.forward()
    l2 = g.residuals(obs[0], obs[1], None, 1)
    g.adjoint()
    g.assemble()
    grads = [g.gradient(3 + k) for k in range(3)]
    assert l2 > 0 and all(np.isfinite(a).all() and np.abs(a).max() > 0 for a in grads)
    print(f"[config3] misfit {l2:.6e}, footprint {g.footprint() / 1e9:.1f} GB", flush=True)
    g.close()

    def misfit(vp=bg[0], vs=bg[1], rho=bg[2]):
        fwd.set_model(vp, vs, rho)
        fwd.forward()
        return fwd.residuals(obs[0], obs[1], None, 1)

    # The adjoint gradient is (rho_rec / DTINV) times the misfit derivative:
    # the residual enters the adjoint run as a force without the 1/rho of
    # the velocity equation, and the correlation runs every DTINV-th step
    # (the derivation is with the oracle pins).  rho varies by about 1% along
    # this receiver line, so the expected ratio uses the mean of 1/rho there.
    rec_rho = bg[2][rec[0] - 1, rec[1] - 1].astype(np.float64)
    expect = common["dtinv"] * float(np.mean(1.0 / rec_rho))

    # vp at 16 seeded cells around the anomaly (vp enters only the cell's
    # own normal-stress update, where the correlation sits)
    ci = rng.integers(int((cx - 500) / dh), int((cx + 500) / dh), 16)
    cj = rng.integers(int((cz - 500) / dh), int((cz + 400) / dh), 16)
    eps = 10.0  # m/s (0.3% of vp)
    fd, ad = [], []
    for a, b in zip(ci, cj):
        vpp, vpm = bg[0].copy(), bg[0].copy()
        vpp[a, b] += eps
        vpm[a, b] -= eps
        fd.append((misfit(vp=vpp) - misfit(vp=vpm)) / (2 * eps))
        ad.append(float(grads[0][a, b]))
    fd, ad = np.array(fd), np.array(ad)
    corr = float(np.corrcoef(fd, ad)[0, 1])
    scale = float(fd @ ad / (ad @ ad))
    resid = float(np.linalg.norm(fd - scale * ad) / np.linalg.norm(fd))
    print(f"[config3] vp at 16 cells: corr {corr:.4f} scale {scale:.4e} (DTINV / rho_rec {expect:.4e}) "
          f"residual {resid:.3f}", flush=True)
    assert corr > 0.95 and resid < 0.2
    assert abs(scale / expect - 1.0) < 0.05

    # vs and rho along smooth directions (the gradient smoothed over 3 cells
    # in a 1 km window around the anomaly): the same ratio
    from synth import gaussian_smooth as gaussian_filter
    ia, ib = int((cx - 500) / dh), int((cx + 500) / dh)
    ja, jb = int((cz - 500) / dh), int((cz + 500) / dh)
    for k, name in ((1, "vs"), (2, "rho")):
        d = np.zeros((nx, ny))
        d[ia:ib, ja:jb] = gaussian_filter(grads[k][ia:ib, ja:jb].astype(np.float64), 3.0)
        d = gaussian_filter(d, 3.0)
        d *= 3e-3 * float(bg[k][ia:ib, ja:jb].mean()) / np.abs(d).max()
        up = {name: (bg[k] + d).astype(np.float32)}
        dn = {name: (bg[k] - d).astype(np.float32)}
        dpl = (misfit(**up) - misfit(**dn)) / 2.0
        pred = float(np.sum(grads[k].astype(np.float64) * d))
        ratio = dpl / pred
        print(f"[config3] {name} along a smooth direction: FD / gradient {ratio:.4e} "
              f"(DTINV / rho_rec {expect:.4e})", flush=True)
        assert abs(ratio / expect - 1.0) < 0.1, name
