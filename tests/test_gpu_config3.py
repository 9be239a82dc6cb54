"""Elastic FWI gradient of one shot at full size (2000 x 600, DH = 10 m,
FD8, NT = 4000, forward wavefield stored every 2nd step): the vp gradient
from the adjoint correlation against central finite differences of the L2
misfit at 16 seeded grid cells.  Observed data come from the true model
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
    grad = g.gradient(3)
    assert l2 > 0 and np.isfinite(grad).all() and np.abs(grad).max() > 0
    print(f"[config3] misfit {l2:.6e}, footprint {g.footprint() / 1e9:.1f} GB", flush=True)
    g.close()

    def misfit(vp):
        fwd.set_model(vp, bg[1], bg[2])
        fwd.forward()
        return fwd.residuals(obs[0], obs[1], None, 1)

    # 16 seeded cells in the region the shot illuminates around the anomaly
    ci = rng.integers(int((cx - 500) / dh), int((cx + 500) / dh), 16)
    cj = rng.integers(int((cz - 500) / dh), int((cz + 400) / dh), 16)
    eps = 10.0  # m/s (0.3% of vp)
    fd, ad = [], []
    for a, b in zip(ci, cj):
        vpp, vpm = bg[0].copy(), bg[0].copy()
        vpp[a, b] += eps
        vpm[a, b] -= eps
        fd.append((misfit(vpp) - misfit(vpm)) / (2 * eps))
        ad.append(float(grad[a, b]))
    fd, ad = np.array(fd), np.array(ad)
    corr = float(np.corrcoef(fd, ad)[0, 1])
    scale = float(fd @ ad / (ad @ ad))
    resid = float(np.linalg.norm(fd - scale * ad) / np.linalg.norm(fd))
    print(f"[config3] FD vs adjoint at 16 cells: corr {corr:.4f} scale {scale:.4e} residual {resid:.3f}",
          flush=True)
    assert corr > 0.9 and scale > 0 and resid < 0.35
