"""HIP path vs the FP32 oracle (reference CPU path restated) on seeded inputs.

Bar: seismograms rel-L2 <= 1e-5 and gradients rel-L2 <= 1e-4 against the
FP32 oracle (the path's accuracy contract); the error of both against the FP64
oracle is printed.  The kernels evaluate the reference's operation order
without contraction, so in practice the HIP results equal the FP32 oracle
bit for bit; `exact` in the printed line reports that.
"""
import numpy as np
import pytest

import dnhip
import orc
from cases import make_case, rel_l2

pytestmark = pytest.mark.gpu

SEIS_TOL = 1e-5
GRAD_TOL = 1e-4
FIELDS = [(dnhip.F_VX, "vx"), (dnhip.F_VY, "vy"), (dnhip.F_SXX, "sxx"), (dnhip.F_SYY, "syy"),
          (dnhip.F_SXY, "sxy")]


def run_oracle(case, prec="f32", nsteps=0):
    p, m, src, sig, rec = case
    o = orc.Oracle(p, m["vp"], m["vs"], m["rho"], m["taus"], m["taup"], prec=prec)
    o.set_sources(src, sig)
    o.set_receivers(rec)
    o.forward(nsteps)
    return o


def run_gpu(case, nsteps=0):
    p, m, src, sig, rec = case
    g = dnhip.Context(p)
    g.set_model(m["vp"], m["vs"], m["rho"], m["taus"], m["taup"])
    g.set_sources(src, sig)
    g.set_receivers(rec)
    g.forward(nsteps)
    return g


def seis_comps(seismo):
    return {0: [], 1: [0, 1], 2: [2], 3: [3, 4], 4: [0, 1, 2, 3, 4]}[seismo]


def check_forward(case, tag, nsteps=0, check_state=True):
    p = case[0]
    o32 = run_oracle(case, "f32", nsteps)
    o64 = run_oracle(case, "f64", nsteps)
    g = run_gpu(case, nsteps)
    for comp in seis_comps(p.seismo):
        ref = o32.seismogram(comp)
        got = g.seismogram(comp)
        e32 = rel_l2(got, ref)
        e64 = rel_l2(got, o64.seismogram(comp))
        r64 = rel_l2(ref, o64.seismogram(comp))
        exact = bool(np.array_equal(got, ref))
        print(f"[{tag}] seis comp {comp}: relL2 vs f32 {e32:.2e}, vs f64 {e64:.2e} "
              f"(oracle f32 vs f64 {r64:.2e}) exact={exact}")
        assert np.isfinite(got).all()
        assert np.abs(ref).max() > 0
        assert e32 <= SEIS_TOL, f"{tag} comp {comp}: {e32}"
    if check_state:
        for fid, name in FIELDS:
            ref = o32.field(fid)
            got = g.field(fid)
            err = rel_l2(got, ref)
            print(f"[{tag}] field {name}: relL2 {err:.2e} maxabs {np.abs(got - ref).max():.2e} "
                  f"exact={bool(np.array_equal(got, ref))}")
            assert err <= SEIS_TOL, f"{tag} field {name}: {err}"
        if p.fw > 0:
            for pid in range(9):
                ref = o32.psi(pid)
                got = g.psi(pid)
                if np.abs(ref).max() == 0:
                    assert np.abs(got).max() == 0
                    continue
                assert rel_l2(got, ref) <= SEIS_TOL, f"{tag} psi {pid}"
        for l in range(1, p.L + 1):
            for which in range(3):
                assert rel_l2(g.memvar(which, l), o32.memvar(which, l)) <= SEIS_TOL, f"{tag} memvar {which} {l}"
    return o32, g


@pytest.mark.parametrize("fdorder", [2, 4, 6, 8, 10, 12])
def test_forward_elastic_fdorders(fdorder):
    case = make_case(nx=96, ny=64, nt=250, fdorder=fdorder, fw=10, free_surf=1)
    check_forward(case, f"el-fd{fdorder}")


@pytest.mark.parametrize("maxrelerror", [0, 2, 4])
def test_forward_taylor_and_holberg_rows(maxrelerror):
    case = make_case(nx=80, ny=60, nt=200, fdorder=6, maxrelerror=maxrelerror)
    check_forward(case, f"el-mre{maxrelerror}")


def test_forward_cpml_four_sides():
    case = make_case(nx=90, ny=70, nt=300, fdorder=8, fw=12, free_surf=0, src_xz=(450.0, 300.0))
    check_forward(case, "el-nofs")


def test_forward_overlapping_cpml_bands():
    # NX < 2*FW: cells in both the left and the right band
    case = make_case(nx=30, ny=40, nt=200, fdorder=4, fw=20, free_surf=0, src_xz=(150.0, 200.0), nrec=8)
    check_forward(case, "el-overlap")


def test_forward_no_cpml():
    case = make_case(nx=70, ny=50, nt=200, fdorder=8, fw=0, free_surf=1)
    check_forward(case, "el-fw0")


@pytest.mark.parametrize("L,fdorder", [(1, 8), (2, 8), (2, 4), (3, 6), (5, 8)])
def test_forward_viscoelastic(L, fdorder):
    case = make_case(nx=96, ny=64, nt=250, fdorder=fdorder, L=L, free_surf=1)
    check_forward(case, f"visc-L{L}-fd{fdorder}")


@pytest.mark.parametrize("tile", [(64, 4, 1, 1), (64, 4, 1, 4)])
def test_forward_viscoelastic_varying_tau(tile):
    """Spatially varying relaxation times (the planes are read per cell);
    uniform ones are passed as scalars -- both must give the oracle."""
    case = make_case(nx=90, ny=64, nt=250, fdorder=8, L=2)
    p, m, src, sig, rec = case
    rng = np.random.default_rng(4)
    m = dict(m)
    m["taus"] = (0.04 * (1 + 0.5 * rng.random((p.nx, p.ny)))).astype(np.float32)
    m["taup"] = (0.03 * (1 + 0.5 * rng.random((p.nx, p.ny)))).astype(np.float32)
    case = (p, m, src, sig, rec)
    o = run_oracle(case)
    g = dnhip.Context(p)
    g.set_tile(*tile)
    g.set_model(m["vp"], m["vs"], m["rho"], m["taus"], m["taup"])
    g.set_sources(src, sig)
    g.set_receivers(rec)
    g.forward()
    for comp in (0, 1):
        assert np.array_equal(g.seismogram(comp), o.seismogram(comp).astype(np.float32)), comp
    for fid, name in FIELDS:
        assert np.array_equal(g.field(fid), o.field(fid).astype(np.float32)), name


def test_forward_viscoelastic_no_free_surface():
    case = make_case(nx=80, ny=64, nt=250, fdorder=8, L=2, free_surf=0, src_xz=(400.0, 300.0))
    check_forward(case, "visc-nofs")


@pytest.mark.parametrize("seismo", [2, 3, 4])
def test_forward_seismo_modes(seismo):
    case = make_case(nx=80, ny=60, nt=200, fdorder=8, seismo=seismo)
    check_forward(case, f"seismo{seismo}")


@pytest.mark.parametrize("stype", [2, 3, 4, 5])
def test_forward_source_types(stype):
    case = make_case(nx=80, ny=60, nt=200, fdorder=8, stype=stype, src_xz=(400.0, 150.0))
    check_forward(case, f"src{stype}")


def test_forward_invmat3():
    case = make_case(nx=80, ny=60, nt=200, fdorder=8, invmat1=3, seismo=4)
    check_forward(case, "invmat3")


def test_forward_partial_run_state():
    case = make_case(nx=96, ny=64, nt=300, fdorder=8)
    check_forward(case, "partial", nsteps=137)


def test_config1_full():
    """Config 1 at full size: 400x200, FD8, DT=1 ms, NT=2000, CPML 20, free
    surface, Ricker 10 Hz at (2000 m, 20 m), 200 receivers every 20 m."""
    import synth
    vp, vs, rho = synth.layered_model(400, 200, 10.0, seed=1)
    p = dnhip.Params(nx=400, ny=200, nt=2000, dh=10.0, dt=1e-3, fdorder=8, fw=20, free_surf=1, ts=0.1)
    src = synth.source_table(2000.0, 20.0, 10.0, fc=10.0)
    sig = synth.ricker(2000, 1e-3, 10.0)[None, :]
    rec = synth.receiver_line(20.0, 20.0, 200, 20.0, 10.0)
    case = (p, dict(vp=vp, vs=vs, rho=rho, taus=None, taup=None), src, sig, rec)
    check_forward(case, "config1", check_state=False)


def test_deterministic_repeat():
    case = make_case(nx=128, ny=96, nt=300, fdorder=8, L=2)
    g = run_gpu(case)
    a = g.seismogram(0)
    g.forward()
    assert np.array_equal(a, g.seismogram(0))


def _gradient_case(grad_form=1, L=0, invmat1=1, dtinv=2, nt=500):
    case = make_case(nx=72, ny=56, nt=nt, fdorder=8, L=L, mode=1, grad_form=grad_form,
                     invmat1=invmat1, dtinv=dtinv, fc=15.0, nrec=30)
    p, m, src, sig, rec = case
    # "observed" data from a model with a local anomaly
    mt = {k: (v.copy() if v is not None else None) for k, v in m.items()}
    mt["vp"][30:42, 8:16] *= 1.08
    mt["vs"][30:42, 8:16] *= 1.08
    o = orc.Oracle(p, mt["vp"], mt["vs"], mt["rho"], mt["taus"], mt["taup"], prec="f32")
    o.set_sources(src, sig)
    o.set_receivers(rec)
    o.forward()
    obs = (o.seismogram(0).astype(np.float32), o.seismogram(1).astype(np.float32))
    return case, obs


def run_gradient(runner, case, obs):
    p, m, src, sig, rec = case
    runner.set_sources(src, sig)
    runner.set_receivers(rec)
    runner.forward()
    l2 = runner.residuals(obs[0], obs[1], None, 1)
    runner.adjoint()
    runner.assemble()
    return l2


@pytest.mark.parametrize("grad_form,L,invmat1,dtinv", [(1, 0, 1, 2), (1, 0, 1, 1), (2, 0, 1, 3),
                                                        (1, 2, 1, 2), (1, 0, 3, 2)])
def test_gradient(grad_form, L, invmat1, dtinv):
    case, obs = _gradient_case(grad_form, L, invmat1, dtinv)
    p, m = case[0], case[1]
    o32 = orc.Oracle(p, m["vp"], m["vs"], m["rho"], m["taus"], m["taup"], prec="f32")
    o64 = orc.Oracle(p, m["vp"], m["vs"], m["rho"], m["taus"], m["taup"], prec="f64")
    g = dnhip.Context(p)
    g.set_model(m["vp"], m["vs"], m["rho"], m["taus"], m["taup"])
    l2o = run_gradient(o32, case, obs)
    run_gradient(o64, case, obs)
    l2g = run_gradient(g, case, obs)
    tag = f"grad-gf{grad_form}-L{L}-inv{invmat1}-dtinv{dtinv}"
    print(f"[{tag}] L2 oracle {l2o:.9e} gpu {l2g:.9e}")
    assert abs(l2g - l2o) <= 1e-5 * abs(l2o)
    for comp in (0, 1):
        assert rel_l2(g.adjoint_sources(comp), o32.adjoint_sources(comp)) <= SEIS_TOL
    for which, name in enumerate(["lam", "mu", "rho", "vp", "vs", "rho_grad"]):
        ref = o32.gradient(which)
        got = g.gradient(which)
        e32 = rel_l2(got, ref)
        e64 = rel_l2(got, o64.gradient(which))
        print(f"[{tag}] {name}: relL2 vs f32 {e32:.2e} vs f64 {e64:.2e} exact={bool(np.array_equal(got, ref))}")
        assert np.abs(ref).max() > 0
        assert e32 <= GRAD_TOL, f"{tag} {name}: {e32}"


@pytest.mark.parametrize("quelltypb", [2, 3, 4, 5, 6, 7])
def test_gradient_adjoint_source_types(quelltypb):
    """QUELLTYPB 2-7 against the oracle: the adjoint forces in y (2, 6, 7)
    and in x (3, 5, 7) of src/PSV/update_v_PML_PSV.c:755-768, the adjoint
    pressure source of src/psource.c:41-52 (4-7) driven by the p residuals
    of src/PSV/calc_res_PSV.c:133-180 (SEISMO 4 records p), and for 5-7 the
    COMP_WEIGHT scaling of the weighted component's synthetic and observed
    samples by C_vp*C_rho (src/calc_res.c:100-123)."""
    case = make_case(nx=72, ny=56, nt=500, fdorder=8, mode=1, seismo=4, quelltypb=quelltypb, dtinv=2,
                     fc=15.0, nrec=30)
    p, m, src, sig, rec = case
    mt = {k: (v.copy() if v is not None else None) for k, v in m.items()}
    mt["vp"][30:42, 8:16] *= 1.08
    mt["vs"][30:42, 8:16] *= 1.08
    t = orc.Oracle(p, mt["vp"], mt["vs"], mt["rho"], prec="f32")
    t.set_sources(src, sig)
    t.set_receivers(rec)
    t.forward()
    obs = [t.seismogram(c).astype(np.float32) for c in (0, 1, 2)]
    w = np.float32(np.float32(m["vp"].mean()) * np.float32(m["rho"].mean()))
    wx = float(w) if quelltypb in (5, 7) else 1.0
    wy = float(w) if quelltypb in (6, 7) else 1.0
    use = [quelltypb in (1, 3, 5, 7), quelltypb in (1, 2, 6, 7), quelltypb >= 4]
    runs = {}
    for name, r in (("f32", orc.Oracle(p, m["vp"], m["vs"], m["rho"], prec="f32")),
                    ("gpu", dnhip.Context(p))):
        if name == "gpu":
            r.set_model(m["vp"], m["vs"], m["rho"])
        r.set_residual_weights(wx, wy)
        r.set_sources(src, sig)
        r.set_receivers(rec)
        r.forward()
        l2 = r.residuals(*[o if u else None for o, u in zip(obs, use)], 1)
        r.adjoint()
        r.assemble()
        runs[name] = (r, l2)
    (o32, l2o), (g, l2g) = runs["f32"], runs["gpu"]
    tag = f"grad-quelltypb{quelltypb}"
    print(f"[{tag}] weights {wx:.6e} {wy:.6e}; L2 oracle {l2o:.9e} gpu {l2g:.9e}")
    assert l2o > 0
    assert abs(l2g - l2o) <= 1e-5 * abs(l2o)
    for comp in (0, 1, 2):
        ref = o32.adjoint_sources(comp)
        got = g.adjoint_sources(comp)
        if not use[comp]:
            assert not np.any(got) and not np.any(ref), comp
            continue
        assert np.abs(ref).max() > 0
        assert rel_l2(got, ref) <= SEIS_TOL, (tag, comp)
    for which, name in enumerate(["lam", "mu", "rho", "vp", "vs", "rho_grad"]):
        ref = o32.gradient(which)
        got = g.gradient(which)
        e32 = rel_l2(got, ref)
        print(f"[{tag}] {name}: relL2 vs f32 {e32:.2e} exact={bool(np.array_equal(got, ref))}")
        assert np.abs(ref).max() > 0
        assert e32 <= GRAD_TOL, f"{tag} {name}: {e32}"


def test_nan_probe_reports_unstable():
    case = make_case(nx=64, ny=48, nt=120, fdorder=4)
    p, m, src, sig, rec = case
    vp = m["vp"].copy()
    vp[32, 24] = np.nan
    g = dnhip.Context(p)
    g.set_model(vp, m["vs"], m["rho"])
    g.set_sources(src, sig)
    g.set_receivers(rec)
    with pytest.raises(dnhip.DnError) as ei:
        g.forward()
    assert ei.value.code == -4


def test_cfl_violation_rejected():
    case = make_case(nx=64, ny=48, nt=50, fdorder=8)
    p, m = case[0], case[1]
    p.dt = p.dt * 3.0
    g = dnhip.Context(p)
    with pytest.raises(dnhip.DnError) as ei:
        g.set_model(m["vp"], m["vs"], m["rho"])
    assert ei.value.code == -6


def test_adjoint_requires_forward_store():
    case = make_case(nx=48, ny=40, nt=60, fdorder=4, mode=1)
    p, m, src, sig, rec = case
    g = dnhip.Context(p)
    g.set_model(m["vp"], m["vs"], m["rho"])
    g.set_sources(src, sig)
    g.set_receivers(rec)
    with pytest.raises(dnhip.DnError) as ei:
        g.adjoint()
    assert ei.value.code == -5


TILES = [(64, 4, 1, 1), (128, 2, 1, 1), (256, 1, 1, 1), (64, 4, 2, 1), (128, 2, 4, 1), (64, 8, 2, 1),
         (64, 4, 1, 4), (64, 4, 2, 4), (64, 2, 4, 4), (64, 4, 1, 2), (128, 2, 2, 2), (64, 4, 2, 2)]


@pytest.mark.parametrize("tile", TILES)
@pytest.mark.parametrize("L", [0, 2])
def test_tile_shapes_agree(tile, L):
    """Every compiled sweep tile shape gives the FP32 oracle's results,
    including tile-boundary sources and receivers and a ragged grid."""
    case = make_case(nx=131, ny=77, nt=220, fdorder=8, L=L, seismo=4, fw=9,
                     src_xz=(640.0, 20.0))
    p, m, src, sig, rec = case
    o = run_oracle(case)
    g = dnhip.Context(p)
    g.set_tile(*tile)
    g.set_model(m["vp"], m["vs"], m["rho"], m["taus"], m["taup"])
    g.set_sources(src, sig)
    g.set_receivers(rec)
    g.forward()
    for comp in range(5):
        assert rel_l2(g.seismogram(comp), o.seismogram(comp)) <= SEIS_TOL, (tile, comp)
    for fid, name in FIELDS:
        assert rel_l2(g.field(fid), o.field(fid)) <= SEIS_TOL, (tile, name)


LEAN_CASES = [(t, 0) for t in TILES] + [(t, 2) for t in [(128, 2, 4, 1), (64, 4, 1, 4), (64, 4, 2, 4), (128, 2, 2, 2)]]


@pytest.mark.parametrize("tile,L", LEAN_CASES)
def test_lean_interior_tiles_exact(tile, L):
    """A grid large enough for interior (lean) tiles in every shape, with an
    explosive source and a receiver line in the interior (so full tiles sit
    inside the lean region too) and a surface line: seismograms, wavefields,
    CPML and memory variables equal the FP32 oracle bit for bit."""
    import synth
    p, m, src, sig, rec = make_case(nx=600, ny=260, nt=240, fdorder=8, L=L, seismo=4, fw=10,
                                    src_xz=(2950.0, 1310.0), nrec=280, seed=11)
    # the surface line plus an interior line 100 m above the source
    rec = np.concatenate([rec, synth.receiver_line(2500.0, 30.0, 30, 1210.0, 10.0)], axis=1)
    case = (p, m, src, sig, rec)
    o = run_oracle(case)
    g = dnhip.Context(p)
    g.set_tile(*tile)
    g.set_model(m["vp"], m["vs"], m["rho"], m["taus"], m["taup"])
    g.set_sources(src, sig)
    g.set_receivers(rec)
    g.forward()
    for comp in range(5):
        assert np.abs(o.seismogram(comp)).max() > 0, comp
        assert np.array_equal(g.seismogram(comp), o.seismogram(comp)), (tile, L, "seis", comp)
    for fid, name in FIELDS:
        assert np.array_equal(g.field(fid), o.field(fid)), (tile, L, name)
    for pid in range(9):
        assert np.array_equal(g.psi(pid), o.psi(pid)), (tile, L, "psi", pid)
    for l in range(1, L + 1):
        for which in range(3):
            assert np.array_equal(g.memvar(which, l), o.memvar(which, l)), (tile, L, "memvar", which, l)


@pytest.mark.parametrize("tile", [(128, 2, 1, 1), (64, 4, 2, 1), (128, 2, 4, 1), (64, 4, 1, 4), (64, 2, 4, 4),
                                  (128, 2, 2, 2)])
def test_tile_shapes_gradient(tile):
    """Every shape, also with different shapes for the two sweeps, gives the
    oracle's gradients."""
    case, obs = _gradient_case(1, 0, 1, 2)
    p, m = case[0], case[1]
    o32 = orc.Oracle(p, m["vp"], m["vs"], m["rho"], m["taus"], m["taup"], prec="f32")
    g = dnhip.Context(p)
    g.set_tile(*tile, sweep=0)
    g.set_tile(*TILES[(TILES.index(tile) + 5) % len(TILES)], sweep=1)
    g.set_model(m["vp"], m["vs"], m["rho"], m["taus"], m["taup"])
    run_gradient(o32, case, obs)
    run_gradient(g, case, obs)
    for which in range(6):
        assert rel_l2(g.gradient(which), o32.gradient(which)) <= GRAD_TOL, (tile, which)


def test_autotune_keeps_results():
    """The autotuner only changes the launch shape: the shot run after it is
    the oracle's, and the chosen shape is a compiled one."""
    case = make_case(nx=150, ny=90, nt=200, fdorder=8, L=2, seismo=4)
    p, m, src, sig, rec = case
    o = run_oracle(case)
    g = dnhip.Context(p)
    g.set_model(m["vp"], m["vs"], m["rho"], m["taus"], m["taup"])
    g.set_sources(src, sig)
    g.set_receivers(rec)
    vt, st = g.autotune(20)
    assert tuple(vt) in TILES and tuple(st) in TILES
    assert tuple(g.tile(0)) == tuple(vt) and tuple(g.tile(1)) == tuple(st)
    g.forward()
    for comp in range(5):
        assert np.array_equal(g.seismogram(comp), o.seismogram(comp).astype(np.float32)), comp


@pytest.mark.parametrize("hint", [1])
@pytest.mark.parametrize("L", [0, 2])
@pytest.mark.parametrize("tile", [(64, 4, 1, 1), (64, 4, 1, 4)])
def test_streaming_stores_keep_results(hint, L, tile):
    """The streaming cache policy (nontemporal loads of the once-per-step
    planes, nontemporal memory-variable stores) changes only how the caches
    are used: results stay the oracle's, in boundary and interior tiles."""
    case = make_case(nx=331, ny=97, nt=200, fdorder=8, L=L, seismo=4, fw=9, src_xz=(1600.0, 500.0))
    p, m, src, sig, rec = case
    o = run_oracle(case)
    g = dnhip.Context(p)
    g.set_tile(*tile)
    g.set_store_hint(hint)
    assert g.store_hint(0) == hint and g.store_hint(1) == hint
    g.set_model(m["vp"], m["vs"], m["rho"], m["taus"], m["taup"])
    g.set_sources(src, sig)
    g.set_receivers(rec)
    g.forward()
    for comp in range(5):
        assert np.array_equal(g.seismogram(comp), o.seismogram(comp).astype(np.float32)), comp
    for fid, name in FIELDS:
        assert np.array_equal(g.field(fid), o.field(fid).astype(np.float32)), name
