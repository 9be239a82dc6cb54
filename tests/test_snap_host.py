"""The wavefield snapshot writer of libdnio (dnio_write_snap, the
reference's snap_ssg) on the CPU: its files equal the numpy restatement
(oracle/snap_oracle.py) value for value, binary and ASCII, for SNAP 1-4 and
IDX/IDY subsampling; snapshots after the first are appended; and the
restatement is pinned: x/y are the fields at the grid points, p = -sxx-syy,
and with Taylor coefficients a velocity field linear in x and y gives the
exact divergence and curl (times sqrt(ppi), sqrt(pu))."""
import os

import numpy as np
import pytest

import dnio
import inpfile
import orc
import snap_oracle as so

F32 = np.float32


def par_for(tmp, snap, fmt, nx=23, ny=17, fd=4, idx=1, idy=1):
    path = os.path.join(tmp, f"s{snap}{fmt}.inp")
    inpfile.write_inp(path, NX=nx, NY=ny, FD_ORDER=fd, DH=5.0, SNAP=snap, SNAP_FORMAT=fmt,
                      SNAP_FILE=os.path.join(tmp, f"snap{snap}{fmt}"), IDX=idx, IDY=idy)
    return dnio.read_par(path)


def fields(rng, nx, ny, fd):
    nd = fd // 2 + 1
    return [rng.standard_normal((ny + 2 * nd, nx + 2 * nd)).astype(F32) for _ in range(4)]


def read(par, suffix, n, fmt):
    ext = ".asc" if fmt == 2 else ".bin"
    path = f"{par.get('snap_file')}{ext}.{suffix}.0.0"
    if fmt == 2:
        return [line.strip() for line in open(path)]
    return np.fromfile(path, F32)


@pytest.mark.parametrize("snap,fmt,idx", [(1, 3, 1), (2, 3, 1), (3, 3, 1), (4, 3, 2), (4, 2, 1)])
def test_snap_files_equal_restatement(tmp_path, snap, fmt, idx):
    rng = np.random.default_rng(snap * 10 + fmt)
    par = par_for(str(tmp_path), snap, fmt, idx=idx, idy=idx)
    vx, vy, sxx, syy = fields(rng, par.nx, par.ny, par.fdorder)
    ppi = rng.uniform(1500, 4000, (par.nx, par.ny)).astype(F32)
    pu = (ppi / F32(1.7)).astype(F32)
    hc = orc.holberg(par.fdorder, 1)
    dnio.write_snap(par, 1, vx, vy, sxx, syy, ppi, pu, hc)
    dnio.write_snap(par, 2, vy, vx, syy, sxx, ppi, pu, hc)        # appended
    want1 = so.snap_values(snap, par.fdorder, par.dh, vx, vy, sxx, syy, ppi, pu, hc, idx, idx)
    want2 = so.snap_values(snap, par.fdorder, par.dh, vy, vx, syy, sxx, ppi, pu, hc, idx, idx)
    assert want1
    for suffix in want1:
        got = read(par, suffix, None, fmt)
        want = np.concatenate([want1[suffix], want2[suffix]])
        if fmt == 2:
            assert got == ["%e" % float(v) for v in want], suffix
        else:
            assert np.array_equal(got, want), suffix


def test_snap_pins(tmp_path):
    par = par_for(str(tmp_path), 4, 3, nx=20, ny=14, fd=4)
    nx, ny, nd, dh = par.nx, par.ny, par.fdorder // 2 + 1, float(par.dh)
    j = np.arange(-nd + 1, ny + nd + 1)[:, None].astype(np.float64)
    i = np.arange(-nd + 1, nx + nd + 1)[None, :].astype(np.float64)
    a, b, c, e = 0.002, -0.003, 0.0015, 0.0005
    vx = (a * i * dh + c * j * dh).astype(F32)            # d vx/dx = a, d vx/dy = c
    vy = (e * i * dh + b * j * dh).astype(F32)            # d vy/dx = e, d vy/dy = b
    sxx = np.ones_like(vx) * F32(2.0)
    syy = np.ones_like(vx) * F32(3.5)
    ppi = np.full((nx, ny), 2500.0, F32)
    pu = np.full((nx, ny), 1600.0, F32)
    hc = orc.holberg(4, 0)                                 # Taylor: exact on linear fields
    out = so.snap_values(4, 4, dh, vx, vy, sxx, syy, ppi, pu, hc)
    assert np.allclose(out["div"], (a + b) * np.sqrt(2500.0), rtol=2e-4)
    assert np.allclose(out["rot"], (c - e) * np.sqrt(1600.0), rtol=2e-4)
    assert np.array_equal(out["x"], np.array([vx[jj + nd - 1, ii + nd - 1] for ii in range(1, nx + 1)
                                              for jj in range(1, ny + 1)], F32))
    p = so.snap_values(2, 4, dh, vx, vy, sxx, syy, ppi, pu, hc)["p"]
    assert np.all(p == F32(-5.5))
    dnio.write_snap(par, 1, vx, vy, sxx, syy, ppi, pu, hc)
    assert np.array_equal(read(par, "div", None, 3), out["div"])


def test_snap_su_format_refused(tmp_path):
    par = par_for(str(tmp_path), 1, 1)
    rng = np.random.default_rng(0)
    vx, vy, sxx, syy = fields(rng, par.nx, par.ny, par.fdorder)
    m = np.ones((par.nx, par.ny), F32)
    with pytest.raises(dnio.IOError_):
        dnio.write_snap(par, 1, vx, vy, sxx, syy, m, m, orc.holberg(4, 1))
