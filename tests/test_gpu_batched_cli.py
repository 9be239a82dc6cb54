"""The executable's multi-shot passes: with several shots, MODE 0 forward
modelling, the MODE 1 gradient passes with their step-length test shots and
the MODE 2 migration run the shots together through the engine, one device
context per shot.  Every output file must equal the one-context shot loop's
(DENISE_HIP_SEQUENTIAL=1) byte for byte, also when a small arena
(DENISE_HIP_ARENA_BYTES) splits the passes into groups of shots."""
import filecmp
import os
import shutil
import subprocess

import numpy as np
import pytest

import inpfile
import synth
from test_gpu_cli import BIN, _files

pytestmark = pytest.mark.gpu


def _exe(tmp, env_extra, *args):
    env = dict(os.environ, **env_extra)
    print(f"[batched cli] {' '.join(args)} {env_extra} in {tmp}", flush=True)
    r = subprocess.run([BIN, *args], cwd=tmp, capture_output=True, text=True, timeout=300, env=env)
    print(r.stdout[-1500:], r.stderr[-1500:], flush=True)
    assert r.returncode == 0, r.stderr
    return r.stdout


def _tree(tmp):
    out = []
    for root, _, files in os.walk(tmp):
        for f in files:
            rel = os.path.relpath(os.path.join(root, f), tmp)
            if not rel.endswith((".inp", ".dat")) or rel.startswith("misfit"):
                out.append(rel)
    return sorted(out)


def _setup(tmp, nshots, mode, eprecond=0, itermax=2):
    nx, ny, dh, dt, nt = 120, 70, 10.0, 1e-3, 360
    bg = synth.layered_model(nx, ny, dh, seed=8, sigma=3.0)
    true = tuple(a.copy() for a in bg)
    true[0][50:70, 15:30] *= 1.07
    sources = [(150.0 + 900.0 * k / max(nshots - 1, 1), 30.0, 14.0, 1.0, 1) for k in range(nshots)]
    receivers = [(40.0 + 20.0 * k, 20.0) for k in range(55)]
    for d in ("wavelet", "taper"):
        os.makedirs(os.path.join(tmp, d), exist_ok=True)
    _files(tmp, nx, ny, dh, true, sources, receivers, mname="model/true")
    inpfile.write_inp(os.path.join(tmp, "true.inp"), NX=nx, NY=ny, DH=dh, DT=dt, TIME=nt * dt, MFILE="model/true",
                      SEIS_FILE_VX="obs/d_x.su", SEIS_FILE_VY="obs/d_y.su", SEIS_FILE_P="obs/d_p.su", SEISMO=4)
    if mode == 0:
        return "true.inp", None
    _exe(tmp, {"DENISE_HIP_SEQUENTIAL": "1"}, "true.inp")
    _files(tmp, nx, ny, dh, bg, sources, receivers, mname="model/start")
    inpfile.write_inp(os.path.join(tmp, "run.inp"), MODE=mode, NX=nx, NY=ny, DH=dh, DT=dt, TIME=nt * dt,
                      MFILE="model/start", DTINV=2, TESTSHOT=f"1, {nshots}, 2", ITERMAX=itermax, INV_MOD_OUT=1,
                      MISFIT_LOG_FILE="misfit.dat", SWS_TAPER_GRAD_HOR=1, GRADT="3, 12, 60, 62",
                      EXP_TAPER_GRAD_HOR=2.0, RTM_SHOT=int(mode == 2))
    inpfile.write_workflow(os.path.join(tmp, "wf.inp"), eprecond=eprecond)
    return "run.inp", "wf.inp"


@pytest.mark.parametrize("mode,eprecond,arena", [(0, 0, 0), (0, 0, 6 << 20), (1, 3, 0), (1, 1, 0),
                                                 (1, 3, 40 << 20), (2, 0, 0)])
def test_batched_passes_equal_the_shot_loop(tmp_path, mode, eprecond, arena):
    nshots = 5
    seq, bat = str(tmp_path / "seq"), str(tmp_path / "bat")
    os.makedirs(seq)
    inp, wf = _setup(seq, nshots, mode, eprecond)
    shutil.copytree(seq, bat)
    args = [inp] + ([wf] if wf else [])
    _exe(seq, {"DENISE_HIP_SEQUENTIAL": "1"}, *args)
    env = {"DENISE_HIP_ARENA_BYTES": str(arena)} if arena else {}
    out = _exe(bat, env, *args)
    assert "multi-shot passes: 5 shots through the engine" in out
    groups = [int(line.split()[1]) for line in out.splitlines() if line.startswith(" engine: ")
              and line.endswith("forward run")]
    # a small arena splits the first pass into groups of shots
    assert (groups[0] < 5) == bool(arena), groups
    files = _tree(seq)
    assert files == _tree(bat)
    assert len(files) > 10
    for rel in files:
        assert filecmp.cmp(os.path.join(seq, rel), os.path.join(bat, rel), shallow=False), rel
    if mode == 1:
        assert filecmp.cmp(os.path.join(seq, "misfit.dat"), os.path.join(bat, "misfit.dat"), shallow=False)
        log = np.loadtxt(os.path.join(bat, "misfit.dat"), ndmin=2)
        assert log.shape[0] >= 1 and np.all(np.isfinite(log))
