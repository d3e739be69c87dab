"""The reconstruction programs end to end on the GPU: HDF5 input written
here, the program run as a child process, its HDF5 output compared with
the library's own reconstruction of the same input through the Python
mirror of the reference mains (trace_amd.recon.sirt/mlem/pml/apmlr).

The programs run one trace_reconstruct call per iteration and the mirror
one call for all iterations, so equality also shows that splitting the
iterations over calls changes no bit.  SIRT is also compared with the
oracle (the compiled-reference restatement) at the tolerance of the other
reconstruction parity tests."""
import os
import subprocess

import numpy as np
import pytest

from trace_amd import h5io, recon
from trace_amd.phantom import angles_deg, angles_rad, noisy_sinogram

pytestmark = pytest.mark.gpu

N, P, S, ITERS = 64, 60, 3, 4


def _inputs(seed=5):
    """Angles in degrees, the same angles in radians (through the
    reference's conversion) and a noisy sinogram [P][S][C] of them."""
    rad = angles_rad(P)
    sino = noisy_sinogram(rad, N, S, seed=seed, I0=1e4, K=8, device="cpu").numpy()
    return angles_deg(P), rad, sino


def _run(prog, args, tmp_path):
    print(f"[cli] {prog} {' '.join(args)}", flush=True)
    r = subprocess.run([h5io.program(prog), *args], capture_output=True, text=True, timeout=120,
                       cwd=str(tmp_path))
    assert r.returncode == 0, r.stdout + r.stderr
    return r


@pytest.mark.parametrize("prog", ["sirt", "mlem", "pml", "apmlr"])
def test_program_matches_library(gpu, tmp_path, prog):
    deg, rad, sino = _inputs()
    f = str(tmp_path / "in.h5")
    # PML reads radians (pml_main.cc:121), the others degrees
    h5io.write_input(f, sino, rad if prog == "pml" else deg)
    extra = {"pml": ["-j", "0.5"],
             "apmlr": ["-j", "0.5", "-k", "0.25", "-l", "0.1", "-m", "0.2", "-w", "0.8"]}
    r = _run(prog, ["-f", f, "-e", f, "-i", str(ITERS), "-o", "out.h5"] + extra.get(prog, []),
             tmp_path)
    assert [l for l in r.stdout.splitlines() if l.startswith("Iteration: ")] == \
        [f"Iteration: {i}" for i in range(ITERS)]
    got = h5io.read(str(tmp_path / "out.h5"), "/data")
    assert got.shape == (S, N, N)
    if prog == "sirt":
        want = recon.sirt(deg, sino, ITERS)
    elif prog == "mlem":
        want = recon.mlem(deg, sino, ITERS)
    elif prog == "pml":
        want = recon.pml(rad, sino, ITERS, beta=0.5)
    else:
        want = recon.apmlr(deg, sino, ITERS, beta0=0.5, beta1=0.25, delta0=0.1, delta1=0.2,
                           regw=0.8)
    assert np.array_equal(got, want), float(np.abs(got - want).max())


def test_unified_program(gpu, tmp_path):
    deg, rad, sino = _inputs(seed=6)
    f = str(tmp_path / "in.h5")
    h5io.write_input(f, sino, deg)
    _run("trace", ["-r", "mlem", "--degree-to-radian", "-f", f, "-i", str(ITERS),
                   "-o", "u.h5"], tmp_path)
    got = h5io.read(str(tmp_path / "u.h5"), "/image")
    assert np.array_equal(got, recon.mlem(deg, sino, ITERS))
    # radians in, no conversion; APMLR with the ghost offset corrected
    g = str(tmp_path / "in_rad.h5")
    h5io.write_input(g, sino, rad)
    _run("trace", ["-r", "apmlr", f"--input-file-path={g}", "-i", "3", "--apmlr-fixed",
                   "--beta-zero=0.5", "-o", "a.h5", "-m", "/x"], tmp_path)
    got = h5io.read(str(tmp_path / "a.h5"), "/x")
    want = recon._run("apmlr", rad, recon.absolute(sino), 1.0, 3, 0.0,
                      recon.RegParams(0.5, 1.0, 1.0, 1.0, 1.0, False), 0)
    assert np.array_equal(got, want)


def test_sirt_program_vs_oracle(gpu, oracle, tmp_path):
    deg, _, sino = _inputs(seed=7)
    f = str(tmp_path / "in.h5")
    h5io.write_input(f, sino, deg)
    _run("sirt", ["-f", f, "-e", f, "-i", str(ITERS), "-c", "31.5"], tmp_path)
    got = h5io.read(str(tmp_path / "output.h5"), "/data")
    th = recon.degree_to_radian(deg)
    want = oracle.reconstruct("sirt", th, sino, np.zeros((S, N, N), np.float32), 31.5, ITERS)
    err = float(np.linalg.norm(got - want) / np.linalg.norm(want))
    assert err <= 1e-4, err
