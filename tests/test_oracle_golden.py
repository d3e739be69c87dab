"""The oracle restatement against the golden vectors produced by the compiled
reference (tests/golden/make_golden.py): everything bit for bit."""
import os

import numpy as np
import pytest

from oracle import Params

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _bits(a):
    return np.asarray(a, np.float32).view(np.int32)


def test_golden_rays(oracle):
    g = np.load(os.path.join(GOLD, "rays.npz"))
    off = g["offsets"]
    n = g["N"].size
    assert n > 2000
    for i in range(n):
        N, ctr, th, col = int(g["N"][i]), float(g["center"][i]), g["theta"][i], int(g["col"][i])
        idx, ln, l2 = oracle.trace_ray(N, N, ctr, th, col)
        a, b = off[i], off[i + 1]
        assert idx.size == b - a, (N, ctr, th, col)
        assert np.array_equal(idx, g["idx"][a:b]), (N, ctr, th, col)
        assert np.array_equal(_bits(ln), _bits(g["len"][a:b])), (N, ctr, th, col)
        assert np.array_equal(_bits(l2), _bits(g["len2"][a:b])), (N, ctr, th, col)


def test_golden_scalars(oracle):
    g = np.load(os.path.join(GOLD, "scalars.npz"))
    assert np.array_equal(_bits(oracle.deg2rad(g["deg"])), _bits(g["rad"]))
    assert [oracle.quadrant(q) for q in g["q"]] == list(g["quadrant"])
    for i, C in enumerate(g["C"]):
        for j, c in enumerate(g["center"]):
            assert np.float32(oracle.mov(int(C), float(c))) == g["mov"][i, j]


def _recon_cases():
    g = np.load(os.path.join(GOLD, "recon.npz"))
    return sorted({k.rsplit("_meta", 1)[0] for k in g.files if k.endswith("_meta")})


@pytest.mark.parametrize("key", _recon_cases())
def test_golden_recon(oracle, key):
    g = np.load(os.path.join(GOLD, "recon.npz"))
    alg = key.split("_")[0]
    N, P, S, iters = (int(v) for v in g[key + "_meta"])
    center = float(g[key + "_center"][0])
    b, b1, d, d1, rw = (float(v) for v in g[key + "_params"])
    prm = Params(beta=b, beta1=b1, delta=d, delta1=d1, regw=rw, apmlr_compat=1)
    sino = g[key + "_sino"]
    th_file = g[key + "_theta_file"]
    # the apps' own preprocessing (sirt_main.cc:121, mlem_main.cc:114-117,
    # pml_main.cc:121-124, apmlr_main.cc:151-154)
    th = th_file if alg == "pml" else oracle.deg2rad(th_file)
    s_in = sino if alg == "sirt" else oracle.absolute(sino)
    init = np.zeros((S, N, N), np.float32) if alg == "sirt" else np.ones((S, N, N), np.float32)
    x, st = oracle.reconstruct(alg, th, s_in, init, center, iters, prm, want_stages=True)
    assert np.array_equal(_bits(x), _bits(g[key + "_recon"]))
    for s in ("num", "den", "F", "G"):
        if key + "_" + s in g.files:
            assert np.array_equal(_bits(st[s]), _bits(g[key + "_" + s])), s
