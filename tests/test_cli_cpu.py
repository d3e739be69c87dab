"""The reconstruction programs' command lines on the CPU: each takes the
reference program's flags (sirt_main.cc:25-46, pml_main.cc:48-49,
apmlr_main.cc:54-63, src/all/trace.cc:35-150) with the same defaults, echoes
its configuration the way the reference does, and fails like TCLAP on bad
command lines (exit 1, "PARSE ERROR").  Every run here stops before any GPU
call (bad flags or unreadable input)."""
import os
import subprocess

import numpy as np
import pytest

from trace_amd import h5io

# (short, long, default or None when required, kind); defaults as printed
MAIN = [("f", "projectionFilePath", None), ("d", "projectionDatasetPath", "/exchange/data"),
        ("e", "thetaFilePath", None), ("q", "thetaDatasetPath", "/exchange/theta"),
        ("o", "reconOutputPath", "./output.h5"), ("r", "reconDatasetPath", "/data"),
        ("i", "iteration", None), ("c", "center", "0"), ("t", "thread", "1")]
EXTRA = {"sirt": [], "mlem": [], "pml": [("j", "beta", "10")],
         "apmlr": [("j", "beta0", "10"), ("k", "beta1", "1"), ("l", "delta0", "1"),
                   ("m", "delta1", "1"), ("w", "regw", "1")]}
UNIFIED = [("", "degree-to-radian", ""), ("", "regw", "1"), ("", "beta-one", "1"),
           ("", "beta-zero", "10"), ("", "delta-one", "1"), ("", "delta-zero", "1"),
           ("p", "projection-file-path", ""), ("d", "projection-dataset-path", "/exchange/data"),
           ("e", "theta-file-path", ""), ("q", "theta-dataset-path", "/exchange/theta"),
           ("m", "output-dataset-path", "/image"), ("c", "center", "0"),
           ("t", "thread-count", "1"), ("i", "iteration-count", None),
           ("o", "output-file-path", "./output.h5"), ("f", "input-file-path", ""),
           ("r", "reconstruction-algorithm", None)]


def run(prog, *args, cwd=None):
    env = dict(os.environ)
    return subprocess.run([h5io.program(prog), *args], capture_output=True, text=True,
                          timeout=60, cwd=cwd, env=env)


@pytest.fixture
def bad_input(tmp_path):
    """A file with a sinogram but no theta dataset: the programs echo their
    configuration, read the sinogram and stop at theta."""
    f = str(tmp_path / "in.h5")
    h5io.write(f, "/exchange/data", np.ones((4, 2, 8), np.float32), truncate=True)
    return f


@pytest.mark.parametrize("prog", ["sirt", "mlem", "pml", "apmlr"])
def test_main_flags_and_defaults(prog):
    r = run(prog, "--help")
    assert r.returncode == 0
    for s, l, d in MAIN + EXTRA[prog]:
        assert f"-{s},  --{l}" in r.stdout, (prog, l)
        block = r.stdout.split(f"--{l}\n", 1)[1].split("\n\n", 1)[0]
        if d is None:
            assert "(required)" in block, (prog, l)
        else:
            assert f"(default {d})" in block, (prog, l, block)


def test_unified_flags_and_defaults():
    r = run("trace", "--help")
    assert r.returncode == 0
    for s, l, d in UNIFIED:
        head = f"-{s},  --{l}" if s else f"--{l}"
        assert head in r.stdout, l
        block = r.stdout.split(f"--{l}\n", 1)[1].split("\n\n", 1)[0]
        if d is None:
            assert "(required)" in block, l
        elif d:
            assert f"(default {d})" in block, (l, block)


@pytest.mark.parametrize("args,text", [
    (["-e", "t.h5", "-i", "3"], "Required argument(s) missing: projectionFilePath"),
    (["-f", "a.h5", "-e", "t.h5"], "Required argument(s) missing: iteration"),
    (["-f", "a.h5", "-e", "t.h5", "-i", "x"], "Couldn't read argument value from string 'x'"),
    (["-f", "a.h5", "-e", "t.h5", "-i", "3", "-c", "1.5.2"], "Couldn't read argument value"),
    (["-f", "a.h5", "-e", "t.h5", "-i", "3", "-z", "1"], "Couldn't find match for argument -z"),
    (["-f", "a.h5", "-f", "b.h5", "-e", "t.h5", "-i", "3"], "Argument already set!"),
    (["-f", "a.h5", "-e", "t.h5", "-i"], "Missing a value for this argument!"),
    (["-f", "a.h5", "-e", "t.h5", "-i", "3", "extra"], "Couldn't find match for argument extra"),
])
def test_main_parse_errors(args, text):
    r = run("sirt", *args)
    assert r.returncode == 1
    assert r.stderr.startswith("PARSE ERROR:") and text in r.stderr, r.stderr


@pytest.mark.parametrize("args,text", [
    (["-i", "3", "-f", "a.h5"], "Required argument(s) missing: reconstruction-algorithm"),
    (["-i", "3", "-f", "a.h5", "-r", "art"], "Value 'art' does not meet constraint"),
    (["-i", "3", "-r", "sirt"], "Only one of them should be given"),
    (["-i", "3", "-r", "sirt", "-f", "a.h5", "-p", "b.h5"], "Only one of them should be given"),
    (["-i", "3", "-r", "sirt", "-p", "b.h5"], "Both projection and theta files need to be set"),
    (["-i", "3", "-r", "sirt", "-f", "a.h5", "--degree-to-radian=1"], "Switch takes no value"),
])
def test_unified_parse_errors(args, text):
    r = run("trace", *args)
    assert r.returncode == 1
    assert text in r.stderr, r.stderr


def test_main_echo_and_input_errors(bad_input, tmp_path):
    r = run("apmlr", "-f", bad_input, "-e", bad_input, "-i", "7", "-c", "3.5", "-t", "16",
            "-j", "0.25", "-k", "2", "-l", "0.5", "-m", "4", "-w", "0.75", "-o", "out.h5",
            "--reconDatasetPath=/rec")
    assert r.returncode == 2
    assert "error: theta dataset: unable to open dataset '/exchange/theta'" in r.stderr
    expect = ["MPI rank:0; MPI size:1", f"Projection file path={bad_input}",
              "Projection dataset path in hdf5=/exchange/data", f"Theta file path={bad_input}",
              "Theta dataset path=/exchange/theta", "Output file path=out.h5",
              "Recon. dataset path=/rec", "Number of iterations=7", "Center value=3.5",
              "Number of threads per process=16", "Beta0=0.25", "Beta1=2", "Delta0=0.5",
              "Delta1=4", "Regw=0.75"]
    assert r.stdout.splitlines() == expect
    r = run("pml", "-f", bad_input, "-e", bad_input, "-i", "1")
    assert r.returncode == 2 and r.stdout.splitlines()[-1] == "Beta=10"


def test_unified_echo(bad_input):
    r = run("trace", "-r", "apmlr", "-i=4", f"--input-file-path={bad_input}",
            "--degree-to-radian", "--beta-zero", "0.5")
    assert r.returncode == 2
    out = r.stdout.splitlines()
    assert out[0] == "Reconstruction algorithm=apmlr"
    assert "Recon. dataset path=/image" in out and "Degree to radian=1" in out
    assert "Beta 0=0.5" in out and "Regw=1" in out


def test_input_shape_errors(tmp_path):
    f = str(tmp_path / "in.h5")
    h5io.write(f, "/exchange/data", np.ones((4, 2, 8), np.float32), truncate=True)
    h5io.write(f, "/exchange/theta", np.ones((2, 2), np.float32))
    h5io.write(f, "/t3", np.ones(3, np.float32))
    h5io.write(f, "/flat", np.ones((4, 8), np.float32))
    cases = [
        (["-f", str(tmp_path / "none.h5"), "-e", f], "unable to open HDF5 file"),
        (["-f", f, "-e", f], "Theta file dimension should be 1!"),
        (["-f", f, "-e", f, "-q", "/t3"], "theta holds 3 angles, the projections 4"),
        (["-f", f, "-e", f, "-d", "/flat", "-q", "/t3"], "must be 3-D"),
    ]
    for args, text in cases:
        r = run("mlem", *args, "-i", "2")
        assert r.returncode == 2 and text in r.stderr, (args, r.stderr)
    r = run("mlem", "-f", f, "-e", f, "-q", "/t3", "-i", "-1")
    assert r.returncode == 2 and "iterations must be >= 0" in r.stderr
