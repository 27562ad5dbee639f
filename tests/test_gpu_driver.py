"""The drop-in driver bin/acum on a generated run directory: the reference's output files, the
same bytes with GPU preprocessing (--gpu-prep) as with host preprocessing, and again when the
surfaces come from the binary cache written by --cache."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ACUM = os.path.join(ROOT, "acum3_periodic_amd", "bin", "acum")
OUTS = ("pT.out", "pL.out", "pA.out")


def run(*args):
    p = subprocess.run([ACUM, *args], capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stderr
    return p.stdout


def read_outs(d):
    return {f: open(os.path.join(d, f), "rb").read() for f in OUTS}


def test_driver_host_and_gpu_preprocessing_agree(tmp_path):
    d = str(tmp_path / "c1")
    run("--generate", "1", d)
    run(d)
    host = read_outs(d)
    assert all(len(v) > 0 for v in host.values())
    for f in OUTS:
        os.remove(os.path.join(d, f))
    run(d, "--gpu-prep", "--cache")
    assert read_outs(d) == host
    assert os.path.exists(os.path.join(d, "surfaces", "surface_1.acumc"))
    # the surfaces now come from the cache
    for f in OUTS:
        os.remove(os.path.join(d, f))
    run(d, "--gpu-prep")
    assert read_outs(d) == host
