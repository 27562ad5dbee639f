"""The C-ABI library without a GPU: it loads, exports every entry point include/acum.h declares,
its structures have the layout the header gives them, and the host-side entry points (file
contract, synthetic inputs, levels, argument and I/O errors) behave as documented."""
import ctypes as C
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

import acum3_periodic_amd as acum

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "acum.h")


def header_functions():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(acum_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_entry_point():
    names = header_functions()
    assert len(names) >= 30
    lib = C.CDLL(acum.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, f"declared in acum.h but not exported: {missing}"
    assert set(names) == set(acum.EXPORTS), "Python binding list out of step with the header"


def test_no_device_needed_to_load():
    assert acum.device_count() >= 0
    if acum.device_count() == 0:
        ctx = C.c_void_p()
        assert acum.lib.acum_ctx_create(C.byref(ctx), 0) == acum.ACUM_ERR_DEVICE
        assert acum.lib.acum_ctx_create(None, 0) == acum.ACUM_ERR_ARG


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_struct_layout_matches_header(tmp_path):
    structs = {"acum_tonal_case": acum.TonalCase, "acum_bpm_case": acum.BpmCase,
               "acum_timing": acum.Timing, "acum_tune": acum.Tune, "acum_inputs": acum.Inputs,
               "acum_synth_params": acum.SynthParams}
    src = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', 'int main(void){']
    for cname, py in structs.items():
        src.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            src.append(f'printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    src.append("return 0;}")
    (tmp_path / "l.c").write_text("\n".join(src))
    subprocess.run(["gcc", "-std=c99", "-o", str(tmp_path / "l"), str(tmp_path / "l.c")], check=True)
    out = subprocess.run([str(tmp_path / "l")], capture_output=True, text=True, check=True).stdout
    for line in out.splitlines():
        cname, field, val = line.split()
        py = structs[cname]
        got = C.sizeof(py) if field == "size" else getattr(py, field).offset
        assert got == int(val), f"{cname}.{field}: ctypes {got} != C {val}"


def test_synth_save_load_round_trip(tmp_path):
    p = acum.synth_params(1, nChord=3, nSpan=2, nTime=16)
    a = acum.Run.synth(p)
    a.save(str(tmp_path / "r"))
    b = acum.Run.load(str(tmp_path / "r"))
    for i in range(a.inputs().nSurf):
        assert np.array_equal(a.surface_raw(i), b.surface_raw(i))
    ia, ib = a.inputs(), b.inputs()
    for f, _ in acum.Inputs._fields_:
        va, vb = getattr(ia, f), getattr(ib, f)
        if f == "vInf":
            va, vb = list(va), list(vb)
        assert va == vb, f
    ca = acum.case_arrays(a.preprocess().tonal_case(acum.FP64))
    cb = acum.case_arrays(b.preprocess().tonal_case(acum.FP64))
    for k in ca:
        assert np.array_equal(ca[k], cb[k]), k


def test_synth_broadband_round_trip(tmp_path):
    a = acum.Run.synth(acum.synth_params(3, nSect=5, nTime=8, nObsAz=2, nObsEl=2))
    a.save(str(tmp_path / "r"))
    b = acum.Run.load(str(tmp_path / "r"))
    ca, cb = a.bpm_case(), b.bpm_case()
    g = lambda c, f, n: np.ctypeslib.as_array(getattr(c, f), shape=(n,)).copy()
    tot = int(g(ca, "nsect", ca.nSurf_BB).sum())
    assert np.array_equal(g(ca, "bbdata", 7 * tot * ca.nTime), g(cb, "bbdata", 7 * tot * cb.nTime))
    assert np.array_equal(g(ca, "ff", ca.nf2), g(cb, "ff", cb.nf2))


def test_io_errors(tmp_path):
    with pytest.raises(acum.AcumError) as e:
        acum.Run.load(str(tmp_path / "absent"))
    assert e.value.code == acum.ACUM_ERR_IO
    d = tmp_path / "r"
    acum.Run.synth(acum.synth_params(1, nChord=3, nSpan=2, nTime=8)).save(str(d))
    # unknown keyword in the input file (io.cu:66 exits; the library returns an error)
    inp = (d / "input").read_text()
    (d / "input").write_text(inp + "\nbogus_keyword 3\n")
    with pytest.raises(acum.AcumError) as e:
        acum.Run.load(str(d))
    assert e.value.code == acum.ACUM_ERR_ARG and "bogus_keyword" in str(e.value)
    (d / "input").write_text(inp)
    # truncated surface file
    s = d / "surfaces" / "surface_1.dat"
    txt = s.read_text()
    s.write_text(txt[: len(txt) // 2])
    with pytest.raises(acum.AcumError) as e:
        acum.Run.load(str(d))
    assert e.value.code == acum.ACUM_ERR_IO
    # missing observer file
    s.write_text(txt)
    os.remove(d / "obs.dat")
    with pytest.raises(acum.AcumError):
        acum.Run.load(str(d))


def test_argument_errors():
    run = acum.Run.synth(acum.synth_params(1, nChord=3, nSpan=2, nTime=8))
    with pytest.raises(acum.AcumError):
        run.tonal_case()  # not preprocessed yet
    run.preprocess()
    with pytest.raises(acum.AcumError):
        run.preprocess()  # twice
    with pytest.raises(acum.AcumError):
        run.bpm_case()  # no broadband inputs
    with pytest.raises(acum.AcumError):
        acum.synth_params(99)


def test_oaspl_parseval():
    """OASPL = 10 log10(mean_t (p - mean p)^2 / (2e-5)^2) per observer."""
    nT = 256
    t = np.arange(nT) / nT
    amps = np.array([1.0, 0.02, 30.0])
    pA = (amps[None, :] * np.sin(2 * np.pi * 5 * t)[:, None] + 3.0).astype(np.float32)
    got = acum.oaspl(pA)
    expect = 10 * np.log10(amps ** 2 / 2 / (2e-5) ** 2)
    np.testing.assert_allclose(got, expect, atol=1e-4)


def test_batch_case_parameters_are_seeded():
    a = acum.batch_params(31004, 7)
    b = acum.batch_params(31004, 7)
    c = acum.batch_params(31004, 8)
    fa = [getattr(a, f) for f, _ in acum.SynthParams._fields_]
    fb = [getattr(b, f) for f, _ in acum.SynthParams._fields_]
    fc = [getattr(c, f) for f, _ in acum.SynthParams._fields_]
    assert fa == fb and fa != fc
    assert 2 <= a.nBlades <= 6 and 256 <= a.nTime <= 2048
