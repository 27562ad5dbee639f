"""GPU preprocessing (acum_preprocess_gpu) against the host FP64 preprocessing, which the CPU
tests pin bit for bit to the oracle's FP64 instance: the results must be bit-identical."""
import numpy as np
import pytest

import acum3_periodic_amd as acum

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    if acum.device_count() < 1:
        pytest.fail("no GPU visible")
    c = acum.Context(0)
    yield c
    c.close()


def _pair(config, **over):
    p = acum.synth_params(config, **over)
    return acum.Run.synth(p), acum.Run.synth(p)


@pytest.mark.parametrize("config,over", [
    (1, dict(nChord=4, nSpan=3, nTime=36)),
    (2, dict(nChord=5, nSpan=4, nTime=1024)),
    (1, dict(nChord=3, nSpan=2, nTime=4096)),
    (1, dict(nChord=3, nSpan=2, nTime=7)),
])
def test_gpu_preprocess_bit_identical(ctx, config, over):
    a, b = _pair(config, **over)
    a.preprocess()
    b.preprocess_gpu(ctx)
    ca = acum.case_arrays(a.tonal_case(acum.FP64))
    cb = acum.case_arrays(b.tonal_case(acum.FP64))
    for k in ("X", "N", "V", "PR"):
        assert np.array_equal(ca[k].view(np.uint32), cb[k].view(np.uint32)), k


def test_gpu_preprocess_staged_in_chunks(ctx):
    """Arrays of 97-145 MB, neither a multiple of the 64 MB pinned staging buffers nor of the
    1 MB host-copy pieces, go up and come back through several staging rounds; the results are
    bit-identical to the host preprocessing."""
    a, b = _pair(2, nChord=40, nSpan=37, nTime=512)
    a.preprocess()
    b.preprocess_gpu(ctx)
    ca = acum.case_arrays(a.tonal_case(acum.FP64))
    assert ca["X"].nbytes > (64 << 20) and ca["X"].nbytes % (1 << 20) != 0
    cb = acum.case_arrays(b.tonal_case(acum.FP64))
    for k in ("X", "N", "V", "PR"):
        assert np.array_equal(ca[k].view(np.uint32), cb[k].view(np.uint32)), k


@pytest.mark.parametrize("lowpass,periodic", [(1, 1), (0, 0), (1, 0)])
def test_gpu_preprocess_filters_and_ends(ctx, tmp_path, lowpass, periodic):
    d = tmp_path / "run"
    acum.Run.synth(acum.synth_params(1, nChord=4, nSpan=3, nTime=40)).save(str(d))
    txt = (d / "input").read_text()
    txt = txt.replace("lowpass 0", f"lowpass {lowpass}").replace("periodic 1", f"periodic {periodic}")
    (d / "input").write_text(txt)
    a = acum.Run.load(str(d)).preprocess()
    b = acum.Run.load(str(d)).preprocess_gpu(ctx)
    ca = acum.case_arrays(a.tonal_case(acum.FP64))
    cb = acum.case_arrays(b.tonal_case(acum.FP64))
    for k in ("X", "N", "V", "PR"):
        assert np.array_equal(ca[k].view(np.uint32), cb[k].view(np.uint32)), k


def test_gpu_preprocess_limits(ctx):
    run = acum.Run.synth(acum.synth_params(1, nChord=2, nSpan=2, nTime=4100))
    with pytest.raises(acum.AcumError) as e:
        run.preprocess_gpu(ctx)
    assert e.value.code == acum.ACUM_ERR_ARG
    run = acum.Run.synth(acum.synth_params(1, nChord=2, nSpan=2, nTime=16)).preprocess()
    with pytest.raises(acum.AcumError):
        run.preprocess_gpu(ctx)


@pytest.mark.parametrize("config,over", [(1, dict(nChord=4, nSpan=3, nTime=36, nObs=8)),
                                         (2, dict(nChord=5, nSpan=4, nTime=64, nObsAz=6, nObsEl=3))])
def test_stage_raw_run_equals_host_preprocessed(ctx, config, over):
    """acum_run_stage on a run that was never preprocessed (preprocessing on the GPU, data kept on
    the device) gives bit-identical histories to staging the host-preprocessed run."""
    raw = acum.Run.synth(acum.synth_params(config, **over))
    host = acum.Run.synth(acum.synth_params(config, **over)).preprocess()
    c = host.tonal_case(acum.MIXED)
    h_raw = ctx.stage_run(raw)
    h_host = ctx.stage([c])
    for prec in (acum.MIXED, acum.FP64):
        ctx.run(h_raw, prec)
        a = ctx.fetch(h_raw, 0, c.nTime, c.nObs)
        ctx.run(h_host, prec)
        b = ctx.fetch(h_host, 0, c.nTime, c.nObs)
        for x, y in zip(a, b):
            assert np.array_equal(x, y)
    ctx.release(h_raw)
    ctx.release(h_host)
