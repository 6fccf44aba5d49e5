"""The autotuner (capow_autotune, capow_batch_autotune): it reports every
candidate it timed and its choice, and the results after tuning equal the
untuned ones -- bitwise when the fused x-pass binning keeps its setting,
within the FP32 rounding level (2e-6 of the bin's scale) when it changes.
The reference has nothing to tune (FFTW_ESTIMATE plans,
src/main.c:1200-1217)."""
import sys

import numpy as np
import pytest

from capow_amd import api, mocks

pytestmark = pytest.mark.gpu

L = 1000.0


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_autotune_reports_and_keeps_results():
    say("autotune 96^3")
    n = 96
    p = api.make_params(nx=n, ny=n, nz=n, xmax=L, ymax=L, zmax=L, interp_order=3, do_interlacing=1,
                        momentum=1, extra_spectra=3, nk=20, maxk=0.2, precision=api.FP32)
    pos, vel = mocks.zeldovich(48, L, 3)
    xyzv = mocks.xyzv(pos, vel, 0).astype(np.float32)
    ps = api.PowerSpectrum(p)
    ps.set_particles(xyzv)
    before = ps.power()
    rep = ps.autotune(reps=2)
    say(rep)
    for k in ("sort_blocks=", "sort_lds=", "dep_halo=", "dep_bank=", "fft_runs=", "fft_order=",
              "chosen:"):
        assert k in rep, k
    xbin_now = "xbin=1" in rep.split("chosen:")[1]
    after = ps.power()
    ps.close()
    assert np.array_equal(before["nmodes"], after["nmodes"])
    for code in before["codes"]:
        a, b = before[code]["pk"], after[code]["pk"]
        if xbin_now:
            assert np.array_equal(a.view(np.uint64), b.view(np.uint64)), code
        else:
            scale = np.abs(a[0]).max()
            assert np.abs(a - b).max() <= 2e-6 * scale, code


def test_batch_autotune_keeps_results():
    import torch
    say("batch autotune")
    Lb = 500.0
    cfgs = [dict(nx=32, ny=32, nz=32, xmax=Lb, ymax=Lb, zmax=Lb, interp_order=2, do_interlacing=1,
                 momentum=0, nk=20, maxk=0.2, precision=api.FP32),
            dict(nx=48, ny=48, nz=48, xmax=Lb, ymax=Lb, zmax=Lb, interp_order=2, do_interlacing=1,
                 momentum=0, nk=30, maxk=0.3, precision=api.FP32)]
    counts = [3000, 5000, 4000, 6000, 2500, 4500, 3500, 5500]
    cats = [mocks.mixed_mock_torch(c, Lb, seed=300 + i) for i, c in enumerate(counts)]
    which = [i % 2 for i in range(len(cats))]
    b = api.Batch(cfgs, max_particles=6000, nstreams=4)
    r0, _, _ = b.run(cats, which)
    rep = b.autotune(cats, which, reps=2)
    say(rep)
    assert "streams=2" in rep and "chosen: streams=" in rep
    assert "chosen: stack=" in rep
    if "chosen: stack=0" not in rep:
        # stacks chosen: each set's stack context is tuned again from the
        # library defaults, every knob (the sort's LDS aggregation and the
        # FFT write runs included) timed on the whole list
        for q in (0, 1):
            for knob in ("sort_blocks=12", "sort_lds=1", "sort_lds=0", "dep_halo=1", "dep_bank=1",
                         "fft_runs=0", "fft_runs=1", "fft_order=5"):
                assert "stack config %d %s:" % (q, knob) in rep, (q, knob)
            assert "chosen: stack config %d sort_blocks=" % q in rep
    r1, _, _ = b.run(cats, which)
    b.close()
    torch.cuda.synchronize()
    for a, c in zip(r0, r1):
        assert np.array_equal(a["nmodes"], c["nmodes"])
        for code in a["codes"]:
            scale = np.abs(a[code]["pk"][0]).max()
            assert np.abs(a[code]["pk"] - c[code]["pk"]).max() <= 2e-6 * scale


def test_batch_autotune_retunes_stacks():
    """Many small catalogues on grids every stage can stack (32^3, 64^3): the
    per-catalogue launches cost more than their work, so stacks win and each set's stack context is tuned again (every
    knob from the library defaults, timed on the whole list); the stacked
    bins after tuning are bitwise those of the untuned per-catalogue run
    (every re-tuned knob keeps the sums; the stack keeps the fused binning)."""
    import torch
    say("batch autotune, 48 small catalogues")
    Lb = 500.0
    cfgs = [dict(nx=32, ny=32, nz=32, xmax=Lb, ymax=Lb, zmax=Lb, interp_order=2, do_interlacing=1,
                 momentum=0, nk=20, maxk=0.2, precision=api.FP32),
            dict(nx=64, ny=64, nz=64, xmax=Lb, ymax=Lb, zmax=Lb, interp_order=2, do_interlacing=1,
                 momentum=0, nk=30, maxk=0.3, precision=api.FP32)]
    counts = [1500 + 37 * i for i in range(48)]
    cats = [mocks.mixed_mock_torch(c, Lb, seed=500 + i) for i, c in enumerate(counts)]
    which = [i % 2 for i in range(len(cats))]
    b = api.Batch(cfgs, max_particles=max(counts), nstreams=4)
    r0, _, _ = b.run(cats, which)
    rep = b.autotune(cats, which, reps=2)
    say(rep)
    assert "chosen: stack=0" not in rep and "chosen: stack=" in rep
    for q in (0, 1):
        for knob in ("sort_blocks=12", "sort_blocks=8", "sort_blocks=4", "sort_lds=1", "sort_lds=0",
                     "dep_halo=1", "dep_halo=0", "dep_bank=1", "dep_bank=0", "fft_runs=0",
                     "fft_runs=1", "fft_order=5", "fft_order=0", "fft_order=7", "fft_order=2"):
            assert "stack config %d %s:" % (q, knob) in rep, (q, knob)
        assert "chosen: stack config %d sort_blocks=" % q in rep
    r1, _, _ = b.run(cats, which)
    b.close()
    torch.cuda.synchronize()
    for i, (a, c) in enumerate(zip(r0, r1)):
        assert np.array_equal(a["nmodes"], c["nmodes"]), i
        for code in a["codes"]:
            assert np.array_equal(a[code]["pk"].view(np.uint64), c[code]["pk"].view(np.uint64)), (i, code)
