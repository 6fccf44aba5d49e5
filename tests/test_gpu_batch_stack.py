"""Stacked batches (capow_batch_set_stack): the catalogues of a parameter
set run in stacks, every stage one launch for the whole stack.  The bins
and sums must be bitwise those of the per-catalogue batch path (the same
per-catalogue kernels' arithmetic: the stack only adds the catalogue index
to each launch), for stacks that hold a whole group, groups split over
several stacks, empty and tiny catalogues, weighted and unweighted
catalogues in one list, and graph replays of a stacked list.  No
counterpart in the reference, which runs one catalogue per process
(src/main.c:28-298)."""
import sys

import numpy as np
import pytest

from capow_amd import api, mocks

pytestmark = pytest.mark.gpu

L = 500.0


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _configs():
    return [dict(nx=32, ny=32, nz=32, xmax=L, ymax=L, zmax=L, interp_order=2, do_interlacing=1,
                 momentum=0, nk=20, maxk=0.3, precision=api.FP32),
            dict(nx=64, ny=64, nz=64, xmax=L, ymax=L, zmax=L, interp_order=3, do_interlacing=1,
                 momentum=1, extra_spectra=3, nk=25, maxk=0.3, odd_multipoles=1,
                 precision=api.FP32)]


def _same(ra, rb, what):
    assert len(ra) == len(rb)
    for i, (a, b) in enumerate(zip(ra, rb)):
        assert np.array_equal(a["nmodes"], b["nmodes"]), (what, i)
        assert a["codes"] == b["codes"], (what, i)
        for code in a["codes"]:
            assert np.array_equal(a[code]["pk"].view(np.uint64), b[code]["pk"].view(np.uint64)), \
                (what, i, code, np.abs(a[code]["pk"] - b[code]["pk"]).max())


def _same_stats(sa, sb, what):
    for i, (a, b) in enumerate(zip(sa, sb)):
        for f in ("nread", "nout", "ngrid", "nsq", "vr", "vr2", "vr3", "vr4", "alpha"):
            assert getattr(a, f) == getattr(b, f), (what, i, f)


def _cats(counts, seed):
    return [mocks.mixed_mock_torch(c, L, seed=seed + i) for i, c in enumerate(counts)]


@pytest.mark.parametrize("stack", [16, 3])
def test_stack_matches_streams(stack):
    say("stack %d: 11 catalogues (one empty, one of 7 particles), 2 parameter sets" % stack)
    counts = [3000, 5000, 0, 6000, 2500, 4500, 7, 5500, 4000, 3500, 6000]
    which = [0, 1, 0, 1, 1, 0, 1, 0, 1, 0, 1]
    cats = [mocks.mixed_mock_torch(c, L, seed=300 + i) if c >= 64 else None
            for i, c in enumerate(counts)]
    import torch
    for i, c in enumerate(counts):
        if cats[i] is None:
            g = torch.Generator().manual_seed(40 + i)
            cats[i] = (torch.rand(c, 4, generator=g, dtype=torch.float32) * L).cuda()
    ref_b = api.Batch(_configs(), max_particles=6000, nstreams=4)
    ref, ref_stats, _ = ref_b.run(cats, which)
    b = api.Batch(_configs(), max_particles=6000, nstreams=4)
    b.set_stack(stack)
    r, st, ms = b.run(cats, which)
    _same(ref, r, "stack %d" % stack)
    _same_stats(ref_stats, st, "stack %d" % stack)
    assert (ms >= 0).all()
    # a second run on the same stacks (buffers reused, counters cleared)
    r2, _, _ = b.run(cats[::-1], which[::-1])
    _same(ref[::-1], r2, "stack %d, reversed list" % stack)
    ref_b.close()
    b.close()


def test_stack_weighted_and_unweighted():
    import torch
    say("stack: weighted and unit-weight catalogues of one parameter set")
    counts = [4000, 3000, 5000, 3500, 4500]
    cats = _cats(counts, 700)
    g = torch.Generator().manual_seed(7)
    ws = [(0.5 + torch.rand(c, generator=g, dtype=torch.float32)).cuda() if i % 2 else None
          for i, c in enumerate(counts)]
    which = [1] * len(counts)
    ref_b = api.Batch(_configs(), max_particles=6000, nstreams=2)
    b = api.Batch(_configs(), max_particles=6000, nstreams=2)
    b.set_stack(4)
    ref, ref_stats, _ = ref_b.run(cats, which, weights=ws)
    r, st, _ = b.run(cats, which, weights=ws)
    _same(ref, r, "weighted stack")
    _same_stats(ref_stats, st, "weighted stack")
    ref_b.close()
    b.close()


def test_stack_graph_replays():
    import torch
    say("stack + graph: replays against the eager stacked run")
    counts = [3000, 5000, 4000, 6000, 2500, 4500]
    which = [i % 2 for i in range(len(counts))]
    cats = _cats(counts, 800)
    b = api.Batch(_configs(), max_particles=6000, nstreams=2)
    b.set_stack(8)
    ref, _, _ = b.run(cats, which)
    g = api.Batch(_configs(), max_particles=6000, nstreams=2)
    g.set_stack(8)
    g.set_graph(True)
    r, _, ms = g.run(cats, which)
    _same(ref, r, "first (eager) run")
    for rep in range(3):
        say("  replay", rep)
        r, _, ms = g.run(cats, which)
        _same(ref, r, "replay %d" % rep)
        assert (ms == -1).all()
        x = torch.randn(1 << 20, device="cuda")
        assert torch.isfinite((x * x).sum()).item()
        torch.cuda.synchronize()
    # contents rewritten in place: the replay reads them
    for i, c in enumerate(counts):
        cats[i].copy_(mocks.mixed_mock_torch(c, L, seed=900 + i))
    ref2, _, _ = b.run(cats, which)
    r, _, _ = g.run(cats, which)
    _same(ref2, r, "replay after rewrite")
    b.close()
    g.close()


def test_stack_unsupported_sets_keep_streams():
    say("stack: a parameter set without the fused binning is refused")
    cfg = [dict(nx=48, ny=48, nz=48, xmax=L, ymax=L, zmax=L, interp_order=2, nk=10, maxk=0.3,
                precision=api.FP32)]
    b = api.Batch(cfg, max_particles=1000, nstreams=2)
    with pytest.raises(api.CapowError):
        b.set_stack(4)
    cats = _cats([1000], 5)
    r, _, _ = b.run(cats, [0])
    assert r[0]["nmodes"].sum() > 0
    b.close()


def test_stack_at_c3_sizes():
    """The C3 geometry (128^3 with NK 80 and 192^3 with NK 120, CIC +
    interlacing, L = 500, density) with lognormal catalogue sizes: stacks
    of 3 catalogues against one catalogue per launch, bitwise."""
    say("stack at C3 sizes: 4 + 4 catalogues of 128^3 / 192^3")
    Lb = 500.0
    cfgs = [dict(nx=g, ny=g, nz=g, xmax=Lb, ymax=Lb, zmax=Lb, interp_order=2, do_interlacing=1,
                 momentum=0, nk=nk, maxk=nk * 0.01, precision=api.FP32)
            for g, nk in ((128, 80), (192, 120))]
    counts = mocks.lognormal_counts(8, seed=77, median=1.2e5, lo=6e4, hi=3e5)
    cats = [mocks.mixed_mock_torch(int(n), Lb, seed=2000 + i) for i, n in enumerate(counts)]
    which = [i % 2 for i in range(len(cats))]
    ref_b = api.Batch(cfgs, max_particles=int(max(counts)), nstreams=2)
    ref, ref_stats, _ = ref_b.run(cats, which)
    b = api.Batch(cfgs, max_particles=int(max(counts)), nstreams=2)
    b.set_stack(3)
    r, st, _ = b.run(cats, which)
    _same(ref, r, "C3 sizes")
    _same_stats(ref_stats, st, "C3 sizes")
    ref_b.close()
    b.close()
