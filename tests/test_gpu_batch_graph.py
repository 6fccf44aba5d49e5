"""Graph mode of the batch runner (capow_batch_set_graph): a catalogue list
runs eagerly once, is captured into one HIP graph over every stream, and is
replayed while the list is unchanged.  Every replay is checked bitwise
against an eager batch (same kernels, same order of sums) on the same
contents, with the program's own eager work (a PowerSpectrum run, torch
kernels) and synchronisations between replays, catalogue contents rewritten
in place between replays, and a change of list (which runs eagerly and is
captured anew).  No counterpart in the reference, which runs one catalogue
per process (src/main.c:28-298)."""
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
            dict(nx=48, ny=48, nz=48, xmax=L, ymax=L, zmax=L, interp_order=3, do_interlacing=1,
                 momentum=1, extra_spectra=3, nk=25, maxk=0.3, odd_multipoles=1,
                 precision=api.FP32)]


def _same(ra, rb, what):
    assert len(ra) == len(rb)
    for i, (a, b) in enumerate(zip(ra, rb)):
        assert np.array_equal(a["nmodes"], b["nmodes"]), (what, i)
        assert a["codes"] == b["codes"], (what, i)
        for code in a["codes"]:
            assert np.array_equal(a[code]["pk"].view(np.uint64), b[code]["pk"].view(np.uint64)), \
                (what, i, code)


def test_batch_graph_replays_match_eager():
    import torch
    say("batch graph: 8 catalogues, 2 parameter sets, 4 streams")
    counts = [3000, 5000, 4000, 6000, 2500, 4500, 3500, 5500]
    cats = [mocks.mixed_mock_torch(c, L, seed=500 + i) for i, c in enumerate(counts)]
    which = [i % 2 for i in range(len(cats))]
    eager = api.Batch(_configs(), max_particles=6000, nstreams=4)
    graph = api.Batch(_configs(), max_particles=6000, nstreams=4)
    graph.set_graph(True)
    ref, ref_stats, _ = eager.run(cats, which)
    # first run of the list: eager, then captured
    r, st, ms = graph.run(cats, which)
    _same(ref, r, "first run")
    assert (ms >= 0).all()
    for rep in range(3):
        say("  replay", rep)
        r, st, ms = graph.run(cats, which)
        _same(ref, r, "replay %d" % rep)
        assert (ms == -1).all()
        for a, b in zip(ref_stats, st):
            assert a.ngrid == b.ngrid and a.nsq == b.nsq and a.nout == b.nout
        # the program's own eager work between replays
        p = api.make_params(**_configs()[1])
        ps = api.PowerSpectrum(p)
        ps.set_particles(cats[1].cpu().numpy())
        alone = ps.power()
        ps.close()
        _same([ref[1]], [alone], "eager PowerSpectrum between replays")
        x = torch.randn(1 << 20, device="cuda")
        assert torch.isfinite((x * x).sum()).item()
        torch.cuda.synchronize()

    # contents rewritten in place: the replay reads the new catalogues
    say("  contents rewritten in place")
    for i, c in enumerate(counts):
        cats[i].copy_(mocks.mixed_mock_torch(c, L, seed=900 + i))
    torch.cuda.synchronize()
    ref2, _, _ = eager.run(cats, which)
    r2, _, ms = graph.run(cats, which)
    assert (ms == -1).all()
    _same(ref2, r2, "replay after rewrite")
    assert not np.array_equal(ref2[0][0]["pk"], ref[0][0]["pk"])

    # another list (fewer catalogues, other order): eager + capture, replay
    say("  another list")
    sub = [cats[5], cats[2], cats[7]]
    subw = [which[5], which[2], which[7]]
    ref3, _, _ = eager.run(sub, subw)
    for rep in range(2):
        r3, _, ms = graph.run(sub, subw)
        _same(ref3, r3, "sub list run %d" % rep)
        assert (ms == -1).all() == (rep == 1)
    # back to the full list: captured again, then replayed
    for rep in range(2):
        r4, _, ms = graph.run(cats, which)
        _same(ref2, r4, "full list again %d" % rep)
        assert (ms == -1).all() == (rep == 1)
    graph.set_graph(False)
    r5, _, ms = graph.run(cats, which)
    _same(ref2, r5, "graph off")
    assert (ms >= 0).all()
    eager.close()
    graph.close()
