"""The drop-in program end to end on the GPU: capow_amd/bin/capow with the
reference's argv, parameter file and ASCII catalogue writes the reference's
P_l(k) file (src/main.c:1058-1101); the oracle's output file for the same
inputs is the expected result.  FP64 plain form (the program's default):
mode counts bit-exact, P_l within 1e-10 of the bin's scale (side dump with
%.17g), and the %g text files equal up to the last printed digit."""
import os
import subprocess
import sys

import numpy as np
import pytest

from capow_amd import mocks
from tests.hostfiles import write_catalogue, write_params

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "capow_amd", "bin", "capow")
L = 1000.0

CASES = [
    # name, tag overrides
    ("tsc_int_mom1_odd", dict(NX=32, NY=32, NZ=32, InterpOrder=3, DoInterlacing=1, Momentum=1,
                              Odd_Multipoles=1, NK=12, Maxk=0.12)),
    ("cic_los3_rsd", dict(NX=32, NY=32, NZ=32, InterpOrder=2, DoInterlacing=0, Momentum=0,
                          LOS=3, NK=10, Maxk=0.1)),
    ("ngp_int_2d", dict(NX=24, NY=24, NZ=24, InterpOrder=1, DoInterlacing=1, Momentum=2,
                        Output2D=1, NMU=5, NK=8, Maxk=0.08)),
]


def _rows(path):
    out = []
    with open(path) as f:
        for line in f:
            if line.startswith("#") or not line.strip():
                continue
            out.append([float(x) for x in line.split()])
    return np.array(out)


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_program_output_matches_oracle(case, oracle, tmp_path):
    name, over = case
    print("host program case", name, flush=True)
    sys.stdout.flush()
    pos, _ = mocks.uniform(15000, L, 17)
    vel = np.random.default_rng(5).normal(0, 200.0, size=pos.shape)
    if over.get("LOS", 0) == 0:
        pos = np.clip(pos, 0, np.nextafter(L, 0))
    params = write_params(str(tmp_path / "run_parameters.dat"), **over)
    cat = write_catalogue(str(tmp_path / "cat.txt"), pos, vel)
    out = str(tmp_path / "pk.txt")
    env = dict(os.environ, CAPOW_PRECISE_OUTPUT="1")
    r = subprocess.run([EXE, params, cat, "unused.txt", out], capture_output=True, text=True,
                       env=env, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    tags = dict(NX=32, NY=32, NZ=32, InterpOrder=2, DoInterlacing=0, Momentum=0, LOS=0, NK=10,
                Mink=0.0, Maxk=0.1, Odd_Multipoles=0, Output2D=0, NMU=4, OutputLog=0,
                Omega_m=0.3089, Redshift=0.1)
    tags.update(over)
    op = oracle.params(nx=tags["NX"], ny=tags["NY"], nz=tags["NZ"], xmax=L, ymax=L, zmax=L,
                       interp_order=tags["InterpOrder"], do_interlacing=tags["DoInterlacing"],
                       los=tags["LOS"], momentum=tags["Momentum"], nk=tags["NK"],
                       nmu=tags["NMU"], output_2d=tags["Output2D"], output_log=tags["OutputLog"],
                       odd_multipoles=tags["Odd_Multipoles"], mink=tags["Mink"],
                       maxk=tags["Maxk"], omega_m=tags["Omega_m"], redshift=tags["Redshift"])
    rc, b, st, alpha, shot, norm = oracle.run(op, pos, vel)
    assert rc == 0
    oracle.finalize(op, shot, norm, b)
    ref = str(tmp_path / "ref.txt")
    oracle.write_output(op, shot, norm, b, ref, precise=True)
    oracle.write_output(op, shot, norm, b, ref + ".g", precise=False)
    got = _rows(out + ".precise")
    exp = _rows(ref)
    assert got.shape == exp.shape
    # columns: k, P0, P2, P4 [, P1, P3], Nmodes ... as written by output_power
    ncol = exp.shape[1]
    nm = exp[:, -1]
    assert np.array_equal(got[:, -1], nm), "mode counts"
    # the bin scale: the largest multipole of the spectrum (for the cross
    # spectrum of Momentum 1 that is the dipole) plus the shot-noise level
    scale = np.abs(exp[:, 1:-1]).max() + abs(shot / norm)
    assert np.abs(got[:, :-1] - exp[:, :-1]).max() <= 1e-10 * scale
    g1, g2 = _rows(out), _rows(ref + ".g")
    assert g1.shape == g2.shape
    assert np.allclose(g1, g2, rtol=2e-5, atol=1e-5 * scale)
    if tags["Output2D"]:
        a2, b2 = _rows(out + "_2D"), _rows(ref + ".g_2D")
        assert a2.shape == b2.shape and np.allclose(a2, b2, rtol=2e-5, atol=1e-5 * scale)


@pytest.mark.parametrize("case", CASES[:2], ids=[c[0] for c in CASES[:2]])
def test_program_fp32_form_within_1e4_of_fp64(case, tmp_path):
    """CAPOW_PRECISION=fp32 runs the program's FP32 form (FP32 grids and FFT;
    on a 64^3 grid the x pass fused with the binning): same mode counts as
    the FP64 program output, P_l within 1e-4 of the bin's scale for
    k < k_Nyq/2, on a catalogue whose numbers are float32-representable."""
    name, over = case
    over = dict(over, NX=64, NY=64, NZ=64)
    print("host program fp32 case", name, flush=True)
    pos, _ = mocks.uniform(40000, L, 23)
    vel = np.random.default_rng(6).normal(0, 200.0, size=pos.shape)
    pos = pos.astype(np.float32).astype(np.float64)
    vel = vel.astype(np.float32).astype(np.float64)
    if over.get("LOS", 0) == 0:
        pos = np.clip(pos, 0, np.float32(np.nextafter(np.float32(L), 0)))
    params = write_params(str(tmp_path / "run_parameters.dat"), **over)
    cat = write_catalogue(str(tmp_path / "cat.txt"), pos, vel)
    outs = {}
    for prec in ("fp64", "fp32"):
        out = str(tmp_path / ("pk_%s.txt" % prec))
        env = dict(os.environ, CAPOW_PRECISE_OUTPUT="1", CAPOW_PRECISION=prec)
        r = subprocess.run([EXE, params, cat, "unused.txt", out], capture_output=True, text=True,
                           env=env, timeout=120)
        assert r.returncode == 0, r.stdout + r.stderr
        outs[prec] = _rows(out + ".precise")
        if prec == "fp64":
            with open(out + ".precise") as f:
                head = [ln for ln in f if ln.startswith("# shot")]
            shot_norm = float(head[0].split("=")[1])
    a, b = outs["fp32"], outs["fp64"]
    assert a.shape == b.shape
    assert np.array_equal(a[:, -1], b[:, -1]), "mode counts"
    knyq = np.pi * 64 / L
    sel = (b[:, 0] < knyq / 2) & (b[:, -1] > 0)
    assert sel.sum() >= 4
    scale = np.abs(b[sel, 1]) + abs(shot_norm)      # |P_0| + shot/norm per bin
    rel = (np.abs(a[sel, 1:-1] - b[sel, 1:-1]).max(axis=1) / scale).max()
    print("  max rel", rel, flush=True)
    assert rel <= 1e-4
