"""The survey branch end to end on the GPU: capow_amd/bin/capow with
`Periodic 0`, `Survey 1` reads a data and a randoms catalogue in (RA, Dec, z),
estimates n(z), applies FKP weights, deposits the data and the -alpha
randoms on the GPU (CAPOW_EDGE_SURVEY), runs the Yamamoto estimator and
writes the reference's P_l(k) file (src/main.c:77-94, 149-204, 226-228,
279-280).  Expected: the oracle's ingest (scipy's QUADPACK for the z-r
integral, scipy's natural spline), the oracle's add_to_grid for the deposit
and its survey estimator.  FP64: mode counts bit-exact, P_l within 1e-9 of
the bin scale (the ingest's table and splines differ from the oracle's at
~1e-12 relative in the positions)."""
import os
import subprocess
import sys

import numpy as np
import pytest

from oracle import survey_ingest as so
from tests.hostfiles import write_params

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "capow_amd", "bin", "capow")
OM = 0.31
TAGS = dict(Periodic=0, Survey=1, NX=32, NY=32, NZ=32, XMIN=-1000.0, XMAX=1000.0, YMIN=-1000.0,
            YMAX=1000.0, ZMIN=-1000.0, ZMAX=1000.0, InterpOrder=3, DoInterlacing=1, NK=12,
            Mink=0.0, Maxk=0.12, Omega_m=OM, Coord_Type=1, REDMININ=0.05, REDMAXIN=0.28,
            NOBJ_Max=100000, NBAR_Column=-1, FKP_Column=-1, SkyArea=41253.0, FKP_Pk=5000.0,
            FKP_Pk_mom=1.0e9, FKP_Pk_cross=1.0e6, X_Origin=0.0, Y_Origin=0.0, Z_Origin=0.0,
            Odd_Multipoles=1)


def _write(path, rng, n, randoms):
    z = rng.uniform(0.05, 0.28, n)
    ra, dec = rng.uniform(0, 360, n), np.degrees(np.arcsin(rng.uniform(-1, 1, n)))
    w = rng.uniform(0.5, 1.5, n)
    cols = [ra, dec, z, w, np.full(n, 1e-4), np.zeros(n), rng.normal(0, 0.05, n),
            rng.uniform(0.01, 0.1, n)]
    if not randoms:
        cols += [np.zeros(n)] * 4
    with open(path, "w") as f:
        f.write("# ra dec z w nbar ...\n")
        for row in zip(*cols):
            f.write(" ".join("%.17g" % v for v in row) + "\n")


def _rows(path):
    return np.array([[float(x) for x in l.split()] for l in open(path)
                     if l.strip() and not l.startswith("#")])


def _oracle(oracle, tags, dfile, rfile):
    g = dict(nx=tags["NX"], ny=tags["NY"], nz=tags["NZ"], xmin=tags["XMIN"], xmax=tags["XMAX"],
             ymin=tags["YMIN"], ymax=tags["YMAX"], zmin=tags["ZMIN"], zmax=tags["ZMAX"],
             interp_order=tags["InterpOrder"], do_interlacing=tags["DoInterlacing"])
    s = dict(coord_type=tags["Coord_Type"], momentum=tags["Momentum"],
             nbar_column=tags["NBAR_Column"], omega_m=OM, redmin_in=tags["REDMININ"],
             redmax_in=tags["REDMAXIN"], sky_area=tags["SkyArea"], fkp_pk=tags["FKP_Pk"],
             fkp_pk_mom=tags["FKP_Pk_mom"], fkp_pk_cross=tags["FKP_Pk_cross"])
    cos = so.Cosmo(OM)
    d, z1 = so.read_survey(dfile, 0, g, s, cos)
    r, z2 = so.read_survey(rfile, 1, g, s, cos)
    so.compute_nbar(s, cos, [min(z1[0], z2[0]), max(z1[1], z2[1])], d, r)
    so.compute_fkp(s, d)
    so.compute_fkp(s, r)
    t = so.survey_sums(s, d, r)
    op = oracle.params(nx=g["nx"], ny=g["ny"], nz=g["nz"], xmin=g["xmin"], xmax=g["xmax"],
                       ymin=g["ymin"], ymax=g["ymax"], zmin=g["zmin"], zmax=g["zmax"],
                       interp_order=g["interp_order"], do_interlacing=g["do_interlacing"], los=3,
                       momentum=tags["Momentum"], nk=tags["NK"], mink=tags["Mink"],
                       maxk=tags["Maxk"], odd_multipoles=tags["Odd_Multipoles"], omega_m=OM)
    m = tags["Momentum"]
    grids = [oracle.new_grid(op), oracle.new_grid(op), None, None]
    if m == 3:
        grids[2], grids[3] = oracle.new_grid(op), oracle.new_grid(op)
    so.deposit(op, s, d, r, t["alpha"], grids, oracle.add_to_grid)
    b = oracle.survey_power(op, [tags["X_Origin"], tags["Y_Origin"], tags["Z_Origin"]], grids)
    return op, b, t, d, r


@pytest.mark.parametrize("momentum", [0, 1])
def test_survey_program_matches_oracle(momentum, oracle, tmp_path):
    print("survey program case: Momentum", momentum, flush=True)
    sys.stdout.flush()
    rng = np.random.default_rng(40 + momentum)
    dfile, rfile = str(tmp_path / "data.txt"), str(tmp_path / "rand.txt")
    _write(dfile, rng, 3000, False)
    _write(rfile, rng, 9000, True)
    tags = dict(TAGS, Momentum=momentum)
    params = write_params(str(tmp_path / "run_parameters.dat"), **tags)
    out = str(tmp_path / "pk.txt")
    env = dict(os.environ, CAPOW_PRECISE_OUTPUT="1")
    r = subprocess.run([EXE, params, dfile, rfile, out], capture_output=True, text=True, env=env,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Computing FKP Weights" in r.stdout and "Calculating Number Density" in r.stdout
    op, b, t, _, _ = _oracle(oracle, tags, dfile, rfile)
    oracle.finalize(op, t["shot"], t["norm"], b)
    ref = str(tmp_path / "ref.txt")
    oracle.write_output(op, t["shot"], t["norm"], b, ref, precise=True)
    got, exp = _rows(out + ".precise"), _rows(ref)
    assert got.shape == exp.shape
    assert np.array_equal(got[:, -1], exp[:, -1]), "mode counts"
    assert exp[:, -1].sum() > 0
    scale = np.abs(exp[:, 1:-1]).max() + abs(t["shot"] / t["norm"])
    assert np.abs(got[:, :-1] - exp[:, :-1]).max() <= 1e-9 * scale


def test_survey_density_and_velocity_fields_match_oracle(oracle, tmp_path):
    """Momentum 3 (density + velocity grids, weight_pv pv in the second):
    the raw estimator sums through the C ABI, since the reference's
    normalisation for this code is zero (src/main.c:166-167, 199-200)."""
    from capow_amd import api
    import ctypes as C
    print("survey ABI case: Momentum 3", flush=True)
    sys.stdout.flush()
    rng = np.random.default_rng(7)
    dfile, rfile = str(tmp_path / "data.txt"), str(tmp_path / "rand.txt")
    _write(dfile, rng, 2000, False)
    _write(rfile, rng, 6000, True)
    tags = dict(TAGS, Momentum=3)
    op, b, t, d, r = _oracle(oracle, tags, dfile, rfile)
    assert t["norm"] == 0.0
    sp = api.SurveyParams(coord_type=1, momentum=3, nbar_column=-1, fkp_column=-1, omega_m=OM,
                          redmin_in=tags["REDMININ"], redmax_in=tags["REDMAXIN"],
                          sky_area=tags["SkyArea"], fkp_pk=tags["FKP_Pk"],
                          fkp_pk_mom=tags["FKP_Pk_mom"], fkp_pk_cross=tags["FKP_Pk_cross"])
    dc = (api.SurveyObj * len(d))(*[api.SurveyObj(**o) for o in d])
    rc_ = (api.SurveyObj * len(r))(*[api.SurveyObj(**o) for o in r])
    xyzv, w = np.zeros(4 * (len(d) + len(r))), np.zeros(len(d) + len(r))
    n = C.c_int64()
    assert api.lib().capow_survey_pack(C.byref(sp), dc, len(d), rc_, len(r), t["alpha"], api.FP64,
                                       xyzv.ctypes.data, w.ctypes.data, C.byref(n)) == 0
    ps = api.PowerSpectrum(nx=32, ny=32, nz=32, xmin=-1000.0, xmax=1000.0, ymin=-1000.0,
                           ymax=1000.0, zmin=-1000.0, zmax=1000.0, interp_order=3,
                           do_interlacing=1, momentum=3, nk=12, mink=0.0, maxk=0.12,
                           odd_multipoles=1, omega_m=OM, precision=api.FP64,
                           edge_mode=api.EDGE_SURVEY)
    ps.set_particles(xyzv.reshape(-1, 4), w)
    ps.assign()
    res = ps.survey_power([0.0, 0.0, 0.0])
    pk, nm = res["pk"], res["nmodes"]
    assert np.array_equal(nm, b.nmodes)
    for l in range(5):
        ref = b.pk[l]
        assert np.abs(pk[l] - ref).max() <= 1e-9 * (np.abs(ref).max() + 1e-300), l
