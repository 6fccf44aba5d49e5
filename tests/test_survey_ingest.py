"""Survey ingest on the host (no GPU): the library's z-r table, splines,
survey reader, n(z), FKP weights, sums and packing against the oracle's
restatement (reference src/read_data.c:184-494, src/main.c:149-204,
1230-1242), with the oracle pinned by closed forms and hand cases."""
import ctypes as C
import math

import numpy as np
import pytest

from capow_amd import api
from oracle import survey_ingest as so

LS = so.LIGHT_SPEED


def _d(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def grid(**kw):
    g = dict(nx=32, ny=32, nz=32, xmin=-2000.0, xmax=2000.0, ymin=-2000.0, ymax=2000.0,
             zmin=-2000.0, zmax=2000.0, interp_order=3, do_interlacing=1)
    g.update(kw)
    return g


def survey(**kw):
    s = dict(coord_type=1, momentum=0, nbar_column=1, fkp_column=-1, omega_m=0.31,
             redmin_in=0.0, redmax_in=0.5, sky_area=41253.0, fkp_pk=1.0e4,
             fkp_pk_mom=5.0e9, fkp_pk_cross=1.0e6)
    s.update(kw)
    return s


def cparams(g):
    return api.make_params(**g)


def csurvey(s):
    return api.SurveyParams(**s)


class Cosmo:
    def __init__(self, om):
        self.h = C.c_void_p()
        assert api.lib().capow_cosmo_create(om, C.byref(self.h)) == 0

    def __del__(self):
        api.lib().capow_cosmo_destroy(self.h)


def read(path, randoms, g, s, cosmo, nobj_max=10 ** 6, zr=None):
    rows = C.POINTER(api.SurveyObj)()
    nk, nr = C.c_int64(), C.c_int64()
    zr = np.array([1e30, -1e30]) if zr is None else zr
    xr = np.zeros(6)
    bad = C.create_string_buffer(256)
    badv = np.zeros(4)
    rc = api.lib().capow_read_survey(path.encode(), randoms, C.byref(cparams(g)),
                                     C.byref(csurvey(s)), cosmo.h, nobj_max, C.byref(rows),
                                     C.byref(nk), C.byref(nr), _d(zr), _d(xr), bad, 256, _d(badv))
    objs = None
    if rc == 0:
        arr = (api.SurveyObj * nk.value)()
        C.memmove(arr, rows, C.sizeof(api.SurveyObj) * nk.value)
        api.lib().capow_free_survey(rows)
        objs = arr
    return rc, objs, zr, nr.value, bad.value.decode(), badv


def as_dicts(arr):
    return [{k: getattr(o, k) for k, _ in api.SurveyObj._fields_} for o in arr]


def to_c(objs):
    arr = (api.SurveyObj * max(len(objs), 1))()
    for i, o in enumerate(objs):
        for k, v in o.items():
            setattr(arr[i], k, v)
    return arr


# ---- the oracle's pins ----------------------------------------------------

@pytest.mark.parametrize("z", [0.01, 0.3, 1.0, 4.7, 10.0])
def test_oracle_distance_closed_forms(z):
    # Omega_m = 1: int (1+z)^-3/2 = 2 (1 - (1+z)^-1/2); Omega_m = 0: z
    assert so.comoving_distance(z, 1.0) == pytest.approx(2 * LS / 100 * (1 - (1 + z) ** -0.5), rel=1e-10)
    assert so.comoving_distance(z, 0.0) == pytest.approx(LS / 100 * z, rel=1e-12)


def test_oracle_spline_is_the_natural_spline():
    rng = np.random.default_rng(3)
    x = np.sort(rng.uniform(0, 5, 30))
    y = rng.normal(size=30)
    s = so.natural_spline(x, y)
    assert np.allclose(s(x), y, rtol=0, atol=1e-13)
    assert abs(s(x[0], 2)) < 1e-10 and abs(s(x[-1], 2)) < 1e-10
    # C2 at the interior nodes: the left piece's second derivative at its end
    # equals the right piece's at its start
    h = np.diff(x)
    left = 6 * s.c[0, :-1] * h[:-1] + 2 * s.c[1, :-1]
    assert np.allclose(left, 2 * s.c[1, 1:], rtol=1e-9, atol=1e-9)
    lin = so.natural_spline(x, 3 * x - 2)
    q = np.linspace(x[0], x[-1], 101)
    assert np.allclose(lin(q), 3 * q - 2, atol=1e-12)


# ---- library against the oracle --------------------------------------------

def test_zr_table_closed_forms_and_oracle():
    n = 10000
    z, r = np.zeros(n), np.zeros(n)
    L = api.lib()
    assert L.capow_zr_table(1.0, n, 10.0, _d(z), _d(r)) == 0
    assert np.array_equal(z, np.arange(n) * (10.0 / (n - 1)))   # i * width, src/main.c:1238
    exact = 2 * LS / 100 * (1 - (1 + z) ** -0.5)
    assert np.max(np.abs(r - exact) / np.maximum(exact, 1e-300)) < 1e-13
    assert L.capow_zr_table(0.0, n, 10.0, _d(z), _d(r)) == 0
    assert np.allclose(r, LS / 100 * z, rtol=1e-14, atol=0)
    assert L.capow_zr_table(0.31, n, 10.0, _d(z), _d(r)) == 0
    idx = np.r_[1:40, 40:n:97, n - 1]
    ref = np.array([so.comoving_distance(z[i], 0.31) for i in idx])
    # the reference's qags stops at 1e-7 relative; on this integrand it lands far closer
    assert np.max(np.abs(r[idx] - ref) / ref) < 1e-11


def test_cspline_matches_oracle_and_rejects_outside():
    rng = np.random.default_rng(5)
    x = np.sort(rng.uniform(-3, 7, 200))
    y = np.sin(x) + 0.1 * rng.normal(size=200)
    m = np.zeros(200)
    L = api.lib()
    assert L.capow_cspline_init(200, _d(x), _d(y), _d(m)) == 0
    s = so.natural_spline(x, y)
    assert np.allclose(m, s(x, 2), rtol=1e-9, atol=1e-9)
    out = C.c_double()
    for q in np.r_[x[0], x[-1], x[17], rng.uniform(x[0], x[-1], 300)]:
        assert L.capow_cspline_eval(200, _d(x), _d(y), _d(m), q, C.byref(out)) == 0
        assert out.value == pytest.approx(float(s(q)), rel=1e-11, abs=1e-12)
    assert L.capow_cspline_eval(200, _d(x), _d(y), _d(m), x[-1] + 1e-9, C.byref(out)) == 1
    assert L.capow_cspline_eval(200, _d(x), _d(y), _d(m), x[0] - 1e-9, C.byref(out)) == 1
    xs = x.copy()
    xs[5] = xs[4]
    assert L.capow_cspline_init(200, _d(xs), _d(y), _d(m)) == 1


def _data_line(x, y, z, w=1.0, nb=1e-4, logd=0.0, loge=0.0):
    return "%r %r %r %r %r %r %r %r 0 0 0 0\n" % (x, y, z, w, nb, 2 * nb, logd, loge)


def test_reader_hand_cases(tmp_path):
    """Omega_m = 1 makes r(z) closed-form, so the kept positions are known."""
    g, s = grid(), survey(omega_m=1.0, coord_type=1, momentum=1)
    cosmo = Cosmo(1.0)
    r = lambda z: 2 * LS / 100 * (1 - (1 + z) ** -0.5)
    f = tmp_path / "d.txt"
    f.write_text("# header\n#another\n" + _data_line(0.0, 0.0, 0.1, logd=0.01, loge=0.002)
                 + _data_line(90.0, 0.0, 0.2) + "\n"             # blank line skipped
                 + _data_line(0.0, 90.0, 0.3, w=2.0)
                 + _data_line(10.0, 10.0, 0.7)                   # above REDMAXIN: cut
                 + _data_line(10.0, 10.0, 0.2, w=0.0)            # zero weight: dropped
                 + _data_line(45.0, 0.0, 0.25).rstrip("\n"))     # no newline: not read
    rc, objs, zr, nread, _, _ = read(str(f), 0, g, s, cosmo)
    assert rc == 0 and nread == 5 and len(objs) == 3
    o = as_dicts(objs)
    assert o[0]["x"] == pytest.approx(r(0.1), rel=1e-12) and abs(o[0]["y"]) < 1e-9
    assert o[1]["y"] == pytest.approx(r(0.2), rel=1e-12) and abs(o[1]["x"]) < 1e-9
    assert o[2]["z"] == pytest.approx(r(0.3), rel=1e-12) and o[2]["weight"] == 2.0
    assert o[0]["nbar"] == 1e-4    # the fifth column (src/read_data.c:250)
    assert tuple(zr) == (0.1, 0.3)
    q = 1.5 - 1.0
    zm = 0.1 * (1 + 0.5 * (1 - q) * 0.1 - (2 - q - 3 * q * q) / 6 * 0.01)
    assert o[0]["pv"] == pytest.approx(LS * zm * math.log(10) * 0.01 / (1 + zm), rel=1e-13)
    assert o[0]["pverr"] == pytest.approx(LS * zm * math.log(10) * 0.002 / (1 + zm), rel=1e-13)


def test_reader_errors(tmp_path):
    g, s = grid(), survey()
    cosmo = Cosmo(0.31)
    f = tmp_path / "d.txt"
    f.write_text(_data_line(0.0, 0.0, 0.1) + _data_line(0.0, 0.0, 0.1).replace("\n", " 7\n"))
    rc, _, _, _, bad, _ = read(str(f), 0, g, s, cosmo)
    assert rc == api.CAPOW_ERR_PARSE and bad.endswith(" 7")
    f.write_text("1 2 3 4 5 6 7\n")
    assert read(str(f), 1, g, s, cosmo)[0] == api.CAPOW_ERR_PARSE
    f.write_text("1 2 0.1 4 5 6 7 8 9 10\n")          # randoms: extra columns ignored
    assert read(str(f), 1, g, s, cosmo)[0] == 0
    f.write_text(_data_line(0.0, 0.0, 0.1) + "# late comment\n")
    assert read(str(f), 0, g, s, cosmo)[0] == api.CAPOW_ERR_PARSE
    f.write_text("\n# after a blank line\n")
    assert read(str(f), 0, g, s, cosmo)[0] == api.CAPOW_ERR_PARSE
    # r(0.45) ~ 1210 Mpc/h: inside +-2000, beyond 1300 less the TSC reach and half cell
    f.write_text(_data_line(0.0, 0.0, 0.45))
    assert read(str(f), 0, g, s, cosmo)[0] == 0
    rc, _, _, _, _, bv = read(str(f), 0, grid(xmin=-1300.0, xmax=1300.0), s, cosmo)
    assert rc == api.CAPOW_ERR_OUT_OF_BOX and bv[3] == 0.45
    f.write_text(_data_line(0.0, 0.0, 0.1) * 3)
    assert read(str(f), 0, g, s, cosmo, nobj_max=2)[0] == 3   # CAPOW_ERR_NOMEM
    assert read(str(tmp_path / "missing"), 0, g, s, cosmo)[0] == api.CAPOW_ERR_IO


def _catalogue(path, rng, n, coord_type, randoms, zlo=0.02, zhi=0.48):
    lines = ["# x y z w ...\n"]
    for _ in range(n):
        z = rng.uniform(zlo, zhi)
        if coord_type == 0:
            rr = so.comoving_distance(z, 0.31)
            u = rng.normal(size=3)
            u *= rr / np.linalg.norm(u)
            a, b, c = u
        elif coord_type == 1:
            a, b, c = rng.uniform(0, 360), rng.uniform(-90, 90), z
        else:
            a, b, c = rng.uniform(0, 2 * np.pi), rng.uniform(-np.pi / 2, np.pi / 2), z
        w = rng.uniform(0.5, 2.0)
        vals = [a, b, c, w, rng.uniform(1e-5, 1e-3), 0.0, rng.normal(0, 0.05),
                rng.uniform(0.01, 0.1)]
        if not randoms:
            vals = vals + [0.0, 0.0, 0.0, 0.0]
        lines.append(" ".join(repr(float(v)) for v in vals) + "\n")
    path.write_text("".join(lines))


@pytest.mark.parametrize("coord_type", [0, 1, 2])
@pytest.mark.parametrize("momentum", [0, 1, 3])
def test_pipeline_matches_oracle(tmp_path, coord_type, momentum):
    rng = np.random.default_rng(11 + coord_type * 3 + momentum)
    g = grid()
    s = survey(coord_type=coord_type, momentum=momentum, nbar_column=-1, redmin_in=0.05,
               redmax_in=0.45, sky_area=20000.0)
    cosmo, ocos = Cosmo(0.31), _oracle_cosmo()
    fd, fr = tmp_path / "d.txt", tmp_path / "r.txt"
    _catalogue(fd, rng, 300, coord_type, False)
    _catalogue(fr, rng, 900, coord_type, True)
    rc, d_c, zr, _, _, _ = read(str(fd), 0, g, s, cosmo)
    assert rc == 0
    rc, r_c, zr, _, _, _ = read(str(fr), 1, g, s, cosmo, zr=zr)
    assert rc == 0
    d_o, zo = so.read_survey(str(fd), 0, g, s, ocos)
    r_o, zo2 = so.read_survey(str(fr), 1, g, s, ocos)
    zo = [min(zo[0], zo2[0]), max(zo[1], zo2[1])]
    assert len(d_c) == len(d_o) and len(r_c) == len(r_o) and len(d_o) > 200
    assert zr[0] == pytest.approx(zo[0], rel=1e-11) and zr[1] == pytest.approx(zo[1], rel=1e-11)
    _compare(as_dicts(d_c), d_o, 1e-10)
    _compare(as_dicts(r_c), r_o, 1e-10)
    # n(z), FKP, sums on the library's objects, oracle on the same objects
    d_o, r_o = as_dicts(d_c), as_dicts(r_c)
    L = api.lib()
    nb = L.capow_survey_nbar_bins(C.byref(csurvey(s)))
    assert nb == 200
    zn, nbar = np.zeros(nb + 2), np.zeros(nb + 2)
    assert L.capow_survey_nbar(C.byref(csurvey(s)), cosmo.h, _d(zr), d_c, len(d_c), r_c, len(r_c),
                               _d(zn), _d(nbar), nb + 2) == 0
    zn_o, nbar_o = so.compute_nbar(s, ocos, list(zr), d_o, r_o)
    assert np.allclose(zn, zn_o, rtol=1e-14, atol=0)
    assert np.allclose(nbar, nbar_o, rtol=1e-10, atol=0)
    _compare(as_dicts(d_c), d_o, 1e-9)
    _compare(as_dicts(r_c), r_o, 1e-9)
    for arr, lst in ((d_c, d_o), (r_c, r_o)):
        assert L.capow_survey_fkp(C.byref(csurvey(s)), arr, len(arr)) == 0
        so.compute_fkp(s, lst)
    _compare(as_dicts(d_c), d_o, 1e-9)
    _compare(as_dicts(r_c), r_o, 1e-9)
    t = api.SurveyTotals()
    assert L.capow_survey_sums(C.byref(csurvey(s)), d_c, len(d_c), r_c, len(r_c), C.byref(t)) == 0
    ot = so.survey_sums(s, d_o, r_o)
    for k in ("alpha", "shot", "norm"):
        assert getattr(t, k) == pytest.approx(ot[k], rel=1e-9, abs=0), k


_OCOSMO = {}


def _oracle_cosmo():
    if 0.31 not in _OCOSMO:
        _OCOSMO[0.31] = so.Cosmo(0.31)
    return _OCOSMO[0.31]


def _compare(a, b, tol):
    for oa, ob in zip(a, b):
        for k in oa:
            assert oa[k] == pytest.approx(ob[k], rel=tol, abs=tol * 1e-3), k


def test_nbar_at_bin_centres_is_count_over_volume():
    """The spline passes through its nodes, so an object at a bin centre gets
    count / shell volume exactly; the end nodes are zero."""
    s = survey(redmin_in=0.1, redmax_in=0.11, sky_area=1000.0)   # 5 bins
    cosmo = Cosmo(1.0)
    r = lambda z: 2 * LS / 100 * (1 - (1 + z) ** -0.5)
    zlo, zhi = 0.1, 0.11
    w = (zhi - zlo) / 5
    counts = [3, 1, 4, 1, 5]
    assert api.lib().capow_survey_nbar_bins(C.byref(csurvey(s))) == 5
    rand = [dict(redshift=zlo + (b + 0.5) * w, weight=1.0) for b, c in enumerate(counts) for _ in range(c)]
    data = [dict(redshift=zlo + (b + 0.5) * w, weight=1.0) for b in range(5)]
    data += [dict(redshift=zlo + 2.5 * w, weight=1.0)] * (len(rand) - 5)
    dc, rc_ = to_c(data), to_c(rand)
    zr = np.array([zlo, zhi])
    zn, nbar = np.zeros(7), np.zeros(7)
    assert api.lib().capow_survey_nbar(C.byref(csurvey(s)), cosmo.h, _d(zr), dc, len(data), rc_,
                                       len(rand), _d(zn), _d(nbar), 7) == 0
    for b, c in enumerate(counts):
        vol = 1000.0 / 180 ** 2 * math.pi ** 2 * (r(zlo + (b + 1) * w) ** 3 - r(zlo + b * w) ** 3) / 3
        assert nbar[b + 1] == pytest.approx(c / vol, rel=1e-9)
        assert dc[b].nbar == pytest.approx(c / vol, rel=1e-9)
    assert nbar[0] == 0 and nbar[6] == 0
    assert zn[0] == pytest.approx(zlo - w / 2) and zn[6] == pytest.approx(zhi + w / 2)


def test_fkp_and_sums_hand_values():
    L = api.lib()
    s0 = survey(momentum=0, fkp_pk=1000.0)
    o = to_c([dict(weight=2.0, nbar=1e-3), dict(weight=1.0, nbar=3e-3)])
    assert L.capow_survey_fkp(C.byref(csurvey(s0)), o, 2) == 0
    assert o[0].weight == pytest.approx(1.0) and o[1].weight == pytest.approx(0.25)
    s3 = survey(momentum=3, fkp_pk=1000.0, fkp_pk_mom=1e7, fkp_pk_cross=0.0)
    o = to_c([dict(weight=1.0, nbar=1e-3, pverr=400.0)])
    assert L.capow_survey_fkp(C.byref(csurvey(s3)), o, 1) == 0
    assert o[0].weight == pytest.approx(0.5)
    assert o[0].weight_pv == pytest.approx(1.0 / (400.0 ** 2 + 300.0 ** 2 + 1e4))
    # alpha = 3/6; shot = 1 + 4 + alpha^2 (1+1+4+4); norm = alpha (1e-3 + 1e-3 + 2*4e-3 ... )
    d = to_c([dict(weight=1.0, nbar=1e-3), dict(weight=2.0, nbar=1e-3)])
    r = to_c([dict(weight=1.0, nbar=1e-3), dict(weight=1.0, nbar=2e-3),
              dict(weight=2.0, nbar=1e-3), dict(weight=2.0, nbar=1e-3)])
    t = api.SurveyTotals()
    assert L.capow_survey_sums(C.byref(csurvey(s0)), d, 2, r, 4, C.byref(t)) == 0
    assert t.alpha == pytest.approx(0.5)
    assert t.shot == pytest.approx(5.0 + 0.25 * 10.0)
    assert t.norm == pytest.approx(0.5 * (1e-3 + 2e-3 + 4e-3 + 4e-3))


@pytest.mark.parametrize("momentum", [0, 1, 3])
def test_pack_rows(momentum):
    s = survey(momentum=momentum)
    d = [dict(x=1.0, y=2.0, z=3.0, weight=2.0, weight_pv=0.5, pv=100.0)]
    r = [dict(x=4.0, y=5.0, z=6.0, weight=3.0), dict(x=7.0, y=8.0, z=9.0, weight=1.0)]
    xyzv, w = np.zeros(12), np.zeros(3)
    n = C.c_int64()
    assert api.lib().capow_survey_pack(C.byref(csurvey(s)), to_c(d), 1, to_c(r), 2, 0.25, api.FP64,
                                       xyzv.ctypes.data, w.ctypes.data, C.byref(n)) == 0
    v = {0: 0.0, 1: 100.0, 3: 0.5 * 100.0 / 2.0}[momentum]
    if momentum == 1:   # velocity only: the randoms are not deposited
        assert n.value == 1 and list(xyzv[:4]) == [1, 2, 3, v] and w[0] == 2.0
    else:
        assert n.value == 3
        assert list(xyzv) == [1, 2, 3, v, 4, 5, 6, 0, 7, 8, 9, 0]
        assert list(w) == [2.0, -0.75, -0.25]
    assert api.lib().capow_survey_pack(C.byref(csurvey(survey(momentum=2))), to_c(d), 1, to_c(r), 2,
                                       0.25, api.FP64, xyzv.ctypes.data, w.ctypes.data,
                                       C.byref(n)) == 5


def test_threaded_reader_keeps_file_order(tmp_path, monkeypatch):
    """The reader splits the text over threads (>= 1 MB each): objects come
    back in file order, and the reported failure is the first in file order,
    as the reference's sequential loop meets it."""
    g, s = grid(), survey()
    cosmo = Cosmo(0.31)
    rng = np.random.default_rng(8)
    n = 60000
    z = rng.uniform(0.05, 0.4, n)
    ra, dec = rng.uniform(0, 360, n), rng.uniform(-80, 80, n)
    lines = [_data_line(float(a), float(b), float(c)) for a, b, c in zip(ra, dec, z)]
    f = tmp_path / "big.txt"
    f.write_text("# h\n" + "".join(lines))
    assert f.stat().st_size > 5 << 20     # five threads of ~12000 lines
    res = {}
    for nt in ("1", "8"):
        monkeypatch.setenv("OMP_NUM_THREADS", nt)
        rc, objs, zr, nread, _, _ = read(str(f), 0, g, s, cosmo)
        assert rc == 0 and nread == n
        res[nt] = (np.array([[o.x, o.y, o.z, o.redshift] for o in objs]), tuple(zr))
    assert np.array_equal(res["1"][0], res["8"][0]) and res["1"][1] == res["8"][1]
    assert np.allclose(res["1"][0][:, 3], z[(z >= 0.0) & (z <= 0.5)])
    bad_parse = "1 2 0.1 1\n"
    far = _data_line(0.0, 0.0, 0.45)
    g_small = grid(xmin=-1300.0, xmax=1300.0)
    monkeypatch.setenv("OMP_NUM_THREADS", "8")
    for first, second, want in ((bad_parse, far, api.CAPOW_ERR_PARSE),
                                (far, bad_parse, api.CAPOW_ERR_OUT_OF_BOX)):
        L2 = list(lines)
        L2[25000] = first
        L2[50000] = second
        f.write_text("".join(L2))
        rc, _, _, nread, badl, _ = read(str(f), 0, g_small, s, cosmo)
        assert rc == want
        if want == api.CAPOW_ERR_PARSE:
            assert badl == bad_parse.strip()
    # NOBJ_Max passed before the first failing line: the capacity error wins
    L2 = list(lines)
    L2[50000] = bad_parse
    f.write_text("".join(L2))
    assert read(str(f), 0, g, s, cosmo, nobj_max=20000)[0] == 3
    assert read(str(f), 0, g, s, cosmo, nobj_max=55000)[0] == api.CAPOW_ERR_PARSE


def test_zero_weight_coord0_row_at_momentum_3(tmp_path):
    """Coord_Type 0 keeps zero-weight objects (src/read_data.c:259-266 has no
    weight test there); at Momentum 3 such an object has weight_pv = 0 too
    (tw, or weight tp tc after FKP, src/read_data.c:460) and deposits nothing
    into either field, so the packed row carries w = 0 and v = 0.  A zero
    weight with a nonzero weight_pv pv cannot be carried as w v: refused."""
    g, s = grid(), survey(coord_type=0, momentum=3, nbar_column=1)
    cosmo = Cosmo(0.31)
    f = tmp_path / "d.txt"
    rr = so.comoving_distance(0.2, 0.31)
    f.write_text("%r 0.0 0.0 1.0 1e-4 2e-4 0.02 0.003 0 0 0 0\n" % rr
                 + "0.0 %r 0.0 0.0 1e-4 2e-4 0.05 0.003 0 0 0 0\n" % rr)
    rc, objs, zr, nread, _, _ = read(str(f), 0, g, s, cosmo)
    assert rc == 0 and nread == 2 and len(objs) == 2
    assert objs[1].weight == 0.0 and objs[1].pv != 0.0
    L = api.lib()
    assert L.capow_survey_fkp(C.byref(csurvey(s)), objs, 2) == 0
    assert objs[1].weight == 0.0 and objs[1].weight_pv == 0.0
    xyzv, w = np.zeros(8), np.zeros(2)
    n = C.c_int64()
    assert L.capow_survey_pack(C.byref(csurvey(s)), objs, 2, to_c([]), 0, 0.5, api.FP64,
                               xyzv.ctypes.data, w.ctypes.data, C.byref(n)) == 0
    assert n.value == 2 and w[1] == 0.0 and xyzv[7] == 0.0
    assert xyzv[3] == pytest.approx(objs[0].weight_pv * objs[0].pv / objs[0].weight, rel=1e-15)
    bad = to_c([dict(x=1.0, y=1.0, z=1.0, weight=0.0, weight_pv=0.5, pv=10.0)])
    assert L.capow_survey_pack(C.byref(csurvey(s)), bad, 1, to_c([]), 0, 0.5, api.FP64,
                               xyzv.ctypes.data, w.ctypes.data, C.byref(n)) == \
        5   # CAPOW_ERR_UNSUPPORTED


def test_threaded_nbar_is_bitwise_thread_independent(monkeypatch):
    """The n(z) spline is evaluated over host threads, one per 65,536
    objects up to OMP_NUM_THREADS: 131,072 data and 262,144 randoms give
    bitwise the same table and per-object nbar with OMP_NUM_THREADS = 1
    (one thread) and 16 (six threads)."""
    s = survey(nbar_column=-1, redmin_in=0.05, redmax_in=0.45, sky_area=20000.0)
    cosmo = Cosmo(0.31)
    rng = np.random.default_rng(12)
    nd, nr = 131072, 262144

    def objs(n):
        arr = (api.SurveyObj * n)()
        z = rng.uniform(0.05, 0.45, n)
        for i in range(n):
            arr[i].redshift = z[i]
            arr[i].weight = 1.0
        return arr

    d0, r0 = objs(nd), objs(nr)
    zr = np.array([0.05, 0.45])
    L = api.lib()
    nb = L.capow_survey_nbar_bins(C.byref(csurvey(s)))
    res = {}
    for nt in ("1", "16"):
        monkeypatch.setenv("OMP_NUM_THREADS", nt)
        d = (api.SurveyObj * nd)()
        r = (api.SurveyObj * nr)()
        C.memmove(d, d0, C.sizeof(d0))
        C.memmove(r, r0, C.sizeof(r0))
        zn, nbar = np.zeros(nb + 2), np.zeros(nb + 2)
        assert L.capow_survey_nbar(C.byref(csurvey(s)), cosmo.h, _d(zr), d, nd, r, nr, _d(zn),
                                   _d(nbar), nb + 2) == 0
        per = np.frombuffer(d, dtype=np.float64).reshape(nd, 9)[:, 6].copy()
        perr = np.frombuffer(r, dtype=np.float64).reshape(nr, 9)[:, 6].copy()
        res[nt] = (zn, nbar, per, perr)
    for a, b in zip(res["1"], res["16"]):
        assert np.array_equal(a.view(np.uint64), b.view(np.uint64))
    assert np.all(res["1"][2] > 0)
