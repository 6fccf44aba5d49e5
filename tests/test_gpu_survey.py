"""GPU parity of the survey (Yamamoto) estimator, capow_survey_power,
against the oracle's restatement of src/main.c:472-967 on the same
real-space grids.

Bars: mode counts bit-exact; FP64 |dP_l(k)| <= 1e-10 * s(k) for every l and
bin, with s(k) the bin's monopole power (cross spectra: the geometric mean
of the two fields' monopoles); the FP32 form within 1e-4 * s(k) for
k < k_Nyq/2.  The context's grids are unchanged by the call, and a second
call gives bit-identical sums.
"""
import sys

import numpy as np
import pytest

from capow_amd import api, mocks

pytestmark = pytest.mark.gpu

L = 1000.0


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _oparams(O, p):
    return O.params(nx=p.nx, ny=p.ny, nz=p.nz, xmin=p.xmin, xmax=p.xmax, ymin=p.ymin,
                    ymax=p.ymax, zmin=p.zmin, zmax=p.zmax, interp_order=p.interp_order,
                    do_interlacing=p.do_interlacing, los=p.los, momentum=p.momentum, nk=p.nk,
                    nmu=p.nmu, output_2d=p.output_2d, output_log=p.output_log,
                    odd_multipoles=p.odd_multipoles, mink=p.mink, maxk=p.maxk)


CASES = [
    ("fp64_cic_int_odd", api.FP64, (32, 48, 64), dict(interp_order=2, do_interlacing=1,
                                                      odd_multipoles=1)),
    ("fp64_tsc_mom3", api.FP64, (48, 32, 32), dict(interp_order=3, momentum=3)),
    ("fp64_ngp_mom4_int_odd", api.FP64, (32, 32, 64), dict(interp_order=1, do_interlacing=1,
                                                            momentum=4, odd_multipoles=1)),
    ("fp64_mom1_density_only", api.FP64, (32, 32, 32), dict(interp_order=2, momentum=1)),
    ("fp64_log_bins", api.FP64, (32, 32, 32), dict(interp_order=3, do_interlacing=1,
                                                   output_log=1, mink=-2.2, maxk=-0.6, nk=16)),
    ("fp32_tsc_int_odd", api.FP32, (64, 48, 32), dict(interp_order=3, do_interlacing=1,
                                                      odd_multipoles=1)),
    # register FFT lengths on every axis (the transposed passes)
    ("fp32_reg_cic_int_odd", api.FP32, (64, 32, 64), dict(interp_order=2, do_interlacing=1,
                                                          odd_multipoles=1)),
    ("fp32_reg_tsc_mom3", api.FP32, (32, 64, 32), dict(interp_order=3, momentum=3)),
]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_survey_matches_oracle(case, oracle):
    name, prec, (nx, ny, nz), kw = case
    say("survey case", name)
    base = dict(nx=nx, ny=ny, nz=nz, xmax=L, ymax=L * ny / nx, zmax=L * nz / nx, nk=24,
                maxk=0.24, precision=prec)
    base.update(kw)
    p = api.make_params(**base)
    op = _oparams(oracle, p)
    origin = np.array([-420.0, 1310.0, 275.0])
    rng = np.random.default_rng(nx + 5 * ny + 11 * nz + p.momentum)
    mom = p.momentum not in (0, 1)
    il = bool(p.do_interlacing)
    ps = api.PowerSpectrum(p)
    og = [None] * 4
    ctx_slots = [True, il, p.momentum not in (0, 2, 5), p.momentum not in (0, 2, 5) and il]
    for s in range(4):
        if not ctx_slots[s]:
            continue
        g = oracle.new_grid(op)
        v = oracle.real_view(op, g)
        # float32-representable, with a smooth large-scale part so that the
        # multipoles are not pure noise
        x = np.arange(nx)[:, None, None] / nx
        v[:nx] = (rng.standard_normal(v[:nx].shape) + 3.0 * np.cos(2 * np.pi * x)).astype(np.float32)
        og[s] = g
        ps.set_grid(s, v[:nx])
    ps.mark_assigned()
    used = [og[0], og[1] if il else None, og[2] if mom else None, og[3] if (mom and il) else None]
    ob = oracle.survey_power(op, origin, used)
    r1 = ps.survey_power(origin)
    r2 = ps.survey_power(origin)
    for s in range(4):
        if og[s] is not None:
            back = ps.get_grid(s).reshape(nx, ny, -1)
            assert np.array_equal(back, oracle.real_view(op, og[s])[:nx].astype(ps.dtype)), s
    ps.close()
    assert np.array_equal(r1["pk"], r2["pk"]), name   # reproducible
    assert np.array_equal(r1["nmodes"], ob.nmodes), name
    if mom:
        pa = oracle.survey_power(oracle.params(**{**_kw(op), "momentum": 0}), origin,
                                 [og[0], og[1] if il else None, None, None]).pk[0]
        pb = oracle.survey_power(oracle.params(**{**_kw(op), "momentum": 0}), origin,
                                 [og[2], og[3] if il else None, None, None]).pk[0]
        scale = np.sqrt(np.abs(pa) * np.abs(pb))
    else:
        scale = np.abs(ob.pk[0])
    ls = (0, 1, 2, 3, 4) if p.odd_multipoles else (0, 2, 4)
    if prec == api.FP64:
        tol, sel = 1e-10, np.ones(p.nk, dtype=bool)
    else:
        tol = 1e-4
        kn = np.pi * nx / L
        edges = p.mink + (np.arange(p.nk) + 1) * (p.maxk - p.mink) / p.nk
        if p.output_log:
            edges = 10.0 ** edges
        sel = edges < kn / 2
    for l in ls:
        d = np.abs(r1["pk"][l] - ob.pk[l])[sel]
        worst = (d / np.maximum(scale[sel], 1e-300)).max()
        say("  l", l, "max rel", worst)
        assert np.all(d <= tol * scale[sel]), (name, l, worst)
    for l in (1, 3):
        if not p.odd_multipoles:
            assert not r1["pk"][l].any()


def _kw(op):
    return {f: getattr(op, f) for f, _ in op._fields_}


def test_survey_rejects_momentum_without_grid():
    p = api.make_params(nx=32, ny=32, nz=32, xmax=L, ymax=L, zmax=L, momentum=2, nk=10,
                        maxk=0.1, precision=api.FP64)
    ps = api.PowerSpectrum(p)
    ps.set_grid(0, np.zeros(ps.grid_elems()))
    ps.mark_assigned()
    with pytest.raises(api.CapowError) as e:
        ps.survey_power([0.0, 0.0, 0.0])
    assert e.value.code == 5   # CAPOW_ERR_UNSUPPORTED
    ps.close()


def test_survey_fp32_matches_fp64_on_a_deposited_catalogue():
    """Size-independent checks at 256^3, beyond the oracle's reach: a
    Zel'dovich catalogue deposited by the GPU itself (TSC + interlacing),
    observer outside the box, odd multipoles.  The survey mode counts equal
    the periodic binning's (the same geometry), and the FP32 form agrees
    with the FP64 form within 1e-4 of the bin's monopole for k < k_Nyq/2."""
    say("survey fp32 vs fp64 at 256^3")
    n, Lb = 256, 1000.0
    pos, vel = mocks.zeldovich(128, Lb, seed=5)
    parts = np.ascontiguousarray(np.concatenate([pos, vel[:, 2:3]], axis=1))   # x y z v_z
    res = {}
    for prec in (api.FP64, api.FP32):
        p = api.make_params(nx=n, ny=n, nz=n, xmax=Lb, ymax=Lb, zmax=Lb, interp_order=3,
                            do_interlacing=1, momentum=0, nk=80, maxk=0.8, odd_multipoles=1,
                            precision=prec)
        ps = api.PowerSpectrum(p)
        ps.set_particles(parts.astype(ps.dtype))
        ps.assign()
        res[prec] = ps.survey_power([-500.0, -500.0, -500.0])
        ps.fft()
        res[(prec, "periodic")] = ps.bin()["nmodes"]
        ps.close()
    a, b = res[api.FP32], res[api.FP64]
    assert np.array_equal(a["nmodes"], b["nmodes"])
    assert np.array_equal(b["nmodes"], res[(api.FP64, "periodic")])
    kn = np.pi * n / Lb
    edges = (np.arange(80) + 1) * 0.01
    sel = edges < kn / 2
    scale = np.abs(b["pk"][0])[sel]
    for l in range(5):
        d = np.abs(a["pk"][l] - b["pk"][l])[sel]
        worst = (d / scale).max()
        say("  l", l, "fp32 vs fp64 max rel", worst)
        assert np.all(d <= 1e-4 * scale), (l, worst)


@pytest.mark.parametrize("prec", [api.FP64, api.FP32])
@pytest.mark.parametrize("mom", [0, 3])
def test_mirror_group_survey_binning_matches_row_binning(prec, mom, monkeypatch):
    """The survey binning over mirror groups (symmetric x and y tables:
    the geometry and the k^ factor once per four modes, the mirrors' signs
    from their flipped axes) against the row-by-row binning
    (CAPOW_BIN_MIRROR=0): the same modes and terms summed in another order,
    so the mode counts are equal and every multipole is within 1e-12 (FP64)
    or 1e-6 (FP32) of the largest |monopole|; odd multipoles, interlacing,
    a non-cubic grid with the observer off every axis."""
    say("survey mirror groups", prec, mom)
    nx, ny, nz, Lb = 64, 32, 64, 1000.0
    pos, vel = mocks.zeldovich(32, Lb, seed=11)
    parts = np.ascontiguousarray(np.concatenate([pos, vel[:, 2:3]], axis=1))
    p = api.make_params(nx=nx, ny=ny, nz=nz, xmax=Lb, ymax=Lb, zmax=Lb, interp_order=3,
                        do_interlacing=1, momentum=mom, nk=24, maxk=0.24, odd_multipoles=1,
                        precision=prec)

    def run(mirror):
        if mirror:
            monkeypatch.delenv("CAPOW_BIN_MIRROR", raising=False)
        else:
            monkeypatch.setenv("CAPOW_BIN_MIRROR", "0")
        ps = api.PowerSpectrum(p)
        ps.set_particles(parts.astype(ps.dtype))
        ps.assign()
        r = ps.survey_power([-300.0, 1700.0, 420.0])
        ps.close()
        return r

    grp, rows = run(True), run(False)
    assert np.array_equal(grp["nmodes"], rows["nmodes"])
    scale = np.abs(rows["pk"][0]).max()
    tol = 1e-12 if prec == api.FP64 else 1e-6
    for l in range(5):
        worst = np.abs(grp["pk"][l] - rows["pk"][l]).max() / scale
        say("  l", l, "max rel", worst)
        assert worst <= tol, (l, worst)
