"""Oracle (test infrastructure only): CPU restatement of the reference's
survey ingest, used by the tests to check the library's host C versions.

Functions and the reference code each follows:
- comoving_distance:  src/read_data.c:363-378 (gsl_integration_qags, epsrel
  1e-7).  scipy.integrate.quad is QUADPACK's qags on a finite interval, the
  same routine GSL restates.  Pinned by the closed forms at Omega_m = 1
  (2 c/100 (1 - (1+z)^-1/2)) and Omega_m = 0 (c z / 100).
- zr_table:           src/main.c:1230-1242.
- natural_spline:     gsl_interp_cspline (natural boundary).  Pinned by its
  defining properties in the tests (through the nodes, zero end curvature,
  exact on linear data, C2 at the nodes).
- read_survey:        src/read_data.c:184-361 (the line rules, Coord_Type
  conversion, redshift cut, bounds).  Pinned by hand-computed objects.
- compute_nbar:       src/read_data.c:383-440.  Pinned by a uniform shell
  catalogue whose density is known.
- compute_fkp:        src/read_data.c:444-463.  Pinned by hand values.
- survey_sums:        src/main.c:149-204.  Pinned by hand values.

The product path never imports this module.
"""
import math

import numpy as np
from scipy import integrate
from scipy.interpolate import CubicSpline

LIGHT_SPEED = 299792.458  # src/vars.h:101


def comoving_distance(z, omega_m):
    f = lambda t: 1.0 / math.sqrt(omega_m * (1 + t) ** 3 + (1.0 - omega_m))
    res, _ = integrate.quad(f, 0.0, z, epsabs=0.0, epsrel=1e-7, limit=1000)
    return LIGHT_SPEED * res / 100.0


def zr_table(omega_m, n=10000, zmax=10.0):
    w = zmax / (n - 1)
    z = np.array([i * w for i in range(n)])
    r = np.array([0.0] + [comoving_distance(z[i], omega_m) for i in range(1, n)])
    return z, r


def natural_spline(x, y):
    return CubicSpline(np.asarray(x, float), np.asarray(y, float), bc_type="natural")


class Cosmo:
    def __init__(self, omega_m):
        self.z, self.r = zr_table(omega_m)
        self.r_of_z = natural_spline(self.z, self.r)
        self.z_of_r = natural_spline(self.r, self.z)


def _numbers(line, maxn):
    out = []
    for tok in line.split():
        if len(out) == maxn:
            break
        try:
            out.append(float(tok))
        except ValueError:
            break
    return out


def read_survey(path, randoms, grid, sp, cosmo):
    """grid: dict nx ny nz xmin.. interp_order do_interlacing; sp: dict of
    survey tags.  Returns (objects as list of dicts, zmin, zmax) or raises
    ValueError('parse' | 'box')."""
    n = [grid["nx"], grid["ny"], grid["nz"]]
    mn = [grid["xmin"], grid["ymin"], grid["zmin"]]
    mx = [grid["xmax"], grid["ymax"], grid["zmax"]]
    lo, hi = [], []
    for a in range(3):
        d = (mx[a] - mn[a]) / n[a]
        lo.append(mn[a] + (grid["interp_order"] - 1) * d)
        h = mx[a] - (grid["interp_order"] - 1) * d
        if grid["do_interlacing"]:
            h -= 0.5 * d
        hi.append(h)
    data = open(path, "rb").read().decode()
    lines = data.split("\n")
    lines = lines[:-1]  # the piece after the last newline is not read
    out, zr = [], [math.inf, -math.inf]
    leading = True
    for line in lines:
        if line == "":
            leading = False
            continue
        if leading and line.startswith("#"):
            continue
        leading = False
        v = _numbers(line, 13)
        if randoms:
            if len(v) < 8:
                raise ValueError("parse")
            tx, ty, tz, tw, tnbar, tlerr = v[0], v[1], v[2], v[3], v[4], v[6]
            tlog = 0.0
        else:
            if len(v) != 12:
                raise ValueError("parse")
            tx, ty, tz, tw, tnbar, tlog, tlerr = v[0], v[1], v[2], v[3], v[4], v[6], v[7]
        if sp["coord_type"] == 0:
            tred = float(cosmo.z_of_r(math.sqrt(tx * tx + ty * ty + tz * tz)))
            if tred < sp["redmin_in"] or tred > sp["redmax_in"]:
                continue
        else:
            tred = tz
            if tred < sp["redmin_in"] or tred > sp["redmax_in"]:
                continue
            if tw == 0:
                continue
            dist = float(cosmo.r_of_z(tred))
            ra, dec = tx, ty
            if sp["coord_type"] == 1:
                ra *= math.pi / 180.0
                dec *= math.pi / 180.0
            tx = dist * math.cos(dec) * math.cos(ra)
            ty = dist * math.cos(dec) * math.sin(ra)
            tz = dist * math.sin(dec)
        p = (tx, ty, tz)
        if any(p[a] < lo[a] or p[a] >= hi[a] for a in range(3)):
            raise ValueError("box")
        o = dict(x=tx, y=ty, z=tz, redshift=tred, weight=tw, weight_pv=tw,
                 nbar=tnbar if sp["nbar_column"] > 0 else 0.0, pv=0.0, pverr=0.0)
        if sp["momentum"]:
            q = 1.5 * sp["omega_m"] - 1.0
            zm = tred * (1.0 + 0.5 * (1.0 - q) * tred - 1.0 / 6.0 * (2.0 - q - 3.0 * q * q) * tred * tred)
            f = LIGHT_SPEED * zm * math.log(10.0) / (1.0 + zm)
            o["pv"] = 0.0 if randoms else f * tlog
            o["pverr"] = f * tlerr
        out.append(o)
        zr = [min(zr[0], tred), max(zr[1], tred)]
    return out, zr


def compute_nbar(sp, cosmo, zr, data, randoms):
    nbins = int(math.ceil((sp["redmax_in"] - sp["redmin_in"]) / 0.002))
    zmin, zmax = zr
    w = (zmax - zmin) / float(np.float32(nbins))
    znbar = np.zeros(nbins + 2)
    nbar = np.zeros(nbins + 2)
    for o in randoms[:len(data)]:
        b = int(math.floor((o["redshift"] - zmin) / w)) + 1
        if b == nbins + 1:
            b -= 1
        nbar[b] += 1
    for i in range(nbins):
        ro = float(cosmo.r_of_z((i + 1) * w + zmin))
        ri = float(cosmo.r_of_z(i * w + zmin))
        vol = sp["sky_area"] / (180.0 * 180.0) * math.pi * math.pi * (ro ** 3 - ri ** 3) / 3.0
        znbar[i + 1] = (i + 0.5) * w + zmin
        nbar[i + 1] /= vol
    znbar[0] = zmin - 0.5 * w
    znbar[nbins + 1] = zmax + 0.5 * w
    s = natural_spline(znbar, nbar)
    for o in data + randoms:
        o["nbar"] = float(s(o["redshift"]))
    return znbar, nbar


def compute_fkp(sp, objs):
    for o in objs:
        vv = o["pverr"] ** 2 + 300.0 ** 2
        if sp["momentum"] == 0:
            o["weight"] *= 1.0 / (1.0 + o["nbar"] * sp["fkp_pk"])
        elif sp["momentum"] == 1:
            o["weight"] *= 1.0 / (vv + o["nbar"] * sp["fkp_pk_mom"])
        else:
            tw = 1.0 / (1.0 + o["nbar"] * sp["fkp_pk"])
            tp = 1.0 / (vv + o["nbar"] * sp["fkp_pk_mom"])
            tc = 1.0 / math.sqrt(1.0 + tw * tp * o["nbar"] ** 2 * sp["fkp_pk_cross"] ** 2)
            o["weight_pv"] = o["weight"] * tp * tc
            o["weight"] *= tw * tc


def survey_sums(sp, data, randoms):
    m = sp["momentum"]
    d_nbw = sum(o["weight"] for o in data)
    d_nbwsq = sum(o["weight"] ** 2 for o in data)
    d_vr = sum(o["weight"] * o["weight_pv"] * o["pv"] for o in data) if m else 0.0
    d_vrsq = sum(o["weight"] ** 2 * o["pv"] ** 2 for o in data) if m else 0.0
    r_nbw = sum(o["weight"] for o in randoms)
    r_nbwsq = sum(o["weight"] ** 2 for o in randoms)
    r_nbsqwsq = sum(o["weight"] ** 2 * o["nbar"] for o in randoms)
    r_pv = sum(o["weight_pv"] ** 2 * o["nbar"] for o in randoms) if m == 2 else 0.0
    alpha = d_nbw / r_nbw
    r_nbwsq *= alpha * alpha
    r_nbsqwsq *= alpha
    shot = norm = 0.0
    if m == 0:
        shot, norm = d_nbwsq + r_nbwsq, r_nbsqwsq
    elif m == 1:
        shot, norm = d_vrsq, r_nbsqwsq
    elif m == 2:
        shot, norm = d_vr, math.sqrt(r_nbsqwsq * r_pv)
    elif m == 3:
        shot, norm = d_vrsq, math.sqrt(r_nbsqwsq * r_pv)
    return dict(alpha=alpha, shot=shot, norm=norm)


def deposit(op, sp, data, randoms, alpha, grids, add_to_grid):
    """assign_survey_data for data (prefactor 1) and randoms (-alpha),
    src/read_data.c:464-486 and src/main.c:94, 228, through the oracle's
    add_to_grid.  grids = [A, A interlaced, B, B interlaced] (None where
    absent)."""
    d = [(op.xmax - op.xmin) / op.nx, (op.ymax - op.ymin) / op.ny, (op.zmax - op.zmin) / op.nz]
    m = sp["momentum"]

    def put(o, w, vr, g0, g1):
        add_to_grid(op, o["x"], o["y"], o["z"], w, vr, g0)
        if op.do_interlacing:
            add_to_grid(op, o["x"] + d[0] / 2, o["y"] + d[1] / 2, o["z"] + d[2] / 2, w, vr, g1)

    for o in data:
        if m == 1:
            put(o, o["weight"], o["pv"], grids[0], grids[1])
        else:
            put(o, o["weight"], 1.0, grids[0], grids[1])
            if m:
                put(o, o["weight_pv"], o["pv"], grids[2], grids[3])
    if m != 1:
        for o in randoms:
            put(o, -alpha * o["weight"], 1.0, grids[0], grids[1])
