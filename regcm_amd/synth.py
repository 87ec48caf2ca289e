"""Synthetic MOLOCH domains: static metrics and initial states.

RegCM builds these from its pre-processed terrain and ICBC files
(Main/mod_params.F90:3316-3395 for the metrics, Main/mod_bdycod.F90:3762-3796
for the hydrostatic Exner function, Main/mod_init.F90:155-220, 941-969 and
1008-1025 for the state and the top sponge).  Without those files this module
synthesises the same arrays from closed forms, with every random draw taken
from a counter-based SplitMix64 stream so that any machine reproduces them.

Array convention (shared with the C-ABI and the oracle): a 3-D field is a
C-ordered numpy array of shape (nk, iy, jx), i.e. Fortran (jx, iy, nk) with j
contiguous, over the full dot-point extent; 2-D fields are (iy, jx); water
species are (nqx, kz, iy, jx).
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

# Share/mod_constants.F90 (non-RCEMIP branch), evaluated in the same order.
EGRAV = 9.80665
RGASMOL = 6.02214076e23 * 1.3806490e-23
AMD = 28.96454
AMW = 18.01528
RGAS = (RGASMOL / AMD) * 1000.0
CPD = 3.5 * RGAS
CVD = 2.5 * RGAS
REGRAV = 1.0 / EGRAV
ROVCP = RGAS * (1.0 / CPD)
CPOVR = CPD / RGAS
GOVCP = EGRAV / CPD
EP1 = AMD / AMW - 1.0
EP2 = AMW / AMD
TZERO = 273.15
P00 = 1.0e5
LRATE = 0.00649
MATHPI = 3.14159265358979323846
DEGRAD = MATHPI / 180.0
EOMEG2 = 2.0 * 7.2921159e-5


# ---------------------------------------------------------------- RNG ----
def splitmix64(seed: int, n: int) -> np.ndarray:
    """n outputs of SplitMix64 started at `seed` (counter form)."""
    with np.errstate(over="ignore"):
        k = np.arange(1, n + 1, dtype=np.uint64)
        z = np.uint64(seed) + k * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform01(seed: int, n: int) -> np.ndarray:
    """Uniform doubles in (0, 1] from the top 53 bits."""
    z = splitmix64(seed, n)
    return ((z >> np.uint64(11)).astype(np.float64) + 1.0) * (1.0 / 9007199254740992.0)


def normal(seed: int, shape, sigma: float = 1.0) -> np.ndarray:
    """Box-Muller normals N(0, sigma^2) of the given shape."""
    n = int(np.prod(shape))
    m = (n + 1) // 2
    u = uniform01(seed, 2 * m)
    u1, u2 = u[0::2], u[1::2]
    r = np.sqrt(-2.0 * np.log(u1))
    z = np.concatenate([r * np.cos(2.0 * math.pi * u2), r * np.sin(2.0 * math.pi * u2)])
    return (sigma * z[:n]).reshape(shape)


# ------------------------------------------------------ zita closed forms --
# Share/mod_zita.F90:54-178
def model_zitaf(kz: int, ztop: float) -> np.ndarray:
    dz = ztop / kz
    zf = np.zeros(kz + 1)
    zf[kz] = 0.0
    zf[0] = ztop
    for k in range(kz, 1, -1):
        zf[k - 1] = zf[k] + dz
    return zf


def model_zitah(kz: int, ztop: float) -> np.ndarray:
    dz = ztop / kz
    zh = np.zeros(kz)
    zh[kz - 1] = dz * 0.5
    zh[0] = ztop - dz * 0.5
    for k in range(kz - 1, 1, -1):
        zh[k - 1] = zh[k] + dz
    return zh


def gzita(zita, ztop, a0):
    ratio = np.asarray(zita, dtype=np.float64) / ztop
    return (((0.0 - 1.0 * a0) * ratio - (3.0 - 2.0 * a0) * (ratio * ratio))
            + (2.0 - 1.0 * a0) * (ratio * ratio * ratio)) + 1.0


def gzitap(zita, ztop, a0):
    ratio = np.asarray(zita, dtype=np.float64) / ztop
    return (((0.0 - 1.0 * a0) * 1.0 - (6.0 - 4.0 * a0) * ratio)
            + (6.0 - 3.0 * a0) * (ratio * ratio)) / ztop


def _exp(x):
    """exp of the C library, as the Fortran runtime calls it (numpy's own
    vector exp differs from it in the last bit for some arguments)."""
    x = np.asarray(x, dtype=np.float64)
    return np.array([math.exp(v) for v in x.ravel()]).reshape(x.shape)


def _zfz(ztop, zh):
    return ztop / (math.exp(ztop / zh) - 1.0)


def bzita(zita, ztop, zh):
    return _zfz(ztop, zh) * (_exp(np.asarray(zita) / zh) - 1.0)


def bzitap(zita, ztop, zh):
    return _zfz(ztop, zh) * _exp(np.asarray(zita) / zh) / zh


def md_fmz(zita, geopot, ztop, zh, a0):
    orog = geopot * REGRAV
    return 1.0 / (gzitap(zita, ztop, a0) * orog + bzitap(zita, ztop, zh))


def md_zeta(zita, geopot, ztop, zh, a0):
    orog = geopot * REGRAV
    return orog * (gzita(zita, ztop, a0) - 1.0) + bzita(zita, ztop, zh)


def pfwsat(t, p):
    """Share/pfwsat.inc, vectorised (used only to build initial humidity)."""
    a = [0.611213476e+03, 0.444007856e+02, 0.143064234e+01, 0.264461437e-01,
         0.305903558e-03, 0.196237241e-05, 0.892344772e-08, -0.373208410e-10,
         0.209339997e-13]
    c = [0.611123516e+03, 0.503109514e+02, 0.188369801e+01, 0.420547422e-01,
         0.614396778e-03, 0.602780717e-05, 0.387940929e-07, 0.149436277e-09,
         0.262655803e-12]
    td = np.minimum(np.maximum(np.asarray(t) - TZERO, -75.0), 100.0)

    def horner(co):
        acc = co[8]
        for n in range(7, -1, -1):
            acc = co[n] + td * acc
        return acc

    es = np.where(td >= 0.0, horner(a), horner(c))
    es = np.minimum(es, 0.15 * np.asarray(p))
    return EP2 * (es / (p - es))


# ------------------------------------------------------------ domain -----
@dataclasses.dataclass
class Grid:
    jx: int
    iy: int
    kz: int
    dx: float
    dtsec: float
    nqx: int = 2
    ipptls: int = 0
    nadv: int = 2
    nsound: int = 5
    do_divdamp: int = 1
    do_divfilter: int = 1
    lrotllr: int = 0
    mo_ztop: float = 30000.0
    mo_h: float = 8000.0
    mo_a0: float = 0.0
    ibltyp: int = 0        # 2: TKE carried through the step
    ntr: int = 0           # chemical tracers (ichem == 1 when > 0)
    tkemin: float = 0.0    # the PBL scheme's TKE floor
    crm: int = 0           # i_crm = 1: doubly periodic, jx x iy cross points

    @property
    def npoints(self) -> int:
        """Grid points of the metric: the cross points times kz, (jx-1)(iy-1)kz
        with lateral boundaries, jx*iy*kz on a doubly periodic domain."""
        if self.crm:
            return self.jx * self.iy * self.kz
        return (self.jx - 1) * (self.iy - 1) * self.kz


STATE_3D = ("u", "v", "w", "pai", "tetav", "t", "tvirt", "ux", "vx", "p", "rho", "qsat")
TEND_3D = ("tten", "uten", "vten")
STATIC_3D = ("fmz", "fmzf", "rfmzu", "rfmzv", "bdywtu", "bdywtv", "bdywtw", "zeta", "dz")
STATIC_2D = ("mx", "mu", "mv", "hx", "hy", "coru", "corv")


def empty_fields(g: Grid) -> dict:
    kz, iy, jx = g.kz, g.iy, g.jx
    f = {}
    for name in STATE_3D + TEND_3D + STATIC_3D:
        nk = kz + 1 if name in ("w", "fmzf") else kz
        f[name] = np.zeros((nk, iy, jx))
    f["qx"] = np.zeros((g.nqx, kz, iy, jx))
    f["qxten"] = np.zeros((g.nqx, kz, iy, jx))
    f["ps"] = np.zeros((iy, jx))
    for name in STATIC_2D:
        f[name] = np.zeros((iy, jx))
    f["rlat"] = np.zeros(iy)
    f["ffilt"] = np.zeros(kz)
    if g.ibltyp == 2:
        f["tke"] = np.zeros((kz + 1, iy, jx))
        f["tketen"] = np.zeros((kz + 1, iy, jx))
    if g.ntr > 0:
        f["trac"] = np.zeros((g.ntr, kz, iy, jx))
        f["chiten"] = np.zeros((g.ntr, kz, iy, jx))
    return f


def _hills(x, y, hills, period=None):
    """Sum of Gaussian hills; with period = (lx, ly) the distance to each
    centre is the shortest one on the torus."""
    h = np.zeros(np.broadcast(x, y).shape)
    for (xc, yc, h0, a) in hills:
        dxh, dyh = x - xc, y - yc
        if period is not None:
            dxh = dxh - period[0] * np.round(dxh / period[0])
            dyh = dyh - period[1] * np.round(dyh / period[1])
        r2 = dxh ** 2 + dyh ** 2
        h = h + h0 * np.exp(-math.log(2.0) * r2 / (a * a))
    return h


def _std_atm_t(z):
    return np.where(z <= 11000.0, 288.0 - 0.0065 * z, 216.5)


def _std_atm_p(z, ps_sea=101325.0):
    expo = EGRAV / (RGAS * 0.0065)
    t = _std_atm_t(z)
    p_low = ps_sea * (t / 288.0) ** expo
    p11 = ps_sea * (216.5 / 288.0) ** expo
    p_high = p11 * np.exp(-EGRAV * (z - 11000.0) / (RGAS * 216.5))
    return np.where(z <= 11000.0, p_low, p_high)


def _bdy_weights(g: Grid, nj: int, ni: int, nspgx: int, anudge: float) -> np.ndarray:
    """1 - hefc(ib,k) over the relaxation zone (Main/mod_bdycod.F90:4033-4047
    with ib from Main/mod_atm_interface.F90:384-512 and the exponential
    hefc(n,k) = exp(-(n-1)/anudge) of Main/mod_bdycod.F90:539-544).  nj, ni
    are the extents of the staggered grid the weights live on."""
    j = np.arange(1, g.jx + 1)[None, :]
    i = np.arange(1, g.iy + 1)[:, None]
    dist = np.minimum(np.minimum(j - 1, nj - j), np.minimum(i - 1, ni - i))
    ib = dist + 1
    hefc = np.zeros(nspgx + 1)
    hefc[1] = 1.0
    for n in range(2, nspgx):
        hefc[n] = math.exp(-(n - 1) / anudge)
    hefc[nspgx] = 0.0
    w = np.where((ib >= 1) & (ib <= nspgx), 1.0 - hefc[np.clip(ib, 0, nspgx)], 1.0)
    return np.broadcast_to(w, (g.kz, g.iy, g.jx)).copy()


def build_domain(g: Grid, hills, *, stress: bool = False, nspgx: int | None = None,
                 anudge: float = 2.0) -> dict:
    """Static metrics of one domain (Main/mod_params.F90:3316-3395,
    Main/mod_moloch.F90:260-262, Main/mod_init.F90:1008-1025).  `hills` is a
    list of (x_centre, y_centre, h0, half_width) in metres.  `stress` turns on
    non-uniform map factors and latitude-dependent Coriolis so that every
    metric array carries information."""
    f = empty_fields(g)
    jx, iy, kz, dx = g.jx, g.iy, g.kz, g.dx
    jj = np.arange(1, jx + 1, dtype=np.float64)[None, :]
    ii = np.arange(1, iy + 1, dtype=np.float64)[:, None]
    # cross point (j,i) centre; u point (j,i) on its west face; v on its south
    xc, yc = (jj - 0.5) * dx, (ii - 0.5) * dx
    xu, yu = (jj - 1.0) * dx, (ii - 0.5) * dx
    xv, yv = (jj - 0.5) * dx, (ii - 1.0) * dx
    lx, ly = jx * dx, iy * dx
    period = (lx, ly) if g.crm else None
    ht = _hills(xc, yc, hills, period) * EGRAV          # geopotential (m2/s2)
    htu = _hills(xu, yu, hills, period) * EGRAV
    htv = _hills(xv, yv, hills, period) * EGRAV
    if stress:
        def mapf(x, y):
            if g.crm:   # periodic in both directions
                return 1.0 + 0.03 * np.sin(2 * math.pi * x / lx) * np.cos(2 * math.pi * y / ly)
            return 1.0 + 0.03 * np.sin(2 * math.pi * x / lx) * np.cos(math.pi * y / ly)
        mx, mu, mv = mapf(xc, yc), mapf(xu, yu), mapf(xv, yv)
        latu = 30.0 + 25.0 * yu / ly + 2.0 * np.sin(2 * math.pi * xu / lx)
        latv = 30.0 + 25.0 * yv / ly + 2.0 * np.sin(2 * math.pi * xv / lx)
    else:
        mx = np.ones((iy, jx)); mu = np.ones((iy, jx)); mv = np.ones((iy, jx))
        latu = np.full((iy, jx), 45.0); latv = np.full((iy, jx), 45.0)
    f["mx"][:], f["mu"][:], f["mv"][:] = mx, mu, mv
    f["coru"][:] = EOMEG2 * np.sin(latu * DEGRAD)
    f["corv"][:] = EOMEG2 * np.sin(latv * DEGRAD)
    f["rlat"][:] = 30.0 + 25.0 * (np.arange(1, iy + 1) - 1.0) / iy
    rdx = 1.0 / dx
    # hx on (jde, ice), hy on (jce, ide): mod_params.F90:3324-3354
    hx = np.zeros((iy, jx)); hy = np.zeros((iy, jx))
    hx[:, 1:] = rdx * REGRAV * mu[:, 1:] * (ht[:, 1:] - ht[:, :-1])
    hy[1:, :] = rdx * REGRAV * mv[1:, :] * (ht[1:, :] - ht[:-1, :])
    if g.crm:   # the west / south neighbour of the first row is the last one
        hx[:, 0] = rdx * REGRAV * mu[:, 0] * (ht[:, 0] - ht[:, -1])
        hy[0, :] = rdx * REGRAV * mv[0, :] * (ht[0, :] - ht[-1, :])
    else:
        hx[:, 0] = 2.0 * rdx * REGRAV * mu[:, 0] * (ht[:, 0] - htu[:, 0])
        hy[0, :] = 2.0 * rdx * REGRAV * mv[0, :] * (ht[0, :] - htv[0, :])
    f["hx"][:], f["hy"][:] = hx, hy
    zf = model_zitaf(kz, g.mo_ztop)
    zh = model_zitah(kz, g.mo_ztop)
    args = (g.mo_ztop, g.mo_h, g.mo_a0)
    f["zeta"][:] = md_zeta(zh[:, None, None], ht[None], *args)
    f["fmz"][:] = md_fmz(zh[:, None, None], ht[None], *args)
    f["rfmzu"][:] = 1.0 / md_fmz(zh[:, None, None], htu[None], *args)
    f["rfmzv"][:] = 1.0 / md_fmz(zh[:, None, None], htv[None], *args)
    f["fmzf"][:] = md_fmz(zf[:, None, None], ht[None], *args)
    zetaf = md_zeta(zf[:, None, None], ht[None], *args)
    f["dz"][:] = zetaf[:-1] - zetaf[1:]
    # top sponge: mod_init.F90:1008-1025 (mo_zfilt_fac = 0.8)
    own = (slice(None), slice(None), slice(None)) if g.crm else \
        (slice(None), slice(0, iy - 1), slice(0, jx - 1))
    gmeanz = f["zeta"][own].mean(axis=(1, 2))
    zzi = (gmeanz - 18000.0) / (g.mo_ztop - 18000.0)
    f["ffilt"][:] = np.where(gmeanz < 18000.0, 0.0, 0.8 * np.sin(0.5 * MATHPI * zzi) ** 2)
    if nspgx is None:
        ds_km = dx / 1000.0
        nspgx = max(min(max(int(6 * 50 / ds_km), 6), min(jx, iy) // 4), 6)
    if g.crm:   # no boundary ring: setup_bdywt gives 1 everywhere
        for name in ("bdywtu", "bdywtv", "bdywtw"):
            f[name][:] = 1.0
    else:
        f["bdywtu"][:] = _bdy_weights(g, jx, iy - 1, nspgx, anudge)
        f["bdywtv"][:] = _bdy_weights(g, jx - 1, iy, nspgx, anudge)
        f["bdywtw"][:] = _bdy_weights(g, jx - 1, iy - 1, nspgx, anudge)
    f["_ht"] = ht
    # the surface geopotential the terrain metrics were formed from
    f["ht"], f["htu"], f["htv"] = ht, htu, htv
    f["_xyc"] = (xc, yc)
    return f


def paicompute(g: Grid, f: dict) -> None:
    """Hydrostatic Exner function: Main/mod_bdycod.F90:3762-3796."""
    kz = g.kz
    z, t, q = f["zeta"], f["t"], f["qx"][0]
    mo_dzita = model_zitaf(kz, g.mo_ztop)[kz - 1]
    zdelta = z[kz - 1] * EGRAV
    tv1 = t[kz - 1] * (1.0 + EP1 * q[kz - 1])
    tv2 = t[kz - 2] * (1.0 + EP1 * q[kz - 2])
    lrt = (tv2 - tv1) / (z[kz - 2] - z[kz - 1])
    lrt = np.where(lrt > GOVCP, GOVCP, np.where(lrt < -0.005, 0.5 * lrt - 0.5 * LRATE, lrt))
    tv = tv1 - 0.5 * z[kz - 1] * lrt
    zz = 1.0 / (RGAS * tv)
    p = f["ps"] * np.exp(-zdelta * zz)
    paikp1 = (p / P00) ** ROVCP
    pai = f["pai"]
    pai[kz - 1] = paikp1
    for k in range(kz - 1, 0, -1):
        tv1 = t[k - 1] * (1.0 + EP1 * q[k - 1])
        tv2 = t[k] * (1.0 + EP1 * q[k])
        zb = 2.0 * EGRAV * mo_dzita / (f["fmzf"][k] * CPD) + tv1 - tv2
        zdelta = np.sqrt(zb ** 2 + 4.0 * tv2 * tv1)
        paikp1 = -paikp1 / (2.0 * tv2) * (zb - zdelta)
        pai[k - 1] = paikp1


def _tv_factor(g: Grid, qx):
    fac = 1.0 + EP1 * qx[0]
    if g.ipptls > 0:
        fac = fac - qx[1]
    if g.ipptls > 1:
        fac = fac - qx[2] - qx[3] - qx[4]
    return fac


def initial_state(g: Grid, f: dict, *, wind: str = "uniform", seed_u: int = 1,
                  seed_v: int = 2, u0: float = 10.0, sigma_u: float = 0.5,
                  sigma_v: float = 0.0, bubble: tuple | None = None,
                  dtemp: tuple | None = None, dwind: tuple | None = None,
                  resting: bool = False, qv_floor: float = 0.0) -> None:
    """US-standard-like initial state on the domain f (in place).

    wind="uniform": u = u0 + N(0, sigma_u^2); v = N(0, sigma_v^2).
    wind="jet":     u = 20 m/s exp(-(ln(p/250 hPa)/0.6)^2) + N(0,1);
                    v = N(0,1)  (seeds seed_u / seed_v).
    bubble = (x, y, z, radius, amplitude) adds a cos^2 warm bubble to T.
    dtemp = (seed, sigma) adds N(0, sigma^2) to T (ensemble members).
    dwind = (seed, sigma) adds N(0, sigma^2) to u (stream seed) and to v
    (stream seed + 1) after the wind profile (ensemble members).
    resting = True gives u = v = w = 0.
    qv_floor raises qv to at least that value before Exner is built (1e-8 is
    the floor status_update applies, Main/mpplib/mod_runparams.F90:184).
    """
    kz, iy, jx = g.kz, g.iy, g.jx
    ht = f["_ht"] * REGRAV
    zabs = f["zeta"] + ht[None]
    t = _std_atm_t(zabs)
    if bubble is not None:
        bx, by, bz, rad, amp = bubble
        xc, yc = f["_xyc"]
        ddx, ddy = xc - bx, yc - by
        if g.crm:   # shortest distance on the torus
            lx, ly = jx * g.dx, iy * g.dx
            ddx, ddy = ddx - lx * np.round(ddx / lx), ddy - ly * np.round(ddy / ly)
        r = np.sqrt(ddx[None] ** 2 + ddy[None] ** 2 + (zabs - bz) ** 2)
        t = t + np.where(r < rad, amp * np.cos(MATHPI * r / (2.0 * rad)) ** 2, 0.0)
    if dtemp is not None:
        t = t + normal(dtemp[0], t.shape, dtemp[1])
    f["t"][:] = t
    f["ps"][:] = 101325.0 * (_std_atm_t(ht) / 288.0) ** (EGRAV / (RGAS * 0.0065))
    papprox = _std_atm_p(zabs)
    f["qx"][:] = 0.0
    f["qx"][0] = np.maximum(np.where(zabs < 8000.0, 0.8 * pfwsat(t, papprox), 0.0), qv_floor)
    paicompute(g, f)
    pai = f["pai"]
    f["tvirt"][:] = f["t"] * _tv_factor(g, f["qx"])
    f["tetav"][:] = f["tvirt"] / pai
    f["p"][:] = pai ** CPOVR * P00
    f["rho"][:] = f["p"] / (RGAS * f["t"])
    f["qsat"][:] = pfwsat(f["t"], f["p"])
    f["w"][:] = 0.0
    if resting:
        f["u"][:] = 0.0
        f["v"][:] = 0.0
    elif wind == "uniform":
        f["u"][:] = u0 + normal(seed_u, f["u"].shape, sigma_u) if sigma_u > 0 else u0
        f["v"][:] = normal(seed_v, f["v"].shape, sigma_v) if sigma_v > 0 else 0.0
    elif wind == "jet":
        jet = 20.0 * np.exp(-(np.log(papprox / 25000.0) / 0.6) ** 2)
        f["u"][:] = jet + normal(seed_u, jet.shape, 1.0)
        f["v"][:] = normal(seed_v, jet.shape, 1.0)
    else:
        raise ValueError(f"unknown wind profile {wind!r}")
    if dwind is not None and not resting:
        f["u"][:] = f["u"] + normal(dwind[0], f["u"].shape, dwind[1])
        f["v"][:] = f["v"] + normal(dwind[0] + 1, f["v"].shape, dwind[1])
    # cross-point winds (Main/mod_moloch.F90:1524-1569 stencil)
    u, v = f["u"], f["v"]
    ux, vx = f["ux"], f["vx"]
    if g.crm:
        ux[:] = 0.5625 * (np.roll(u, -1, 2) + u) - 0.0625 * (np.roll(u, -2, 2) + np.roll(u, 1, 2))
        vx[:] = 0.5625 * (np.roll(v, -1, 1) + v) - 0.0625 * (np.roll(v, -2, 1) + np.roll(v, 1, 1))
        tke_tracers(g, f, zabs)
        return
    ux[:, :, 1:jx - 2] = 0.5625 * (u[:, :, 2:jx - 1] + u[:, :, 1:jx - 2]) - \
        0.0625 * (u[:, :, 3:jx] + u[:, :, 0:jx - 3])
    ux[:, :, 0] = 0.5 * (u[:, :, 0] + u[:, :, 1])
    ux[:, :, jx - 2] = 0.5 * (u[:, :, jx - 1] + u[:, :, jx - 2])
    vx[:, 1:iy - 2, :] = 0.5625 * (v[:, 2:iy - 1, :] + v[:, 1:iy - 2, :]) - \
        0.0625 * (v[:, 3:iy, :] + v[:, 0:iy - 3, :])
    vx[:, 0, :] = 0.5 * (v[:, 0, :] + v[:, 1, :])
    vx[:, iy - 2, :] = 0.5 * (v[:, iy - 1, :] + v[:, iy - 2, :])
    tke_tracers(g, f, zabs)


def tke_tracers(g: Grid, f: dict, zabs: np.ndarray, seed: int = 500) -> None:
    """TKE (ibltyp 2) decaying from 1 m2/s2 near the ground over 1 km, on the
    full levels, and ntr tracers: plumes of height-dependent mixing ratio
    (1e-9 .. 1e-7) with 20% seeded noise, tracer n from seed + n."""
    if g.ibltyp == 2:
        zf = np.concatenate([zabs[:1] + (zabs[:1] - zabs[1:2]) * 0.5,
                             0.5 * (zabs[1:] + zabs[:-1]),
                             zabs[-1:] - (zabs[-2:-1] - zabs[-1:]) * 0.5])
        base = 0.05 + np.exp(-np.maximum(zf - zf[-1:], 0.0) / 1000.0)
        f["tke"][:] = base * (1.0 + 0.2 * normal(seed, base.shape))
        f["tke"][:] = np.maximum(f["tke"], 1e-4)
    for n in range(g.ntr):
        h = 1000.0 * (n + 1)
        prof = 1e-7 * np.exp(-np.maximum(zabs - zabs[-1:], 0.0) / h) + 1e-9
        f["trac"][n] = np.maximum(prof * (1.0 + 0.2 * normal(seed + 1 + n, prof.shape)), 0.0)


def random_hills(seed: int, n: int, lx: float, ly: float,
                 hrange=(200.0, 2000.0), wrange=(20e3, 80e3)):
    """n Gaussian hills with centres uniform over the middle 60% of the domain."""
    u = uniform01(seed, 4 * n).reshape(n, 4)
    hills = []
    for m in range(n):
        xc = lx * (0.2 + 0.6 * u[m, 0])
        yc = ly * (0.2 + 0.6 * u[m, 1])
        h0 = hrange[0] + (hrange[1] - hrange[0]) * u[m, 2]
        a = wrange[0] + (wrange[1] - wrange[0]) * u[m, 3]
        hills.append((xc, yc, h0, a))
    return hills


# ------------------------------------------------- lateral boundaries ----
def boundary_area(g: Grid, ldotx: bool, ldoty: bool, nspgx: int, nspgd: int) -> np.ndarray:
    """Ring index ibnd(j,i) of a boundary area (Main/mod_atm_interface.F90:
    384-512, single domain, no band/CRM): > 0 inside the relaxation zone,
    -1 elsewhere.  Returned as an (iy, jx) array (0-based storage)."""
    jx, iy = g.jx, g.iy
    jcx, nspx = (0, nspgd) if ldotx else (1, nspgx)
    icy, nspy = (0, nspgd) if ldoty else (1, nspgx)
    igbb1, igbb2 = 2, nspy - 1
    jgbl1, jgbl2 = 2, nspx - 1
    igbt1, igbt2 = iy - icy - nspy + 2, iy - 1 - icy
    jgbr1, jgbr2 = jx - jcx - nspx + 2, jx - 1 - jcx
    ib = np.full((iy, jx), -1, dtype=np.int64)
    south = np.zeros((iy, jx), bool)
    north = np.zeros((iy, jx), bool)
    for i in range(1, iy + 1):
        for j in range(1, jx + 1):
            if igbb1 <= i <= igbb2 and jgbl1 <= j <= jgbr2:
                if not ((j <= jgbl2 and i >= j) or (j >= jgbr1 and i >= (jgbr2 - j + 2))):
                    ib[i - 1, j - 1] = i - igbb1 + 2
                    south[i - 1, j - 1] = True
    for i in range(1, iy + 1):
        for j in range(1, jx + 1):
            if igbt1 <= i <= igbt2 and jgbl1 <= j <= jgbr2:
                if not ((j <= jgbl2 and j <= (igbt2 - i + 2)) or
                        (j >= jgbr1 and (igbt2 - i) >= (jgbr2 - j))):
                    ib[i - 1, j - 1] = igbt2 - i + 2
                    north[i - 1, j - 1] = True
    for i in range(1, iy + 1):
        for j in range(1, jx + 1):
            if north[i - 1, j - 1] or south[i - 1, j - 1] or i < igbb1 or i > igbt2:
                continue
            if jgbl1 <= j <= jgbl2:
                ib[i - 1, j - 1] = j - jgbl1 + 2
    for i in range(1, iy + 1):
        for j in range(1, jx + 1):
            if north[i - 1, j - 1] or south[i - 1, j - 1] or i < igbb1 or i > igbt2:
                continue
            if jgbr1 <= j <= jgbr2:
                ib[i - 1, j - 1] = jgbr2 - j + 2
    return ib


def hefc_table(g: Grid, nspgx: int, anudge: float) -> np.ndarray:
    """hefc(n,k), n = 1..nspgx (row n of the result; row 0 unused): 1 on the
    outer ring, 0 on the innermost, exp(-(n-1)/anudge) between
    (Main/mod_bdycod.F90:522-545, exponential nudging)."""
    h = np.zeros((nspgx + 1, g.kz))
    h[1, :] = 1.0
    for n in range(2, nspgx):
        h[n, :] = math.exp(-float(n - 1) / anudge)
    h[nspgx, :] = 0.0
    return h


def lbc_weights(g: Grid, ib: np.ndarray, hefc: np.ndarray, j1: int, j2: int,
                i1: int, i2: int):
    """(relaxation weight, bdywt mask) over j1..j2 x i1..i2 of one area:
    weight = hefc(ib,k) where ib > 0 and -1 elsewhere (morelax skips it);
    mask = 1 - hefc(ib,k) or 1 (setup_bdywt, Main/mod_bdycod.F90:4033-4047);
    both 0 / -1 outside the loop range, as their zero-filled arrays."""
    xf = np.full((g.kz, g.iy, g.jx), -1.0)
    mask = np.zeros((g.kz, g.iy, g.jx))
    sub = ib[i1 - 1:i2, j1 - 1:j2]
    inside = sub > 0
    hk = hefc[np.clip(sub, 0, hefc.shape[0] - 1)]          # (ni, nj, kz)
    hk = np.moveaxis(hk, -1, 0)
    xf[:, i1 - 1:i2, j1 - 1:j2] = np.where(inside[None], hk, -1.0)
    mask[:, i1 - 1:i2, j1 - 1:j2] = np.where(inside[None], 1.0 - hk, 1.0)
    return xf, mask


def lbc_inputs(g: Grid, f: dict, *, nspgx: int = 6, nspgd: int = 7, anudge: float = 2.0,
               seed: int = 77, amp: float = 0.5, top_nudge: bool = False,
               present_qc: bool = True, present_qi: bool = True) -> dict:
    """Boundary data of a real-case run on the domain f: b0 = the current
    state, b1 = b0 with N(0, amp^2) perturbations (winds in m/s, T in K,
    scaled for pai and the species), the Davies weights of the three
    boundary areas, and a top-nudging profile.  Also rewrites f's bdywt
    masks from the same rings, as setup_bdywt does."""
    kz, iy, jx = g.kz, g.iy, g.jx
    hefc = hefc_table(g, nspgx, anudge)
    ib_cr = boundary_area(g, False, False, nspgx, nspgd)
    ib_ud = boundary_area(g, True, False, nspgx, nspgd)
    ib_vd = boundary_area(g, False, True, nspgx, nspgd)
    xfu, mu_ = lbc_weights(g, ib_ud, hefc, 2, jx - 1, 2, iy - 2)
    xfv, mv_ = lbc_weights(g, ib_vd, hefc, 2, jx - 2, 2, iy - 1)
    xfc, mc_ = lbc_weights(g, ib_cr, hefc, 2, jx - 2, 2, iy - 2)
    f["bdywtu"][:], f["bdywtv"][:], f["bdywtw"][:] = mu_, mv_, mc_
    b = {"present_qc": present_qc, "present_qi": present_qi,
         "top_nudge": top_nudge, "xfu": xfu, "xfv": xfv, "xfc": xfc}
    src = {"u": f["u"], "v": f["v"], "t": f["t"], "pai": f["pai"], "qv": f["qx"][0],
           "qc": f["qx"][1] if g.nqx > 1 else np.zeros_like(f["t"]),
           "qi": f["qx"][2] if g.nqx > 2 else np.zeros_like(f["t"])}
    scale = {"u": amp, "v": amp, "t": amp, "pai": amp * 1e-4, "qv": amp * 1e-4,
             "qc": amp * 1e-5, "qi": amp * 1e-5}
    for n, (name, a) in enumerate(src.items()):
        b[name + "0"] = np.ascontiguousarray(a, dtype=np.float64).copy()
        pert = normal(seed + n, a.shape, scale[name])
        b1 = a + pert
        if name in ("qv", "qc", "qi"):
            b1 = np.abs(b1)
        b[name + "1"] = np.ascontiguousarray(b1)
    b["ps0"] = f["ps"].copy()
    b["ps1"] = f["ps"] + normal(seed + 20, f["ps"].shape, 50.0 * amp)
    own = (slice(None), slice(None), slice(None)) if g.crm else \
        (slice(None), slice(0, iy - 1), slice(0, jx - 1))
    gmeanz = f["zeta"][own].mean(axis=(1, 2))
    zztop = 18000.0
    nztop = int(np.sum(gmeanz > zztop))
    tn = np.zeros(kz)
    for k in range(kz):
        if k < nztop:
            tn[k] = math.sin(0.5 * MATHPI * (gmeanz[k] - zztop) / (g.mo_ztop - zztop)) ** 2
    b["tnudge"] = tn
    b["nztop"] = nztop
    return b


def spectral_inputs(g: Grid, f: dict, *, dtrad: float = 1800.0, dtbdys: float = 21600.0,
                    km: int | None = None, lm: int | None = None) -> dict:
    """Spectral-nudging inputs of lowpass_init (Main/mod_bdycod.F90:3844-3896):
    the sine bases bvx(jx, 2km), bvy(iy, 2lm) (stored as (2km, jx) and
    (2lm, iy) arrays, i.e. Fortran order) and cnudge(kz).  km and lm default
    to the cut-off wave numbers of 1500 km / 750 km."""
    njcross, nicross = g.jx - 1, g.iy - 1
    ds = g.dx / 1000.0
    if km is None:
        km = max(int(round(njcross * ds / 1500.0)), 1)
    if lm is None:
        lm = max(int(round(nicross * ds / 750.0)), 1)
    dxs = MATHPI / float(njcross - 1)
    dys = MATHPI / float(nicross - 1)
    bvx = np.zeros((2 * km, g.jx))
    bvy = np.zeros((2 * lm, g.iy))
    for k in range(1, 2 * km + 1):
        px = math.exp(-(float(k) / float(km)) ** 2)
        for j in range(1, g.jx + 1):
            bvx[k - 1, j - 1] = math.sqrt(2.0 / float(g.jx - 1) * px) * math.sin(k * (j - 2) * dxs)
    for l in range(1, 2 * lm + 1):
        py = math.exp(-(float(l) / float(lm)) ** 2)
        for i in range(1, g.iy + 1):
            bvy[l - 1, i - 1] = math.sqrt(2.0 / float(g.iy - 1) * py) * math.sin(l * (i - 2) * dys)
    gmeanz = f["zeta"][:, : g.iy - 1, : g.jx - 1].mean(axis=(1, 2))
    cn0 = dtrad / dtbdys
    cnudge = cn0 * np.minimum((gmeanz / g.mo_h) ** 2, 1.0)
    return {"km": km, "lm": lm, "dtrad": dtrad, "tspectral": 0.0, "bvx": bvx, "bvy": bvy,
            "cnudge": np.ascontiguousarray(cnudge)}


def slice_inputs(g: Grid, f: dict, *, calday: float = 200.5, dayspy: float = 365.0,
                 rhmin: float = 0.01, rhmax: float = 1.01, icldmstrat: int = 1,
                 irceideal: int = 0, ibltyp: int = 1) -> dict:
    """Inputs of mkslice beyond the state (Main/mod_slice.F90:112-382): the
    full-level heights mo_atm%zetaf (kz+1, iy, jx) of the domain's terrain
    (Main/mod_moloch.F90:260-262), a latitude field spanning both
    hemispheres so that both tropopause fits are used, and the namelist
    limits and switches (RH limits: Main/mod_params.F90:381-382)."""
    zf = model_zitaf(g.kz, g.mo_ztop)
    zetaf = md_zeta(zf[:, None, None], f["_ht"][None], g.mo_ztop, g.mo_h, g.mo_a0)
    jj = np.arange(1, g.jx + 1, dtype=np.float64)[None, :]
    ii = np.arange(1, g.iy + 1, dtype=np.float64)[:, None]
    xlat = -50.0 + 100.0 * (ii - 1.0) / max(g.iy - 1, 1) + 3.0 * np.sin(2 * MATHPI * jj / g.jx)
    return {"zetaf": np.ascontiguousarray(zetaf), "xlat": np.ascontiguousarray(xlat),
            "calday": calday, "dayspy": dayspy, "rhmin": rhmin, "rhmax": rhmax,
            "icldmstrat": icldmstrat, "irceideal": irceideal, "ibltyp": ibltyp}


SLICE_OUTPUTS = ("pf", "th3d", "rhb3d", "wpx3d", "rhox2d", "tp2d", "th700", "ptrop",
                 "ktrop", "kmxpbl")


# ------------------------------------------------ horizontal diffusion ----
@dataclasses.dataclass
class DiffGrid:
    """Descriptor of the MM-core horizontal-diffusion operator
    (Main/mod_diffusion.F90; namelist defaults Main/mod_params.F90:296,
    :797-804)."""
    jx: int
    iy: int
    kz: int
    dx: float = 12000.0
    dtsec: float = 30.0
    idiffu: int = 1
    idynamic: int = 2
    diffu_hgtf: int = 0
    nspec: int = 3
    ckh: float = 1.0
    adyndif: float = 1.0


def diffusion_inputs(dg: DiffGrid, seed: int = 7) -> dict:
    """Synthetic inputs of the operator: terrain geopotential ht, map factor
    msfd, psb and psdotb, dot-point winds ud/vd (a jet with shear and
    noise), w on kz+1 levels, a kz+1-level multi-species field f and random
    tendencies."""
    jx, iy, kz, dx = dg.jx, dg.iy, dg.kz, dg.dx
    jj = np.arange(1, jx + 1, dtype=np.float64)[None, :]
    ii = np.arange(1, iy + 1, dtype=np.float64)[:, None]
    lx, ly = jx * dx, iy * dx
    x, y = (jj - 1.0) * dx, (ii - 1.0) * dx
    ht = _hills(x, y, random_hills(seed, 4, lx, ly, wrange=(2 * dx, 6 * dx))) * EGRAV
    msfd = 1.0 + 0.03 * np.sin(2 * MATHPI * x / lx) * np.cos(MATHPI * y / ly)
    pc = 95.0 + 3.0 * np.cos(2 * MATHPI * x / lx) * np.sin(MATHPI * y / ly) - 1e-3 * ht / EGRAV
    pd = 95.0 + 3.0 * np.cos(2 * MATHPI * (x - 0.5 * dx) / lx) * \
        np.sin(MATHPI * (y - 0.5 * dx) / ly)
    lev = np.linspace(0.3, 1.0, kz)[:, None, None]
    ud = 15.0 * lev * np.exp(-((y - 0.5 * ly) / (0.25 * ly)) ** 2)[None] + \
        normal(seed + 1, (kz, iy, jx), 0.8)
    vd = 4.0 * np.sin(2 * MATHPI * x / lx)[None] * lev + normal(seed + 2, (kz, iy, jx), 0.8)
    wx = normal(seed + 3, (kz + 1, iy, jx), 0.2)
    f = 280.0 + 5.0 * np.sin(2 * MATHPI * x / lx)[None, None] + \
        normal(seed + 4, (dg.nspec, kz + 1, iy, jx), 1.0)
    out = {"ht": ht, "msfd": msfd, "pc": pc, "pd": pd, "ud": ud, "vd": vd, "wx": wx,
           "f": f, "ften": normal(seed + 5, (dg.nspec, kz + 1, iy, jx), 1e-3),
           "uten": normal(seed + 6, (kz, iy, jx), 1e-3),
           "vten": normal(seed + 7, (kz, iy, jx), 1e-3)}
    return {k: np.ascontiguousarray(v) for k, v in out.items()}


# ------------------------------------------------------ named configs ----
def uneven_domains(n: int = 24, seed: int = 5):
    """(jx, iy, kz) of the uneven batch: jx, iy from {96, 128, 192, 256, 320,
    384}, kz from {23, 30, 41}, drawn with SplitMix64 from `seed`."""
    sizes, levels = (96, 128, 192, 256, 320, 384), (23, 30, 41)
    u = uniform01(seed, 3 * n).reshape(n, 3)
    return [(sizes[int(a * 6) % 6], sizes[int(b * 6) % 6], levels[int(c * 3) % 3])
            for a, b, c in u]


def uneven_grid(member: int) -> Grid:
    """Grid of domain `member` of the uneven batch: dx spans 1500 km, dt is
    2.5 dx[km] s."""
    jx, iy, kz = uneven_domains()[member]
    dx = 1.5e6 / (jx - 1)
    return Grid(jx, iy, kz, dx, 2.5 * dx / 1000.0, nqx=5, ipptls=2)


def config(name: str, **over) -> tuple[Grid, dict, int]:
    """(grid, fields, nsteps) of a named workload.

    cfg1  64x64x23,  dx 12 km, dt 30 s, 10 steps:  one 1 km hill, u = 10 + N(0, 0.5^2)
    cfg2  256x256x41, dx 5 km, dt 15 s, 100 steps: 8 random hills, jet, 5 species
    cfg3  512x512x41, dx 3 km, dt 10 s, 100 steps: cfg2 recipe + 3 K warm bubble
    cfg4  128x128x30, dx 12 km, dt 30 s, 200 steps: ensemble member `member`
          (cfg2 terrain/wind recipe; T += N(0, 0.1^2) seed 1000+m,
          u, v += N(0, 0.2^2) seeds 2000+2m, 2001+2m)
    cfg5  uneven domain `member` of 24 (see uneven_domains())
    small 24x20x10 parity case with every metric non-trivial (stress=True)
    crm   100x50x30 doubly periodic (Testing/isc24_small.in), dx 2 km, dt 30 s,
          10 steps: hills and a warm bubble across the seams
    crm-small 24x20x10 doubly periodic with every metric non-trivial
    isc24 500x100x60 doubly periodic (Testing/isc24.in), dx 2 km, dt 30 s,
          100 steps: hills and a warm bubble across the seams
    """
    member = over.pop("member", 0)
    stress = over.pop("stress", None)
    if name == "cfg1":
        g = Grid(64, 64, 23, 12000.0, 30.0, nqx=2, ipptls=0)
        lx, ly = (g.jx - 1) * g.dx, (g.iy - 1) * g.dx
        f = build_domain(g, [(lx / 2, ly / 2, 1000.0, 80e3)], stress=bool(stress))
        initial_state(g, f, wind="uniform", seed_u=1, u0=10.0, sigma_u=0.5)
        return g, f, 10
    if name in ("cfg2", "cfg3", "cfg4"):
        if name == "cfg2":
            g = Grid(256, 256, 41, 5000.0, 15.0, nqx=5, ipptls=2)
            steps = 100
        elif name == "cfg3":
            g = Grid(512, 512, 41, 3000.0, 10.0, nqx=5, ipptls=2)
            steps = 100
        else:
            g = Grid(128, 128, 30, 12000.0, 30.0, nqx=5, ipptls=2)
            steps = 200
        for k, v in over.items():
            setattr(g, k, v)
        lx, ly = (g.jx - 1) * g.dx, (g.iy - 1) * g.dx
        f = build_domain(g, random_hills(0, 8, lx, ly), stress=bool(stress))
        bubble = (lx / 2, ly / 2, 1500.0, 10e3, 3.0) if name == "cfg3" else None
        if name == "cfg4":
            initial_state(g, f, wind="jet", seed_u=2, seed_v=3,
                          dtemp=(1000 + member, 0.1), dwind=(2000 + 2 * member, 0.2))
        else:
            initial_state(g, f, wind="jet", seed_u=2, seed_v=3, bubble=bubble)
        return g, f, steps
    if name == "cfg5":
        g = uneven_grid(member)
        for k, v in over.items():
            setattr(g, k, v)
        lx, ly = (g.jx - 1) * g.dx, (g.iy - 1) * g.dx
        f = build_domain(g, random_hills(100 + member, 8, lx, ly), stress=bool(stress))
        initial_state(g, f, wind="jet", seed_u=10000 + 2 * member, seed_v=10001 + 2 * member,
                      dtemp=(20000 + member, 0.1))
        return g, f, 100
    if name in ("crm", "crm-small", "isc24"):
        if name == "isc24":
            # the doubly periodic domain of Testing/isc24.in and ideal.in:
            # 500 x 100 x 60, ds 2 km, dt 30 s, ipptls 2
            g = Grid(500, 100, 60, 2000.0, 30.0, nqx=5, ipptls=2, crm=1)
            hills = [(0.0, 0.5 * 100 * 2000.0, 600.0, 12e3),
                     (0.6 * 500 * 2000.0, 0.0, 800.0, 20e3)]
            bubble = (0.0, 0.0, 1500.0, 10e3, 3.0)
            stress_default, steps = False, 100
        elif name == "crm":
            # the doubly periodic domain of Testing/isc24_small.in: 100 x 50
            # x 30, ds 2 km, dt 30 s, ipptls 2
            g = Grid(100, 50, 30, 2000.0, 30.0, nqx=5, ipptls=2, crm=1)
            hills = [(0.0, 0.5 * 50 * 2000.0, 600.0, 12e3),
                     (0.7 * 100 * 2000.0, 0.0, 400.0, 8e3)]
            bubble = (0.0, 0.0, 1500.0, 10e3, 3.0)
            stress_default, steps = False, 10
        else:
            g = Grid(24, 20, 10, 10000.0, 30.0, nqx=5, ipptls=2, nadv=1, nsound=2, crm=1)
            hills = [(0.0, 0.55 * 20 * 10000.0, 1200.0, 50e3)]
            bubble = None
            stress_default, steps = True, 3
        for k, v in over.items():
            setattr(g, k, v)
        f = build_domain(g, hills, stress=stress_default if stress is None else stress)
        initial_state(g, f, wind="uniform", seed_u=21, seed_v=22, u0=8.0, sigma_u=1.0,
                      sigma_v=1.0, bubble=bubble)
        if g.nqx > 1:
            f["qx"][1] = np.abs(normal(23, f["qx"][1].shape, 1e-4))
        f["tvirt"][:] = f["t"] * _tv_factor(g, f["qx"])
        f["tetav"][:] = f["tvirt"] / f["pai"]
        return g, f, steps
    if name == "small":
        g = Grid(24, 20, 10, 10000.0, 30.0, nqx=5, ipptls=2, nadv=1, nsound=2)
        for k, v in over.items():
            setattr(g, k, v)
        lx, ly = (g.jx - 1) * g.dx, (g.iy - 1) * g.dx
        f = build_domain(g, [(lx * 0.45, ly * 0.55, 1200.0, 50e3)],
                         stress=True if stress is None else stress, nspgx=4)
        initial_state(g, f, wind="jet", seed_u=11, seed_v=12)
        if g.nqx > 1:
            f["qx"][1] = np.abs(normal(13, f["qx"][1].shape, 1e-4))
        if g.nqx > 3:
            f["qx"][3] = np.abs(normal(14, f["qx"][3].shape, 5e-5))
        f["tvirt"][:] = f["t"] * _tv_factor(g, f["qx"])
        f["tetav"][:] = f["tvirt"] / f["pai"]
        return g, f, 3
    raise ValueError(f"unknown config {name!r}")
