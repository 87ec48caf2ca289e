"""Host-side mirror of mod_moloch's public interface over libmoloch_hip.

The reference exposes three argument-less subroutines on module state
(Main/mod_moloch.F90:127): allocate_moloch, init_moloch and moloch.  The
class below keeps the same three phases -- `Moloch(grid)` allocates the
device state, `init_moloch()` derives init_moloch's metrics on the device,
`moloch(nsteps)` advances the step -- with state upload/download in between
(the reference's `mo_atm` arrays).  Every call goes through the C-ABI
declared in include/moloch_hip.h; there is no other compute path.
"""
from __future__ import annotations

import ctypes
import pathlib

import numpy as np

LIB_PATH = pathlib.Path(__file__).resolve().parent / "lib" / "libmoloch_hip.so"

MLC_FP64, MLC_MIXED = 0, 1
MLC_PARAM_TKEMIN, MLC_PARAM_MO_H = 1, 2
MLC_VAR_WAF_H, MLC_VAR_WAF_ROWS, MLC_VAR_WAF_GROUP, MLC_VAR_LEVEL_CHUNK = 0, 1, 2, 3
MLC_VAR_COLUMN, MLC_VAR_WAFV_GROUP, MLC_VAR_XCD_MAP, MLC_VAR_WAF_TILE = 4, 5, 6, 7
MLC_VAR_WAF_HGROUP, MLC_VAR_COLUMN_J, MLC_VAR_MIXED_MERGE, MLC_VAR_TERRAIN = 8, 9, 10, 11
MLC_VAR_MOM_CHUNK, MLC_VAR_STAGGER_CHUNK = 12, 13
VARIANT_NAMES = {MLC_VAR_WAF_H: "waf_h", MLC_VAR_WAF_ROWS: "waf_rows",
                 MLC_VAR_WAF_GROUP: "waf_group", MLC_VAR_LEVEL_CHUNK: "level_chunk",
                 MLC_VAR_COLUMN: "column", MLC_VAR_WAFV_GROUP: "wafv_group",
                 MLC_VAR_XCD_MAP: "xcd_map", MLC_VAR_WAF_TILE: "waf_tile",
                 MLC_VAR_WAF_HGROUP: "waf_hgroup", MLC_VAR_COLUMN_J: "column_j",
                 MLC_VAR_MIXED_MERGE: "mixed_merge", MLC_VAR_TERRAIN: "terrain",
                 MLC_VAR_MOM_CHUNK: "mom_chunk", MLC_VAR_STAGGER_CHUNK: "stag_chunk"}

FIELD_IDS = {
    "u": 1, "v": 2, "w": 3, "pai": 4, "tetav": 5, "t": 6, "tvirt": 7, "qx": 8,
    "ux": 9, "vx": 10, "p": 11, "rho": 12, "qsat": 13, "ps": 14, "tten": 15,
    "uten": 16, "vten": 17, "qxten": 18,
    "tke": 30, "tketen": 31, "trac": 32, "chiten": 33,
    "fmz": 100, "fmzf": 101, "rfmzu": 102, "rfmzv": 103, "bdywtu": 104,
    "bdywtv": 105, "bdywtw": 106, "zeta": 107, "dz": 108, "mx": 109, "mu": 110,
    "mv": 111, "hx": 112, "hy": 113, "coru": 114, "corv": 115, "rlat": 116,
    "ffilt": 117,
    "ps0": 200, "ps1": 201, "u0": 202, "u1": 203, "v0": 204, "v1": 205, "t0": 206,
    "t1": 207, "pai0": 208, "pai1": 209, "qv0": 210, "qv1": 211, "qc0": 212, "qc1": 213,
    "qi0": 214, "qi1": 215, "xfu": 216, "xfv": 217, "xfc": 218, "tnudge": 219,
    "zetaf": 120, "xlat": 121, "ht": 122, "htu": 123, "htv": 124,
    "pf": 300, "th3d": 301, "rhb3d": 302, "wpx3d": 303, "rhox2d": 304, "tp2d": 305,
    "th700": 306, "ptrop": 307, "ktrop": 308, "kmxpbl": 309,
}
SLICE_FIELDS = ("pf", "th3d", "rhb3d", "wpx3d", "rhox2d", "tp2d", "th700", "ptrop",
                "ktrop", "kmxpbl")
LBC_FIELDS = ("ps0", "ps1", "u0", "u1", "v0", "v1", "t0", "t1", "pai0", "pai1", "qv0",
              "qv1", "qc0", "qc1", "qi0", "qi1", "xfu", "xfv", "xfc", "tnudge")
STATIC_FIELDS = ("fmz", "fmzf", "rfmzu", "rfmzv", "bdywtu", "bdywtv", "bdywtw",
                 "zeta", "dz", "mx", "mu", "mv", "hx", "hy", "coru", "corv",
                 "rlat", "ffilt")
SLICE_STATICS = ("zetaf", "xlat")
# surface geopotential: with mo_h, lets the kernels form the terrain metrics
TERRAIN_STATICS = ("ht", "htu", "htv")
STATE_FIELDS = ("u", "v", "w", "pai", "tetav", "t", "tvirt", "qx", "ux", "vx",
                "p", "rho", "qsat", "ps")
SAV_FIELDS = ("u", "v", "w", "t", "pai", "qx", "ps")


class MolochError(RuntimeError):
    pass


class LbcDesc(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in ("present_qc", "present_qi", "top_nudge", "nztop",
                                            "spectral_nudge", "km", "lm", "reserved")] + \
        [("dtrad", ctypes.c_double)]


class SliceDesc(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in ("rhmin", "rhmax", "calday", "dayspy")] + \
        [(n, ctypes.c_int) for n in ("icldmstrat", "irceideal", "ibltyp", "reserved")]


class DiffDesc(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in ("jx", "iy", "kz", "idiffu", "idynamic",
                                            "diffu_hgtf", "nspec", "reserved")] + \
        [(n, ctypes.c_double) for n in ("dx", "dtsec", "ckh", "adyndif")]


DIFF_FIELD_IDS = {"ht": 400, "msfd": 401, "pc": 402, "pd": 403, "ud": 404, "vd": 405,
                  "wx": 406, "uten": 407, "vten": 408, "f": 409, "ften": 410, "hgfact": 411,
                  "xkc": 412, "xkd": 413, "xkcf": 414}
MLC_DIFFU_X3D, MLC_DIFFU_X3DF, MLC_DIFFU_X4D = 0, 1, 2


class GridDesc(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in (
        "jx", "iy", "kz", "nqx", "ipptls", "nadv", "nsound", "do_divdamp",
        "do_divfilter", "lrotllr", "precision", "ibltyp", "ntr", "crm")] + \
        [("reserved", ctypes.c_int * 2)] + \
        [(n, ctypes.c_double) for n in ("dx", "dtsec", "mo_ztop", "mo_a0")]


_lib = None
_I4 = ctypes.c_int * 4


def load_library() -> ctypes.CDLL:
    """Load libmoloch_hip.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise MolochError(
            f"{LIB_PATH} is missing: build it with `python -m regcm_amd.build` "
            "(there is no CPU fallback)")
    lib = ctypes.CDLL(str(LIB_PATH))
    vp, dp, ip = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)
    sig = {
        "mlc_create": ([ctypes.POINTER(GridDesc), ctypes.POINTER(vp)], ctypes.c_int),
        "mlc_destroy": ([vp], ctypes.c_int),
        "mlc_set_static": ([vp, ctypes.c_int, dp, ip, ip], ctypes.c_int),
        "mlc_put_state": ([vp, ctypes.c_int, dp, ip, ip], ctypes.c_int),
        "mlc_get_state": ([vp, ctypes.c_int, dp, ip, ip], ctypes.c_int),
        "mlc_init": ([vp], ctypes.c_int),
        "mlc_step": ([vp, ctypes.c_int], ctypes.c_int),
        "mlc_reset_tendencies": ([vp], ctypes.c_int),
        "mlc_dynamical_core": ([vp], ctypes.c_int),
        "mlc_diagnostics": ([vp], ctypes.c_int),
        "mlc_status_update": ([vp, ctypes.c_double], ctypes.c_int),
        "mlc_set_param": ([vp, ctypes.c_int, ctypes.c_double], ctypes.c_int),
        "mlc_massck": ([vp, dp], ctypes.c_int),
        "mlc_massck_noalloc": ([vp, dp], ctypes.c_int),
        "mlc_check_finite": ([vp, dp, dp], ctypes.c_int),
        "mlc_synchronize": ([vp], ctypes.c_int),
        "mlc_profile": ([vp, ctypes.c_int], ctypes.c_int),
        "mlc_kernel_count": ([], ctypes.c_int),
        "mlc_kernel_name": ([ctypes.c_int], ctypes.c_char_p),
        "mlc_kernel_stats": ([vp, ctypes.c_int, ctypes.POINTER(ctypes.c_longlong), dp], ctypes.c_int),
        "mlc_npoints": ([vp], ctypes.c_longlong),
        "mlc_constant": ([ctypes.c_char_p], ctypes.c_double),
        "mlc_last_error": ([], ctypes.c_char_p),
        "mlc_set_variant": ([vp, ctypes.c_int, ctypes.c_int], ctypes.c_int),
        "mlc_lbc_configure": ([vp, ctypes.POINTER(LbcDesc)], ctypes.c_int),
        "mlc_lbc_spectral": ([vp, dp, dp, dp], ctypes.c_int),
        "mlc_init_restart": ([vp], ctypes.c_int),
        "mlc_boundary": ([vp, ctypes.c_double], ctypes.c_int),
        "mlc_step_lbc": ([vp, ctypes.c_int, dp], ctypes.c_int),
        "mlc_mkslice": ([vp, ctypes.POINTER(SliceDesc)], ctypes.c_int),
        "mlc_diff_create": ([ctypes.POINTER(DiffDesc), ctypes.POINTER(vp)], ctypes.c_int),
        "mlc_diff_destroy": ([vp], ctypes.c_int),
        "mlc_diff_put": ([vp, ctypes.c_int, dp, ip, ip], ctypes.c_int),
        "mlc_diff_get": ([vp, ctypes.c_int, dp, ip, ip], ctypes.c_int),
        "mlc_diff_initialize": ([vp], ctypes.c_int),
        "mlc_diff_calc_coeff": ([vp], ctypes.c_int),
        "mlc_diff_d": ([vp], ctypes.c_int),
        "mlc_diff_x": ([vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double],
                       ctypes.c_int),
        "mlc_diff_synchronize": ([vp], ctypes.c_int),
        "mlc_autotune": ([vp, ctypes.c_int, dp], ctypes.c_int),
        "mlc_batch_autotune": ([vp, ctypes.c_int], ctypes.c_int),
        "mlc_batch_set_streams": ([vp, ctypes.c_int], ctypes.c_int),
        "mlc_batch_streams": ([vp], ctypes.c_int),
        "mlc_get_variant": ([vp, ctypes.c_int], ctypes.c_int),
        "mlc_batch_create": ([ctypes.POINTER(GridDesc), ctypes.c_int, ctypes.POINTER(vp)],
                             ctypes.c_int),
        "mlc_batch_member": ([vp, ctypes.c_int, ctypes.POINTER(vp)], ctypes.c_int),
        "mlc_batch_step": ([vp, ctypes.c_int], ctypes.c_int),
        "mlc_batch_synchronize": ([vp], ctypes.c_int),
        "mlc_batch_npoints": ([vp], ctypes.c_longlong),
        "mlc_batch_destroy": ([vp], ctypes.c_int),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _lib = lib
    return lib


def exported_symbols() -> list[str]:
    return ["mlc_create", "mlc_destroy", "mlc_set_static", "mlc_put_state",
            "mlc_get_state", "mlc_init", "mlc_step", "mlc_reset_tendencies",
            "mlc_dynamical_core", "mlc_diagnostics", "mlc_status_update", "mlc_set_param",
            "mlc_massck", "mlc_massck_noalloc", "mlc_check_finite", "mlc_synchronize", "mlc_profile",
            "mlc_kernel_count", "mlc_kernel_name", "mlc_kernel_stats",
            "mlc_npoints", "mlc_constant", "mlc_last_error", "mlc_set_variant",
            "mlc_get_variant", "mlc_autotune", "mlc_batch_create", "mlc_batch_autotune",
            "mlc_batch_set_streams", "mlc_batch_streams",
            "mlc_lbc_configure", "mlc_lbc_spectral", "mlc_boundary", "mlc_step_lbc",
            "mlc_init_restart", "mlc_mkslice",
            "mlc_diff_create", "mlc_diff_destroy", "mlc_diff_put", "mlc_diff_get",
            "mlc_diff_initialize", "mlc_diff_calc_coeff", "mlc_diff_d", "mlc_diff_x",
            "mlc_diff_synchronize",
            "mlc_batch_member", "mlc_batch_step", "mlc_batch_synchronize",
            "mlc_batch_npoints", "mlc_batch_destroy"]


def _check(rc: int, what: str) -> None:
    if rc != 0:
        msg = load_library().mlc_last_error().decode(errors="replace")
        raise MolochError(f"{what} failed ({rc}): {msg}")


def _extent(g, name: str):
    """(shape, ub) of the full-extent host array of a field."""
    kz, iy, jx = g.kz, g.iy, g.jx
    if name == "qx" or name == "qxten":
        return (g.nqx, kz, iy, jx), (jx, iy, kz, g.nqx)
    if name == "trac" or name == "chiten":
        return (g.ntr, kz, iy, jx), (jx, iy, kz, g.ntr)
    if name in ("w", "fmzf", "zetaf", "pf", "tke", "tketen"):
        return (kz + 1, iy, jx), (jx, iy, kz + 1, 1)
    if name in ("ps", "mx", "mu", "mv", "hx", "hy", "coru", "corv", "ht", "htu", "htv"):
        return (iy, jx), (jx, iy, 1, 1)
    if name == "rlat":
        return (iy,), (1, iy, 1, 1)
    if name in ("ffilt", "tnudge"):
        return (kz,), (1, 1, kz, 1)
    if name in ("ps0", "ps1", "xlat", "rhox2d", "tp2d", "th700", "ptrop", "ktrop", "kmxpbl"):
        return (iy, jx), (jx, iy, 1, 1)
    return (kz, iy, jx), (jx, iy, kz, 1)


def grid_desc(grid, precision: int) -> GridDesc:
    return GridDesc(jx=grid.jx, iy=grid.iy, kz=grid.kz, nqx=grid.nqx,
                    ipptls=grid.ipptls, nadv=grid.nadv, nsound=grid.nsound,
                    do_divdamp=grid.do_divdamp, do_divfilter=grid.do_divfilter,
                    lrotllr=grid.lrotllr, precision=precision,
                    ibltyp=getattr(grid, "ibltyp", 0), ntr=getattr(grid, "ntr", 0),
                    crm=getattr(grid, "crm", 0), dx=grid.dx, dtsec=grid.dtsec, mo_ztop=grid.mo_ztop,
                    mo_a0=grid.mo_a0)


class Moloch:
    """One MOLOCH domain resident on the GPU (allocate_moloch)."""

    def __init__(self, grid, precision: int = MLC_FP64, *, _handle=None):
        self.lib = load_library()
        self.grid = grid
        self._owned = _handle is None
        if _handle is not None:
            self.h = _handle
            return
        desc = grid_desc(grid, precision)
        h = ctypes.c_void_p()
        _check(self.lib.mlc_create(ctypes.byref(desc), ctypes.byref(h)), "mlc_create")
        self.h = h

    def close(self) -> None:
        if getattr(self, "h", None):
            if self._owned:
                self.lib.mlc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- transfers ---------------------------------------------------------
    def _box(self, name, lb=None, ub=None):
        shape, full_ub = _extent(self.grid, name)
        lb = lb or (1, 1, 1, 1)
        ub = ub or full_ub
        return shape, _I4(*lb), _I4(*ub)

    def put(self, name: str, arr: np.ndarray) -> None:
        shape, lb, ub = self._box(name)
        a = np.ascontiguousarray(arr, dtype=np.float64)
        if a.shape != shape:
            raise MolochError(f"{name}: expected shape {shape}, got {a.shape}")
        fn = self.lib.mlc_set_static if name in STATIC_FIELDS + SLICE_STATICS + TERRAIN_STATICS \
            else self.lib.mlc_put_state
        _check(fn(self.h, FIELD_IDS[name], a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                  lb, ub), f"put {name}")

    def get(self, name: str) -> np.ndarray:
        shape, lb, ub = self._box(name)
        a = np.zeros(shape)
        _check(self.lib.mlc_get_state(self.h, FIELD_IDS[name],
                                      a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                      lb, ub), f"get {name}")
        return a

    def put_box(self, name: str, arr: np.ndarray, lb, ub) -> None:
        a = np.ascontiguousarray(arr, dtype=np.float64)
        fn = self.lib.mlc_set_static if name in STATIC_FIELDS + SLICE_STATICS + TERRAIN_STATICS \
            else self.lib.mlc_put_state
        _check(fn(self.h, FIELD_IDS[name], a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                  _I4(*lb), _I4(*ub)), f"put {name}")

    def get_box(self, name: str, lb, ub) -> np.ndarray:
        shape = tuple(ub[m] - lb[m] + 1 for m in (3, 2, 1, 0))
        a = np.zeros(shape)
        _check(self.lib.mlc_get_state(self.h, FIELD_IDS[name],
                                      a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                      _I4(*lb), _I4(*ub)), f"get {name}")
        return a

    def load(self, fields: dict) -> None:
        """Upload every static metric and the initial state of a domain."""
        for name in STATIC_FIELDS:
            self.put(name, fields[name])
        if all(n in fields for n in TERRAIN_STATICS) and getattr(self.grid, "mo_h", 0) > 0:
            for name in TERRAIN_STATICS:
                self.put(name, fields[name])
            self.set_param(MLC_PARAM_MO_H, self.grid.mo_h)
        for name in STATE_FIELDS:
            self.put(name, fields[name])
        if getattr(self.grid, "ibltyp", 0) == 2:
            self.set_param(MLC_PARAM_TKEMIN, getattr(self.grid, "tkemin", 0.0))
            self.put("tke", fields["tke"])
        if getattr(self.grid, "ntr", 0) > 0:
            self.put("trac", fields["trac"])

    def set_param(self, param: int, value: float) -> None:
        _check(self.lib.mlc_set_param(self.h, param, float(value)), "mlc_set_param")

    # -- restart ------------------------------------------------------------------
    def save_variables(self) -> dict:
        """The SAV variables RegCM's write_savefile takes for MOLOCH
        (Main/mod_savefile.F90:618-628): u, v, w, t, pai, qx, ps, and tke
        with ibltyp 2."""
        return {n: self.get(n) for n in SAV_FIELDS + (("tke",) if self._has_tke() else ())}

    def init_restart(self) -> None:
        """init's restart branch (Main/mod_init.F90:941-955) on the device:
        p, qsat, rho, tvirt and tetav from the SAV variables."""
        _check(self.lib.mlc_init_restart(self.h), "mlc_init_restart")

    def _has_tke(self) -> bool:
        return getattr(self.grid, "ibltyp", 0) == 2

    # -- real-case lateral boundaries -------------------------------------------
    def lbc_configure(self, present_qc=True, present_qi=True, top_nudge=False, nztop=0,
                      spectral=None):
        d = LbcDesc(present_qc=int(present_qc), present_qi=int(present_qi),
                    top_nudge=int(top_nudge), nztop=int(nztop))
        if spectral:
            d.spectral_nudge, d.km, d.lm = 1, int(spectral["km"]), int(spectral["lm"])
            d.dtrad = float(spectral["dtrad"])
        _check(self.lib.mlc_lbc_configure(self.h, ctypes.byref(d)), "mlc_lbc_configure")
        if spectral:
            arrs = [np.ascontiguousarray(spectral[n], dtype=np.float64)
                    for n in ("bvx", "bvy", "cnudge")]
            ptrs = [a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)) for a in arrs]
            _check(self.lib.mlc_lbc_spectral(self.h, *ptrs), "mlc_lbc_spectral")

    def load_lbc(self, b: dict) -> None:
        """Configure and upload a boundary dict (regcm_amd.synth.lbc_inputs,
        with an optional "spectral" entry from synth.spectral_inputs)."""
        self.lbc_configure(b["present_qc"], b["present_qi"], b["top_nudge"], b["nztop"],
                           b.get("spectral"))
        for name in LBC_FIELDS:
            self.put(name, b[name])

    def boundary(self, x1: float) -> None:
        _check(self.lib.mlc_boundary(self.h, x1), "mlc_boundary")

    def moloch_lbc(self, x1s) -> None:
        x = np.ascontiguousarray(x1s, dtype=np.float64)
        _check(self.lib.mlc_step_lbc(self.h, len(x),
                                     x.ctypes.data_as(ctypes.POINTER(ctypes.c_double))),
               "mlc_step_lbc")

    # -- physics input slices ----------------------------------------------
    def mkslice(self, si: dict) -> None:
        """mkslice on the device (Main/mod_slice.F90:112-382) with the switches
        and limits of `si` (regcm_amd.synth.slice_inputs); zetaf and, unless
        irceideal = 1, xlat must have been put.  Read the outputs with
        get(name) for name in SLICE_FIELDS."""
        d = SliceDesc(rhmin=si["rhmin"], rhmax=si["rhmax"], calday=si["calday"],
                      dayspy=si["dayspy"], icldmstrat=int(si["icldmstrat"]),
                      irceideal=int(si["irceideal"]), ibltyp=int(si["ibltyp"]))
        _check(self.lib.mlc_mkslice(self.h, ctypes.byref(d)), "mlc_mkslice")

    def load_slice(self, si: dict) -> None:
        self.put("zetaf", si["zetaf"])
        if int(si["irceideal"]) != 1:
            self.put("xlat", si["xlat"])

    # -- the reference's phases ---------------------------------------------
    def init_moloch(self) -> None:
        _check(self.lib.mlc_init(self.h), "mlc_init")

    def moloch(self, nsteps: int = 1) -> None:
        _check(self.lib.mlc_step(self.h, nsteps), "mlc_step")

    def reset_tendencies(self) -> None:
        _check(self.lib.mlc_reset_tendencies(self.h), "mlc_reset_tendencies")

    def dynamical_core(self) -> None:
        _check(self.lib.mlc_dynamical_core(self.h), "mlc_dynamical_core")

    def diagnostics(self) -> None:
        _check(self.lib.mlc_diagnostics(self.h), "mlc_diagnostics")

    def status_update(self, dtinc: float) -> None:
        _check(self.lib.mlc_status_update(self.h, dtinc), "mlc_status_update")

    def synchronize(self) -> None:
        _check(self.lib.mlc_synchronize(self.h), "mlc_synchronize")

    def massck(self) -> dict:
        out = (ctypes.c_double * 4)()
        _check(self.lib.mlc_massck(self.h, out), "mlc_massck")
        return {"drymass": out[0], "qmass": out[1], "dryadv": out[2], "qadv": out[3]}

    def check_finite(self):
        mx, mn = ctypes.c_double(), ctypes.c_double()
        rc = self.lib.mlc_check_finite(self.h, ctypes.byref(mx), ctypes.byref(mn))
        _check(rc, "mlc_check_finite")
        return mx.value, mn.value

    # -- launch shapes -----------------------------------------------------------
    def set_variant(self, stage: int, value: int) -> None:
        _check(self.lib.mlc_set_variant(self.h, stage, value), "mlc_set_variant")

    def get_variant(self, stage: int) -> int:
        v = self.lib.mlc_get_variant(self.h, stage)
        if v < 0:
            _check(v, "mlc_get_variant")
        return v

    def variants(self) -> dict:
        return {name: self.get_variant(st) for st, name in VARIANT_NAMES.items()}

    def autotune(self, nsteps: int = 2) -> float:
        """Pick the fastest launch shapes for this domain (state unchanged);
        returns the best step time in ms."""
        ms = ctypes.c_double()
        _check(self.lib.mlc_autotune(self.h, nsteps, ctypes.byref(ms)), "mlc_autotune")
        return ms.value

    # -- timing ------------------------------------------------------------------
    def profile(self, enable, slots=None) -> None:
        """Per-kernel event timing: on (all kernels, or only the kernel
        names in `slots`) or off; resets the totals."""
        if not enable:
            code = 0
        elif slots is None:
            code = 1
        else:
            names = [self.lib.mlc_kernel_name(i).decode()
                     for i in range(self.lib.mlc_kernel_count())]
            code = 0
            for n in slots:
                code |= 2 << names.index(n)
        _check(self.lib.mlc_profile(self.h, code), "mlc_profile")

    def kernel_stats(self) -> dict:
        out = {}
        for slot in range(self.lib.mlc_kernel_count()):
            n, ms = ctypes.c_longlong(), ctypes.c_double()
            _check(self.lib.mlc_kernel_stats(self.h, slot, ctypes.byref(n), ctypes.byref(ms)),
                   "mlc_kernel_stats")
            if n.value:
                out[self.lib.mlc_kernel_name(slot).decode()] = (n.value, ms.value)
        return out

    @property
    def npoints(self) -> int:
        return int(self.lib.mlc_npoints(self.h))


class MolochBatch:
    """Batched form: independent domains stepped together, one graph launch
    per step (mlc_batch_*).  `member(m)` is a non-owning Moloch view of
    member m for uploads, downloads, init and the diagnostics."""

    def __init__(self, grids, precision: int = MLC_FP64):
        self.lib = load_library()
        self.grids = list(grids)
        descs = (GridDesc * len(self.grids))(*[grid_desc(g, precision) for g in self.grids])
        b = ctypes.c_void_p()
        _check(self.lib.mlc_batch_create(descs, len(self.grids), ctypes.byref(b)),
               "mlc_batch_create")
        self.b = b
        self._members = []
        for m, g in enumerate(self.grids):
            h = ctypes.c_void_p()
            _check(self.lib.mlc_batch_member(self.b, m, ctypes.byref(h)), "mlc_batch_member")
            self._members.append(Moloch(g, precision, _handle=h))

    def member(self, m: int) -> Moloch:
        return self._members[m]

    def __len__(self) -> int:
        return len(self._members)

    def step(self, nsteps: int = 1) -> None:
        _check(self.lib.mlc_batch_step(self.b, nsteps), "mlc_batch_step")

    def autotune(self, nsteps: int = 2) -> None:
        _check(self.lib.mlc_batch_autotune(self.b, nsteps), "mlc_batch_autotune")

    def set_streams(self, k: int) -> None:
        _check(self.lib.mlc_batch_set_streams(self.b, k), "mlc_batch_set_streams")

    @property
    def streams(self) -> int:
        return int(self.lib.mlc_batch_streams(self.b))

    def synchronize(self) -> None:
        _check(self.lib.mlc_batch_synchronize(self.b), "mlc_batch_synchronize")

    @property
    def npoints(self) -> int:
        return int(self.lib.mlc_batch_npoints(self.b))

    def close(self) -> None:
        if getattr(self, "b", None):
            for m in self._members:
                m.h = None
            self.lib.mlc_batch_destroy(self.b)
            self.b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class Diffusion:
    """The MM-core horizontal-diffusion operator on the device
    (Main/mod_diffusion.F90): `Diffusion(dg)` with a descriptor like
    regcm_amd.synth.DiffGrid, `load(inputs)`, `initialize()`,
    `calc_coeff()`, `diffu_d()`, `diffu_x(form, n1, n2, fac)`, `get(name)`."""

    def __init__(self, dg):
        self.lib = load_library()
        self.dg = dg
        d = DiffDesc(jx=dg.jx, iy=dg.iy, kz=dg.kz, idiffu=dg.idiffu, idynamic=dg.idynamic,
                     diffu_hgtf=dg.diffu_hgtf, nspec=dg.nspec, dx=dg.dx, dtsec=dg.dtsec,
                     ckh=dg.ckh, adyndif=dg.adyndif)
        h = ctypes.c_void_p()
        _check(self.lib.mlc_diff_create(ctypes.byref(d), ctypes.byref(h)), "mlc_diff_create")
        self.h = h

    def _shape(self, name):
        dg = self.dg
        kz, iy, jx = dg.kz, dg.iy, dg.jx
        if name in ("ht", "msfd", "pc", "pd", "hgfact"):
            return (iy, jx), (jx, iy, 1, 1)
        if name in ("wx", "xkcf"):
            return (kz + 1, iy, jx), (jx, iy, kz + 1, 1)
        if name in ("f", "ften"):
            return (dg.nspec, kz + 1, iy, jx), (jx, iy, kz + 1, dg.nspec)
        return (kz, iy, jx), (jx, iy, kz, 1)

    def put(self, name: str, arr: np.ndarray, ub=None) -> None:
        shape, full = self._shape(name)
        a = np.ascontiguousarray(arr, dtype=np.float64)
        ub = ub or full
        if a.size != int(np.prod(ub)):
            raise MolochError(f"{name}: {a.shape} does not fill the box (1..{ub})")
        _check(self.lib.mlc_diff_put(self.h, DIFF_FIELD_IDS[name],
                                     a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                     _I4(1, 1, 1, 1), _I4(*ub)), f"put {name}")

    def get(self, name: str, ub=None) -> np.ndarray:
        shape, full = self._shape(name)
        ub = ub or full
        out_shape = (ub[3], ub[2], ub[1], ub[0])[4 - len(shape):] if len(shape) > 2 \
            else (ub[1], ub[0])
        a = np.zeros(out_shape)
        _check(self.lib.mlc_diff_get(self.h, DIFF_FIELD_IDS[name],
                                     a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                     _I4(1, 1, 1, 1), _I4(*ub)), f"get {name}")
        return a

    def load(self, inp: dict) -> None:
        for name in ("ht", "msfd", "pc", "pd", "ud", "vd", "wx", "uten", "vten", "f", "ften"):
            if name in inp:
                self.put(name, inp[name])

    def initialize(self) -> None:
        _check(self.lib.mlc_diff_initialize(self.h), "mlc_diff_initialize")

    def calc_coeff(self) -> None:
        _check(self.lib.mlc_diff_calc_coeff(self.h), "mlc_diff_calc_coeff")

    def diffu_d(self) -> None:
        _check(self.lib.mlc_diff_d(self.h), "mlc_diff_d")

    def diffu_x(self, form: int, n1: int = 1, n2: int = 1, fac: float = 1.0) -> None:
        _check(self.lib.mlc_diff_x(self.h, form, n1, n2, fac), "mlc_diff_x")

    def synchronize(self) -> None:
        _check(self.lib.mlc_diff_synchronize(self.h), "mlc_diff_synchronize")

    def close(self) -> None:
        if getattr(self, "h", None):
            self.lib.mlc_diff_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def relative_l2(ref: np.ndarray, got: np.ndarray) -> float:
    """||got - ref||_2 / ||ref||_2 (0 when both vanish)."""
    den = float(np.linalg.norm(ref))
    num = float(np.linalg.norm(got - ref))
    return num / den if den > 0 else num


def constant(name: str) -> float:
    return load_library().mlc_constant(name.encode())
