"""ctypes wrapper of the oracle library (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the kestrel_amd package.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from kestrel_amd.config import KstConfig
from kestrel_amd import scenario

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_LIB_PERT = None

_dp = ctypes.POINTER(ctypes.c_double)


def _ptr(a):
    if a is None:
        return None
    return a.ctypes.data_as(_dp)


class KstStats(ctypes.Structure):
    _fields_ = [("steps", ctypes.c_int64), ("rejected", ctypes.c_int64),
                ("rhs_evals", ctypes.c_int64), ("cell_updates", ctypes.c_int64),
                ("active_tiles", ctypes.c_int32), ("ghost_tiles", ctypes.c_int32),
                ("t", ctypes.c_double), ("last_dt", ctypes.c_double),
                ("min_dt", ctypes.c_double), ("max_dt", ctypes.c_double),
                ("redist_cells", ctypes.c_int64), ("redist_ghost", ctypes.c_int64)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


def lib(perturbed: bool = False):
    """liboracle.so, or with perturbed=True liboracle_pert.so, whose per-cell
    transcendental functions return one ulp up (the conditioning probe)."""
    global _LIB, _LIB_PERT
    if perturbed:
        if _LIB_PERT is None:
            _LIB_PERT = _load("liboracle_pert.so")
        return _LIB_PERT
    if _LIB is None:
        _LIB = _load("liboracle.so")
    return _LIB


def _load(name: str):
    path = os.path.join(_HERE, name)
    if not os.path.exists(path):
        raise RuntimeError(f"oracle/{name} is not built (make -C oracle)")
    L = ctypes.CDLL(path)
    L.ko_create.restype = ctypes.c_void_p
    L.ko_create.argtypes = [ctypes.POINTER(KstConfig)]
    L.ko_destroy.argtypes = [ctypes.c_void_p]
    L.ko_error.restype = ctypes.c_char_p
    L.ko_error.argtypes = [ctypes.c_void_p]
    for name in ["ko_set_topography", "ko_add_caps", "ko_add_flux_sources", "ko_finish_sources",
                 "ko_set_state", "ko_integrate_to", "ko_step_n", "ko_get_state",
                 "ko_get_maxima", "ko_get_fields13", "ko_get_ghost_slopes", "ko_rhs",
                 "ko_tile_lists", "ko_set_maxima", "ko_set_cell_topography"]:
        getattr(L, name).restype = ctypes.c_int
    L.ko_set_topography.argtypes = [ctypes.c_void_p, _dp]
    L.ko_set_maxima.argtypes = [ctypes.c_void_p, ctypes.POINTER(_dp)]
    L.ko_set_cell_topography.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, ctypes.POINTER(ctypes.c_uint8)]
    L.ko_add_caps.argtypes = [ctypes.c_void_p, ctypes.c_int] + [_dp] * 7 + [ctypes.POINTER(ctypes.c_int)]
    L.ko_add_flux_sources.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int32)]
    L.ko_finish_sources.argtypes = [ctypes.c_void_p]
    L.ko_set_state.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, _dp, _dp,
                               ctypes.POINTER(ctypes.c_uint8), ctypes.c_double, _dp]
    L.ko_integrate_to.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.POINTER(KstStats)]
    L.ko_step_n.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.POINTER(KstStats)]
    L.ko_get_state.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, _dp, _dp, _dp,
                               ctypes.POINTER(ctypes.c_uint8), _dp]
    L.ko_get_maxima.argtypes = [ctypes.c_void_p, ctypes.POINTER(_dp)]
    L.ko_get_fields13.argtypes = [ctypes.c_void_p, _dp]
    L.ko_get_ghost_slopes.argtypes = [ctypes.c_void_p, _dp, _dp]
    L.ko_rhs.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double, _dp, _dp, _dp]
    L.ko_tile_lists.argtypes = [ctypes.c_void_p] + [ctypes.POINTER(ctypes.c_int32)] * 4
    L.ko_volume.restype = ctypes.c_int
    L.ko_volume.argtypes = [ctypes.c_void_p, _dp]
    L.ko_volume_terms.restype = ctypes.c_int
    L.ko_volume_terms.argtypes = [ctypes.c_void_p, _dp, ctypes.c_int64]
    for name in ["ko_get_tile_bt", "ko_morpho_rhs", "ko_probe_count", "ko_probe_get",
                 "ko_probe_candidates"]:
        getattr(L, name).restype = ctypes.c_int
    L.ko_get_tile_bt.argtypes = [ctypes.c_void_p, _dp]
    L.ko_step_log.restype = ctypes.c_int64
    L.ko_step_log.argtypes = [ctypes.c_void_p, _dp, ctypes.c_int64]
    L.ko_limiter.restype = ctypes.c_double
    L.ko_limiter.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double]
    L.ko_morpho_rhs.argtypes = [ctypes.c_void_p, _dp, _dp]
    L.ko_set_tile_bt.restype = ctypes.c_int
    L.ko_set_tile_bt.argtypes = [ctypes.c_void_p, _dp]
    L.ko_set_fields13.restype = ctypes.c_int
    L.ko_set_fields13.argtypes = [ctypes.c_void_p, _dp, _dp]
    L.ko_redistribute.restype = ctypes.c_int
    L.ko_redistribute.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, _dp, ctypes.c_int,
                                  ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32),
                                  ctypes.POINTER(ctypes.c_int32), _dp, _dp, _dp,
                                  ctypes.POINTER(ctypes.c_int)]
    L.ko_probe_enable.restype = None
    L.ko_probe_enable.argtypes = [ctypes.c_void_p, ctypes.c_int]
    L.ko_probe_count.argtypes = [ctypes.c_void_p]
    L.ko_probe_dt.restype = ctypes.c_double
    L.ko_probe_dt.argtypes = [ctypes.c_void_p]
    L.ko_probe_get.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp]
    i32p = ctypes.POINTER(ctypes.c_int32)
    L.ko_probe_candidates.argtypes = [ctypes.c_void_p, i32p, i32p, i32p, _dp,
                                      ctypes.POINTER(ctypes.c_uint8)]
    return L


SHAPES = {"flat": 0, "para": 1, "level": 2}


class Oracle:
    """One simulation in the oracle."""

    def __init__(self, cfg: KstConfig, perturbed: bool = False):
        self.cfg = cfg
        self.L = lib(perturbed)
        self.h = self.L.ko_create(ctypes.byref(cfg))
        self._check(0)

    def close(self):
        if self.h:
            self.L.ko_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _check(self, rc):
        msg = self.L.ko_error(self.h).decode()
        if rc != 0 or msg:
            raise RuntimeError(f"oracle error {rc}: {msg}")

    def set_topography(self, b0v):
        self.b0v = np.ascontiguousarray(b0v, dtype=np.float64)
        self._check(self.L.ko_set_topography(self.h, _ptr(self.b0v)))

    def add_caps(self, caps):
        n = len(caps)
        arrs = [np.array([getattr(c, f) for c in caps], dtype=np.float64)
                for f in ("x", "y", "radius", "height", "psi", "u", "v")]
        shp = (ctypes.c_int * max(n, 1))(*[SHAPES[c.shape] for c in caps])
        self._arrs = arrs
        self._check(self.L.ko_add_caps(self.h, n, *[_ptr(a) for a in arrs], shp))

    def add_flux_sources(self):
        n = max(1, self.cfg.n_sources)
        out = (ctypes.c_int32 * n)()
        self._check(self.L.ko_add_flux_sources(self.h, out))
        return list(out)[: self.cfg.n_sources]

    def finish_sources(self):
        self._check(self.L.ko_finish_sources(self.h))

    def set_state(self, st: "scenario.InitialState", t=0.0, btv=None):
        ts = np.ascontiguousarray(st.tile_state, dtype=np.uint8)
        bt = None if btv is None else np.ascontiguousarray(btv, dtype=np.float64)
        hm = None if st.hnmax0 is None else np.ascontiguousarray(st.hnmax0)
        self._check(self.L.ko_set_state(
            self.h, _ptr(np.ascontiguousarray(st.w)), _ptr(np.ascontiguousarray(st.rhoHnu)),
            _ptr(np.ascontiguousarray(st.rhoHnv)), _ptr(np.ascontiguousarray(st.Hnpsi)), _ptr(bt),
            ts.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), float(t), _ptr(hm)))
        if getattr(st, "psimax0", None) is not None:
            mx = self.maxima()
            mx[8] = np.ascontiguousarray(st.psimax0, dtype=np.float64)
            self.set_maxima(mx)

    def set_maxima(self, arrs):
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in arrs]
        ptrs = (_dp * 11)(*[_ptr(a) for a in arrs])
        self._check(self.L.ko_set_maxima(self.h, ptrs))

    def set_cell_topography(self, btc, dbdx, dbdy, cells):
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (btc, dbdx, dbdy)]
        m = np.ascontiguousarray(cells, dtype=np.uint8)
        self._check(self.L.ko_set_cell_topography(self.h, *[_ptr(a) for a in arrs],
                                                  m.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))))

    def integrate_to(self, tend):
        st = KstStats()
        self._check(self.L.ko_integrate_to(self.h, float(tend), ctypes.byref(st)))
        return st.as_dict()

    def step_n(self, n):
        st = KstStats()
        self._check(self.L.ko_step_n(self.h, int(n), ctypes.byref(st)))
        return st.as_dict()

    def state(self):
        c = self.cfg
        shp = (c.NY, c.NX)
        out = {k: np.zeros(shp) for k in ("w", "rhoHnu", "rhoHnv", "Hnpsi", "Hn")}
        out["btv"] = np.zeros((c.NY + 1, c.NX + 1))
        ts = np.zeros(c.ntiles, dtype=np.uint8)
        t = ctypes.c_double()
        self._check(self.L.ko_get_state(self.h, _ptr(out["w"]), _ptr(out["rhoHnu"]), _ptr(out["rhoHnv"]),
                                        _ptr(out["Hnpsi"]), _ptr(out["Hn"]), _ptr(out["btv"]),
                                        ts.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), ctypes.byref(t)))
        out["tile_state"] = ts
        out["t"] = t.value
        return out

    def maxima(self):
        c = self.cfg
        arrs = [np.zeros((c.NY, c.NX)) for _ in range(11)]
        ptrs = (_dp * 11)(*[_ptr(a) for a in arrs])
        self._check(self.L.ko_get_maxima(self.h, ptrs))
        return arrs

    def fields13(self):
        c = self.cfg
        out = np.zeros((13, c.NY, c.NX))
        self._check(self.L.ko_get_fields13(self.h, _ptr(out)))
        return out

    def volume(self) -> np.ndarray:
        """CalculateVolume's Volume.txt columns 2-7 (src/Output.f90:617-735)."""
        out = np.zeros(6)
        self._check(self.L.ko_volume(self.h, _ptr(out)))
        return out

    def volume_terms(self) -> np.ndarray:
        """Per-cell (flow volume, mass, bed, solids) terms in summation order."""
        cap = self.cfg.NX * self.cfg.NY
        t = np.zeros((cap, 4))
        n = self.L.ko_volume_terms(self.h, _ptr(t), cap)
        if n < 0:
            raise RuntimeError("ko_volume_terms: more cells than the grid holds")
        return t[:n]

    def ghost_slopes(self):
        c = self.cfg
        sx = np.zeros((c.NY, c.NX)); sy = np.zeros((c.NY, c.NX))
        self._check(self.L.ko_get_ghost_slopes(self.h, _ptr(sx), _ptr(sy)))
        return sx, sy

    def rhs(self, substep=1, t=0.0):
        c = self.cfg
        e = np.zeros((4, c.NY, c.NX)); i = np.zeros((4, c.NY, c.NX))
        adv = ctypes.c_double()
        self._check(self.L.ko_rhs(self.h, substep, float(t), _ptr(e), _ptr(i), ctypes.byref(adv)))
        return e, i, adv.value

    def _nv(self):
        c = self.cfg
        return (c.ny_per_tile + 1, c.nx_per_tile + 1)

    def tile_bt(self):
        out = np.zeros((self.cfg.ntiles,) + self._nv())
        self._check(self.L.ko_get_tile_bt(self.h, _ptr(out)))
        return out

    def morpho_rhs(self):
        c = self.cfg
        emd = np.zeros((c.NY, c.NX))
        dtbt = np.zeros((c.ntiles,) + self._nv())
        self._check(self.L.ko_morpho_rhs(self.h, _ptr(emd), _ptr(dtbt)))
        return emd, dtbt

    def set_tile_bt(self, bt_tiles):
        btt = np.ascontiguousarray(bt_tiles, dtype=np.float64)
        self._check(self.L.ko_set_tile_bt(self.h, _ptr(btt)))

    def set_fields13(self, f13, bt_tiles=None):
        f13 = np.ascontiguousarray(f13, dtype=np.float64)
        btt = None if bt_tiles is None else np.ascontiguousarray(bt_tiles, dtype=np.float64)
        self._check(self.L.ko_set_fields13(self.h, _ptr(f13), _ptr(btt)))

    def redistribute(self, i0, i3, cand):
        """AddToRedistList + RedistributeGrid on the given containers."""
        f0, bt0 = [np.ascontiguousarray(a, dtype=np.float64) for a in i0]
        f3, bt3 = [np.ascontiguousarray(a, dtype=np.float64) for a in i3]
        k, i, j = [np.ascontiguousarray(a, dtype=np.int32) for a in cand[:3]]
        ex = np.ascontiguousarray(cand[3], dtype=np.float64)
        fout = np.zeros_like(f3); btout = np.zeros_like(bt3)
        failed = ctypes.c_int()
        i32p = ctypes.POINTER(ctypes.c_int32)
        self._check(self.L.ko_redistribute(self.h, _ptr(f0), _ptr(bt0), _ptr(f3), _ptr(bt3), len(k),
                                           k.ctypes.data_as(i32p), i.ctypes.data_as(i32p),
                                           j.ctypes.data_as(i32p), _ptr(ex), _ptr(fout), _ptr(btout),
                                           ctypes.byref(failed)))
        return fout, btout, failed.value

    def probe_enable(self, on=True):
        """on=True: first redistribution; on=2: first morphodynamic step."""
        self.L.ko_probe_enable(self.h, int(on))

    def probe(self):
        """The first redistribution recorded since probe_enable, or None."""
        n = self.L.ko_probe_count(self.h)
        if n < 0:
            return None
        c = self.cfg
        res = {}
        for which, name in enumerate(("i0", "i3_before", "i3_after")):
            f13 = np.zeros((13, c.NY, c.NX))
            btt = np.zeros((c.ntiles,) + self._nv())
            self._check(self.L.ko_probe_get(self.h, which, _ptr(f13), _ptr(btt)))
            res[name] = (f13, btt)
        k = np.zeros(n, np.int32); i = np.zeros(n, np.int32); j = np.zeros(n, np.int32)
        ex = np.zeros(n)
        i32p = ctypes.POINTER(ctypes.c_int32)
        ts = np.zeros(c.ntiles, np.uint8)
        self._check(self.L.ko_probe_candidates(self.h, k.ctypes.data_as(i32p), i.ctypes.data_as(i32p),
                                               j.ctypes.data_as(i32p), _ptr(ex),
                                               ts.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))))
        res["cand"] = (k, i, j, ex)
        res["tile_state"] = ts
        res["dt"] = self.L.ko_probe_dt(self.h)
        return res

    def step_log(self, cap: int = 0):
        """Start recording (t, dt) of up to cap accepted steps (cap = 0 stops);
        returns the pairs the previous recording received."""
        prev = getattr(self, "_steplog", None)
        buf = np.zeros((max(cap, 1), 2)) if cap > 0 else None
        n = self.L.ko_step_log(self.h, _ptr(buf) if buf is not None else None, int(cap))
        self._steplog = buf
        return prev[:n].copy() if prev is not None else np.zeros((0, 2))

    def tile_lists(self):
        n = self.cfg.ntiles
        a = (ctypes.c_int32 * n)(); g = (ctypes.c_int32 * n)()
        na = ctypes.c_int32(); ng = ctypes.c_int32()
        self._check(self.L.ko_tile_lists(self.h, a, ctypes.byref(na), g, ctypes.byref(ng)))
        return list(a)[: na.value], list(g)[: ng.value]


# ---------------------------------------------------------------------------
# oracle/_ref: the reference's own hydraulic operator (flang build)
# ---------------------------------------------------------------------------
_REF = None


def ref_lib():
    global _REF
    if _REF is None:
        path = os.path.join(_HERE, "_ref", "libkestrel_ref.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/_ref/libkestrel_ref.so is not built (make -C oracle ref)")
        L = ctypes.CDLL(path)
        u8 = ctypes.POINTER(ctypes.c_int8)
        L.kref_rhs.restype = None
        L.kref_rhs.argtypes = [ctypes.POINTER(KstConfig), u8, u8, _dp, _dp, _dp, _dp, _dp,
                               ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_double,
                               ctypes.c_double, _dp, _dp, _dp]
        L.kref_morpho_rhs.restype = None
        L.kref_morpho_rhs.argtypes = [ctypes.POINTER(KstConfig), u8, _dp, _dp, _dp, _dp, _dp]
        i32p = ctypes.POINTER(ctypes.c_int32)
        L.kref_redistribute.restype = None
        L.kref_redistribute.argtypes = [ctypes.POINTER(KstConfig), u8, _dp, _dp, _dp, _dp, _dp,
                                        ctypes.c_int32, i32p, i32p, i32p, _dp, _dp, _dp, i32p]
        L.kref_bicubic.restype = None
        L.kref_bicubic.argtypes = [_dp, ctypes.c_int, ctypes.c_int, _dp, _dp, ctypes.c_int, _dp]
        L.kref_topog.restype = None
        L.kref_topog.argtypes = [ctypes.c_int, _dp, ctypes.c_int, _dp, ctypes.c_int, _dp, ctypes.c_int, _dp]
        _REF = L
    return _REF


REF_TOPOG = ["x2slopes", "usgs", "flume", "channel power law", "channel trapezium", "xbislope",
             "xtrislope", "xhump", "xtanh", "xyparab"]


def ref_topog(name: str, params, xv, yv):
    """The reference's TopogFunc `name` at vertex coordinates xv, yv:
    b0[len(yv), len(xv)] (row-major, y outer)."""
    L = ref_lib()
    prm = np.ascontiguousarray(params, dtype=np.float64)
    x = np.ascontiguousarray(xv, dtype=np.float64)
    y = np.ascontiguousarray(yv, dtype=np.float64)
    out = np.zeros((len(y), len(x)))     # Fortran b0(nx, ny) in column order = [ny][nx] row-major
    L.kref_topog(REF_TOPOG.index(name) + 1, _ptr(prm), len(prm), _ptr(x), len(x), _ptr(y), len(y), _ptr(out))
    return out


def ref_available() -> bool:
    return os.path.exists(os.path.join(_HERE, "_ref", "libkestrel_ref.so"))


def ref_rhs(cfg: KstConfig, tile_state, has_src, b0v, f13, sx, sy, *, recompute_ghost=False,
            substep=1, t=0.0, grid_t=0.0, bt_tiles=None):
    """CalculateHydraulicRHS of the reference on the given global state
    (per-tile vertex bt views if given, else bt = 0)."""
    L = ref_lib()
    NX, NY = cfg.NX, cfg.NY
    ts = np.ascontiguousarray(tile_state, dtype=np.int8)
    hs = np.ascontiguousarray(has_src, dtype=np.int8)
    btt = np.zeros(1) if bt_tiles is None else bt_tiles
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (b0v, f13, sx, sy, btt)]
    e = np.zeros((4, NY, NX)); i = np.zeros((4, NY, NX)); adv = ctypes.c_double()
    u8 = ctypes.POINTER(ctypes.c_int8)
    L.kref_rhs(ctypes.byref(cfg), ts.ctypes.data_as(u8), hs.ctypes.data_as(u8),
               *[_ptr(a) for a in arrs], int(bt_tiles is not None), int(recompute_ghost),
               int(substep), float(t), float(grid_t), _ptr(e), _ptr(i), ctypes.byref(adv))
    return e, i, adv.value


def ref_morpho_rhs(cfg: KstConfig, tile_state, b0v, f13, bt_tiles):
    """CalculateMorphodynamicRHS of the reference: (EminusD, ddtExplicitBt per tile)."""
    L = ref_lib()
    ts = np.ascontiguousarray(tile_state, dtype=np.int8)
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (b0v, f13, bt_tiles)]
    emd = np.zeros((cfg.NY, cfg.NX))
    dtbt = np.zeros_like(arrs[2])
    L.kref_morpho_rhs(ctypes.byref(cfg), ts.ctypes.data_as(ctypes.POINTER(ctypes.c_int8)),
                      *[_ptr(a) for a in arrs], _ptr(emd), _ptr(dtbt))
    return emd, dtbt


def ref_redistribute(cfg: KstConfig, tile_state, b0v, i0, i3, cand):
    """AddToRedistList over the candidates, then RedistributeGrid, in the
    reference: returns (fields13, bt per tile) of intermed3 and the refine flag."""
    L = ref_lib()
    ts = np.ascontiguousarray(tile_state, dtype=np.int8)
    f0, bt0 = [np.ascontiguousarray(a, dtype=np.float64) for a in i0]
    f3, bt3 = [np.ascontiguousarray(a, dtype=np.float64) for a in i3]
    k, i, j, ex = [np.ascontiguousarray(a) for a in cand]
    fout = np.zeros_like(f3); btout = np.zeros_like(bt3)
    ref = ctypes.c_int32()
    i32p = ctypes.POINTER(ctypes.c_int32)
    L.kref_redistribute(ctypes.byref(cfg), ts.ctypes.data_as(ctypes.POINTER(ctypes.c_int8)),
                        _ptr(np.ascontiguousarray(b0v, dtype=np.float64)), _ptr(f0), _ptr(bt0),
                        _ptr(f3), _ptr(bt3), len(k), k.astype(np.int32).ctypes.data_as(i32p),
                        i.astype(np.int32).ctypes.data_as(i32p), j.astype(np.int32).ctypes.data_as(i32p),
                        _ptr(ex.astype(np.float64)), _ptr(fout), _ptr(btout), ctypes.byref(ref))
    return fout, btout, ref.value


def _kahan_add(x, s, c):
    y = x - c
    t = s + y
    c = (t - s) - y
    return t, c


def volume_from_state(cfg: KstConfig, b0v: np.ndarray, state: dict) -> dict:
    """CalculateVolume (src/Output.f90:617-735) restated over a downloaded
    state: Hn and rho re-derived as ComputeDesingularisedVariables does
    (src/HydraulicRHS.f90:762-878), then Kahan sums (src/utilities.f90:439-448)
    over the active tiles in index order, cells row by row.  The cell bed is
    taken from the state's single-valued vertex bt, so with morphodynamics
    (where tiles' copies of a shared vertex may differ in the last bits) it
    can differ from tileContainer's in the last bits."""
    b0c, btc, bx, by = scenario.cell_topography(cfg, b0v, state.get("btv"))
    gam = scenario.gamma(cfg, bx, by)
    w, hnpsi = state["w"], state["Hnpsi"]
    # derived Hn and rho as ComputeDesingularisedVariables leaves them
    hn = (-btc + (w - b0c)) * gam
    hn = np.where(hn < 0.0, 0.0, hn)
    hp = np.where(hnpsi < 0.0, 0.0, hnpsi)
    eps = cfg.height_threshold
    den = hn * hn + np.maximum(hn * hn, eps * eps)
    psi = np.minimum(2.0 * hn * hp / den, cfg.max_pack)
    rho = cfg.rhow + (cfg.rhos - cfg.rhow) * psi
    vol = mass = bed = solids = 0.0
    cv = cm = cb = cs = 0.0
    ts = np.asarray(state["tile_state"]).reshape(-1)
    nx, ny = cfg.nx_per_tile, cfg.ny_per_tile
    for k in np.nonzero(ts == 1)[0]:
        ti, tj = k % cfg.nx_tiles, k // cfg.nx_tiles
        sl = (slice(tj * ny, (tj + 1) * ny), slice(ti * nx, (ti + 1) * nx))
        dv = ((w[sl] - b0c[sl] - btc[sl]) * gam[sl] * gam[sl]).ravel()
        dm = (rho[sl] * hn[sl] * gam[sl]).ravel()
        db = btc[sl].ravel()
        ds = (hnpsi[sl] * gam[sl]).ravel()
        for q in range(dv.size):
            vol, cv = _kahan_add(dv[q], vol, cv)
            mass, cm = _kahan_add(dm[q], mass, cm)
            bed, cb = _kahan_add(db[q], bed, cb)
            solids, cs = _kahan_add(ds[q], solids, cs)
    a = cfg.dx * cfg.dy
    return {"volume": vol * cfg.dx * cfg.dy, "mass": mass * cfg.dx * cfg.dy,
            "bed": bed * cfg.dx * cfg.dy, "solids": solids * cfg.dx * cfg.dy, "cell_area": a}


def ref_bicubic(z, x, y):
    """The reference's Bicubic of z[nx, ny] (Fortran z(x, y), 1-based data
    coordinates) at the points x, y."""
    L = ref_lib()
    zf = np.asfortranarray(z, dtype=np.float64)
    xs = np.ascontiguousarray(x, dtype=np.float64)
    ys = np.ascontiguousarray(y, dtype=np.float64)
    out = np.zeros(len(xs))
    L.kref_bicubic(zf.ctypes.data_as(_dp), zf.shape[0], zf.shape[1], _ptr(xs), _ptr(ys), len(xs), _ptr(out))
    return out
