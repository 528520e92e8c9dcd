"""ctypes binding of libkestrel_amd.so (the C-ABI of include/kestrel_step.h).

``Stepper`` is the host-side mirror of the reference's step driver: one
handle per simulation, created after the initial conditions are painted,
then ``integrate_to(tend)`` per output interval as Run does with IntegrateTo
(src/TimeStepper.f90:73-113).  The HIP library is required: there is no CPU
fallback, and a missing library raises.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from .config import KstConfig

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libkestrel_amd.so")
_LIB = None
_dp = ctypes.POINTER(ctypes.c_double)

# Kernel ids of kst_profile_read (enum kst_profile_kernel).
PROFILE_KERNELS = 16
PROFILE_NAMES = ["stage_kernel<1>", "stage_kernel<2>", "stage_kernel<3>", "stage_kernel<4>",
                 "stage_kernel<1> (H2)", "stage_kernel<2> (H2)", "stage_kernel<3> (H2)",
                 "stage_kernel<4> (H2)", "morpho_kernel<1>", "morpho_kernel<2>", "morpho_kernel<3>",
                 "redistribute_kernel", "face_topo_kernel", "commit_kernel", "activate_kernel",
                 "finalize_kernel"]
_u8p = ctypes.POINTER(ctypes.c_uint8)

EXPORTS = ["kst_config_defaults", "kst_config_derive", "kst_create", "kst_set_topography",
           "kst_set_state", "kst_integrate_to", "kst_step_n", "kst_get_state", "kst_get_maxima",
           "kst_last_error", "kst_destroy", "kst_profile", "kst_profile_read", "kst_get_tile_bt",
           "kst_batch_create", "kst_batch_attach", "kst_batch_detach", "kst_batch_start",
           "kst_batch_run", "kst_batch_stats", "kst_batch_last_error", "kst_batch_destroy",
           "kst_get_cell_topography", "kst_set_maxima", "kst_get_output", "kst_volume", "kst_set_grid_cap",
           "kst_set_redistribution_chunk", "kst_autotune", "kst_autotune_choices",
           "kst_volume_noalloc", "kst_set_cell_topography", "kst_set_activation_candidates"]


class KstStats(ctypes.Structure):
    _fields_ = [("steps", ctypes.c_int64), ("rejected", ctypes.c_int64),
                ("rhs_evals", ctypes.c_int64), ("cell_updates", ctypes.c_int64),
                ("active_tiles", ctypes.c_int32), ("ghost_tiles", ctypes.c_int32),
                ("t", ctypes.c_double), ("last_dt", ctypes.c_double),
                ("min_dt", ctypes.c_double), ("max_dt", ctypes.c_double),
                ("redist_cells", ctypes.c_int64), ("redist_ghost", ctypes.c_int64)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class KestrelError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"kestrel_amd error {code}: {msg}")
        self.code = code


def lib():
    """Load the HIP library; raises if it is not built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m kestrel_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        L.kst_config_defaults.argtypes = [ctypes.POINTER(KstConfig)]
        L.kst_config_defaults.restype = None
        L.kst_config_derive.argtypes = [ctypes.POINTER(KstConfig)]
        L.kst_config_derive.restype = None
        L.kst_create.argtypes = [ctypes.POINTER(KstConfig), ctypes.POINTER(ctypes.c_void_p)]
        L.kst_set_topography.argtypes = [ctypes.c_void_p, _dp]
        L.kst_set_state.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, _dp, _dp, _u8p, ctypes.c_double, _dp]
        L.kst_integrate_to.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.POINTER(KstStats)]
        L.kst_step_n.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.POINTER(KstStats)]
        L.kst_get_state.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, _dp, _dp, _dp, _u8p, _dp]
        L.kst_get_maxima.argtypes = [ctypes.c_void_p, ctypes.POINTER(_dp)]
        L.kst_last_error.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int]
        L.kst_profile.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.kst_profile.restype = ctypes.c_int
        L.kst_set_grid_cap.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.kst_set_grid_cap.restype = ctypes.c_int
        L.kst_set_redistribution_chunk.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
        L.kst_set_redistribution_chunk.restype = ctypes.c_int
        L.kst_set_activation_candidates.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.kst_set_activation_candidates.restype = ctypes.c_int
        L.kst_autotune.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.kst_autotune.restype = ctypes.c_int
        i32p = ctypes.POINTER(ctypes.c_int32)
        L.kst_autotune_choices.argtypes = [ctypes.c_void_p, ctypes.c_int, i32p, i32p, i32p, _dp]
        L.kst_autotune_choices.restype = ctypes.c_int
        L.kst_volume.argtypes = [ctypes.c_void_p, _dp]
        L.kst_volume.restype = ctypes.c_int
        L.kst_volume_noalloc.argtypes = [ctypes.c_void_p, _dp]
        L.kst_volume_noalloc.restype = ctypes.c_int
        L.kst_get_output.argtypes = [ctypes.c_void_p, ctypes.c_int32, _dp, ctypes.POINTER(ctypes.c_int32),
                                     ctypes.POINTER(ctypes.c_int32)]
        L.kst_get_output.restype = ctypes.c_int
        L.kst_set_maxima.argtypes = [ctypes.c_void_p, ctypes.POINTER(_dp)]
        L.kst_set_maxima.restype = ctypes.c_int
        L.kst_profile_read.argtypes = [ctypes.c_void_p, _dp, ctypes.POINTER(ctypes.c_int64)]
        L.kst_profile_read.restype = ctypes.c_int
        L.kst_get_tile_bt.argtypes = [ctypes.c_void_p, _dp]
        L.kst_get_cell_topography.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, _dp]
        L.kst_set_cell_topography.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, _u8p]
        L.kst_set_cell_topography.restype = ctypes.c_int
        L.kst_destroy.argtypes = [ctypes.c_void_p]
        L.kst_destroy.restype = None
        L.kst_batch_create.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]
        L.kst_batch_attach.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)]
        L.kst_batch_detach.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.kst_batch_start.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_int64]
        L.kst_batch_run.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.POINTER(ctypes.c_int)]
        L.kst_batch_stats.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(KstStats)]
        L.kst_batch_last_error.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int]
        L.kst_batch_destroy.argtypes = [ctypes.c_void_p]
        L.kst_batch_destroy.restype = None
        for n in ["kst_batch_create", "kst_batch_attach", "kst_batch_detach", "kst_batch_start",
                  "kst_batch_run", "kst_batch_stats", "kst_batch_last_error"]:
            getattr(L, n).restype = ctypes.c_int
        for n in ["kst_create", "kst_set_topography", "kst_set_state", "kst_integrate_to",
                  "kst_step_n", "kst_get_state", "kst_get_maxima", "kst_last_error",
                  "kst_get_tile_bt", "kst_get_cell_topography"]:
            getattr(L, n).restype = ctypes.c_int
        _LIB = L
    return _LIB


def _ptr(a):
    return None if a is None else a.ctypes.data_as(_dp)


class Stepper:
    """One simulation resident on the current HIP device."""

    def __init__(self, cfg: KstConfig):
        self.cfg = cfg
        self.L = lib()
        h = ctypes.c_void_p()
        rc = self.L.kst_create(ctypes.byref(cfg), ctypes.byref(h))
        if rc != 0:
            raise KestrelError(rc, "kst_create failed (configuration or device)")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.L.kst_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != 0:
            buf = ctypes.create_string_buffer(512)
            self.L.kst_last_error(self.h, buf, 512)
            raise KestrelError(rc, buf.value.decode())

    def set_topography(self, b0v: np.ndarray):
        self._b0v = np.ascontiguousarray(b0v, dtype=np.float64)
        self._check(self.L.kst_set_topography(self.h, _ptr(self._b0v)))

    def set_state(self, init, t: float = 0.0, btv: np.ndarray | None = None):
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (init.w, init.rhoHnu, init.rhoHnv, init.Hnpsi)]
        ts = np.ascontiguousarray(init.tile_state, dtype=np.uint8)
        hm = None if init.hnmax0 is None else np.ascontiguousarray(init.hnmax0, dtype=np.float64)
        bt = None if btv is None else np.ascontiguousarray(btv, dtype=np.float64)
        self._check(self.L.kst_set_state(self.h, *[_ptr(a) for a in arrs], _ptr(bt),
                                         ts.ctypes.data_as(_u8p), float(t), _ptr(hm)))
        if getattr(init, "psimax0", None) is not None:
            # a 1D cap also paints the running maximum of the solids fraction
            mx = self.maxima()
            mx[8] = np.ascontiguousarray(init.psimax0, dtype=np.float64)
            self.set_maxima(mx)

    def integrate_to(self, tend: float) -> dict:
        st = KstStats()
        self._check(self.L.kst_integrate_to(self.h, float(tend), ctypes.byref(st)))
        return st.as_dict()

    def step_n(self, n: int) -> dict:
        st = KstStats()
        self._check(self.L.kst_step_n(self.h, int(n), ctypes.byref(st)))
        return st.as_dict()

    def state(self) -> dict:
        c = self.cfg
        out = {k: np.zeros((c.NY, c.NX)) for k in ("w", "rhoHnu", "rhoHnv", "Hnpsi", "Hn")}
        out["btv"] = np.zeros((c.NY + 1, c.NX + 1))
        ts = np.zeros(c.ntiles, dtype=np.uint8)
        t = ctypes.c_double()
        self._check(self.L.kst_get_state(self.h, _ptr(out["w"]), _ptr(out["rhoHnu"]), _ptr(out["rhoHnv"]),
                                         _ptr(out["Hnpsi"]), _ptr(out["Hn"]), _ptr(out["btv"]),
                                         ts.ctypes.data_as(_u8p), ctypes.byref(t)))
        out["tile_state"] = ts
        out["t"] = t.value
        return out

    def tile_bt(self) -> np.ndarray:
        c = self.cfg
        out = np.zeros((c.ntiles, c.ny_per_tile + 1, c.nx_per_tile + 1))
        self._check(self.L.kst_get_tile_bt(self.h, _ptr(out)))
        return out

    def cell_topography(self) -> dict:
        c = self.cfg
        out = {k: np.zeros((c.NY, c.NX)) for k in ("b0c", "btc", "dbdx", "dbdy")}
        self._check(self.L.kst_get_cell_topography(self.h, *[_ptr(out[k]) for k in ("b0c", "btc", "dbdx", "dbdy")]))
        return out

    def set_cell_topography(self, btc, dbdx, dbdy, cells) -> None:
        """Cell bt, db/dx, db/dy (NY x NX) of the cells where `cells` is true,
        as a restart reads them from a solution file (kst_set_cell_topography)."""
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (btc, dbdx, dbdy)]
        m = np.ascontiguousarray(cells, dtype=np.uint8)
        self._check(self.L.kst_set_cell_topography(self.h, *[_ptr(a) for a in arrs],
                                                   m.ctypes.data_as(_u8p)))

    def volume(self) -> np.ndarray:
        """The Volume.txt columns 2-7 (flow volume, total bed volume, total
        mass, bed mass, total solids mass, bed solids mass), on the device."""
        out = np.zeros(6)
        self._check(self.L.kst_volume(self.h, _ptr(out)))
        return out

    def volume_noalloc(self) -> np.ndarray:
        """kst_volume_noalloc: the same sums with the library's fixed device
        scratch (no device allocation in the call)."""
        out = np.zeros(6)
        self._check(self.L.kst_volume_noalloc(self.h, _ptr(out)))
        return out

    def output_fields(self, max_tiles: int | None = None):
        """(tile ids, fields[n, 14, ny_per_tile, nx_per_tile]): the per-cell
        columns 6-19 of the solution text file for the active tiles, in
        active-list order, computed on the device."""
        c = self.cfg
        cap = c.ntiles if max_tiles is None else int(max_tiles)
        f = np.zeros((cap, 14, c.ny_per_tile, c.nx_per_tile))
        t = np.zeros(cap, dtype=np.int32)
        n = ctypes.c_int32(0)
        self._check(self.L.kst_get_output(self.h, cap, _ptr(f), t.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                          ctypes.byref(n)))
        return t[:n.value].copy(), f[:n.value].copy()

    def set_grid_cap(self, tiles: int):
        """Fix the launch grids' active-tile cap (0: follow the active set)."""
        self._check(self.L.kst_set_grid_cap(self.h, int(tiles)))

    def set_redistribution_chunk(self, chunk: int, lds_sort: int = 0):
        """Walk at most `chunk` redistribution candidates at once and sort at
        most `lds_sort` in LDS (0: defaults); results do not depend on them."""
        self._check(self.L.kst_set_redistribution_chunk(self.h, int(chunk), int(lds_sort)))

    def set_activation_candidates(self, cap: int):
        """Walk at most `cap` activation candidates from LDS, more through the
        active list itself (0: default); results do not depend on it."""
        self._check(self.L.kst_set_activation_candidates(self.h, int(cap)))

    def autotune(self, enable: bool = True):
        """Choose the launch shape per active-set bucket from timed trials."""
        self._check(self.L.kst_autotune(self.h, int(enable)))

    def autotune_choices(self) -> list:
        n = 64
        b = np.zeros(n, np.int32); sm = np.zeros(n, np.int32); gr = np.zeros(n, np.int32)
        us = np.zeros((n, 4))
        i32p = ctypes.POINTER(ctypes.c_int32)
        k = self.L.kst_autotune_choices(self.h, n, b.ctypes.data_as(i32p), sm.ctypes.data_as(i32p),
                                        gr.ctypes.data_as(i32p), _ptr(us))
        if k < 0:
            raise KestrelError(k, "kst_autotune_choices")
        names = ("16x16 default grid", "16x16 resident grid", "16x16 twice resident", "16x4")
        return [{"active_tiles_bucket": int(b[q]),
                 "shape": "16x4" if sm[q] else ("16x16" if gr[q] == 0 else f"16x16 grid {int(gr[q])}"),
                 "stage1_us": {names[c]: round(float(us[q, c]), 2) for c in range(4) if us[q, c] > 0}}
                for q in range(k)]

    def profile(self, enable: bool = True):
        self._check(self.L.kst_profile(self.h, int(enable)))

    def profile_read(self) -> dict:
        """{kernel name: (mean ms of the launches that did work, their count)}."""
        ms = (ctypes.c_double * PROFILE_KERNELS)()
        n = (ctypes.c_int64 * PROFILE_KERNELS)()
        self._check(self.L.kst_profile_read(self.h, ms, n))
        return {name: (ms[k], n[k]) for k, name in enumerate(PROFILE_NAMES) if n[k]}

    def set_maxima(self, maxima) -> None:
        """Resume the running maxima from 11 NY x NX arrays (None: leave)."""
        arrs = [None if m is None else np.ascontiguousarray(m, dtype=np.float64) for m in maxima]
        if len(arrs) != 11:
            raise ValueError("maxima: 11 arrays")
        ptrs = (_dp * 11)(*[_ptr(a) for a in arrs])
        self._check(self.L.kst_set_maxima(self.h, ptrs))

    def maxima(self) -> list:
        c = self.cfg
        arrs = [np.zeros((c.NY, c.NX)) for _ in range(11)]
        ptrs = (_dp * 11)(*[_ptr(a) for a in arrs])
        self._check(self.L.kst_get_maxima(self.h, ptrs))
        return arrs


SLOT_FREE, SLOT_RUNNING, SLOT_DONE, SLOT_ERROR = -1, 0, 1, 2


class Batch:
    """Scenarios resident together and advanced by one launch per kernel
    (kst_batch_*).  Attached Steppers are advanced only through the batch."""

    def __init__(self, max_slots: int):
        self.L = lib()
        b = ctypes.c_void_p()
        rc = self.L.kst_batch_create(int(max_slots), ctypes.byref(b))
        if rc != 0:
            raise KestrelError(rc, "kst_batch_create failed")
        self.b = b
        self.max_slots = int(max_slots)
        self.members: dict[int, Stepper] = {}

    def _check(self, rc):
        if rc != 0:
            buf = ctypes.create_string_buffer(512)
            self.L.kst_batch_last_error(self.b, buf, 512)
            raise KestrelError(rc, buf.value.decode())

    def attach(self, stepper: Stepper) -> int:
        slot = ctypes.c_int()
        self._check(self.L.kst_batch_attach(self.b, stepper.h, ctypes.byref(slot)))
        self.members[slot.value] = stepper
        return slot.value

    def detach(self, slot: int):
        self._check(self.L.kst_batch_detach(self.b, int(slot)))
        self.members.pop(slot, None)

    def start(self, slot: int, tend: float = 0.0, nsteps: int = 0):
        self._check(self.L.kst_batch_start(self.b, int(slot), float(tend), int(nsteps)))

    def run(self, max_iters: int = 1 << 40) -> list:
        st = (ctypes.c_int * self.max_slots)()
        self._check(self.L.kst_batch_run(self.b, int(max_iters), st))
        return list(st)

    def stats(self, slot: int) -> dict:
        out = KstStats()
        rc = self.L.kst_batch_stats(self.b, int(slot), ctypes.byref(out))
        if rc != 0:
            self.members[slot]._check(rc)
        return out.as_dict()

    def close(self):
        if getattr(self, "b", None):
            self.L.kst_batch_destroy(self.b)
            self.b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
