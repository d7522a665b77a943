"""ctypes mirrors of the C ABI structs declared in include/nmc.h.

`RunParams` and `TissueParams` are byte-compatible with the reference's
structs (NMC/NMCCode/definitions.h:186-218); the remaining structs carry the
settings NMC fixed at compile time (definitions.h:16-82) and the results.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

NTISS = 19            # definitions.h:57
STRLEN = 32           # definitions.h:58
THRESHOLD = 1e-4      # definitions.h:66
CHANCE = 0.1          # definitions.h:67
EXT_VERSION = 1

SIM_MCSUB = 0         # definitions.h:153-155 (MC_SIM_TYPE)
SIM_MCXYZ = 1
SIM_MCXYZ_TT = 2

DET_REFLECTANCE = 0   # definitions.h:167-171
DET_TRANSMITTANCE = 1

FLAG_NO_FLUENCE = 1
FLAG_COUNT_VISITS = 2

OK = 0
ERR_ARG = -1
ERR_STATE = -2
ERR_DEVICE = -3
ERR_NOMEM = -4
ERR_UNSUPPORTED = -5


class RunParams(C.Structure):
    """definitions.h:186-209 (172 bytes)."""

    _fields_ = [
        ("myname", C.c_char * STRLEN),
        ("mcflag", C.c_int32), ("launchflag", C.c_int32), ("boundaryflag", C.c_int32),
        ("speckleflag", C.c_int32), ("semiflag", C.c_int32),
        ("xfocus", C.c_float), ("yfocus", C.c_float), ("zfocus", C.c_float),
        ("ux0", C.c_float), ("uy0", C.c_float), ("uz0", C.c_float),
        ("radius", C.c_float), ("det_radius", C.c_float), ("xd", C.c_float),
        ("waist", C.c_float),
        ("zsurf", C.c_float),
        ("n1", C.c_float), ("n2", C.c_float),
        ("na", C.c_float), ("lambda_", C.c_float), ("lc", C.c_float),
        ("det_state", C.c_int32),
        ("ph_quant", C.c_int32),
        ("dx", C.c_float), ("dy", C.c_float), ("dz", C.c_float),
        ("Nx", C.c_int32), ("Ny", C.c_int32), ("Nz", C.c_int32), ("Nt", C.c_int32),
        ("xs", C.c_float), ("ys", C.c_float), ("zs", C.c_float),
        ("time_min", C.c_float),
        ("mckernelflag", C.c_int32),
    ]


class TissueParams(C.Structure):
    """definitions.h:211-218 (608 bytes); index 0 is unused, types are 1..Nt."""

    _fields_ = [(name, C.c_float * NTISS)
                for name in ("muav", "musv", "gv", "gf", "af", "gb", "ab", "CC")]


class ExtConfig(C.Structure):
    """nmc_ext_config: run-time replacements for definitions.h:16-82."""

    _fields_ = [
        ("struct_size", C.c_uint32), ("version", C.c_uint32),
        ("seed", C.c_uint64), ("stream_id", C.c_uint32), ("flags", C.c_uint32),
        ("threshold", C.c_float), ("chance", C.c_float), ("max_events", C.c_uint32),
        ("nr", C.c_int32), ("nz", C.c_int32), ("na", C.c_int32),
        ("dr", C.c_float), ("dz", C.c_float),
        ("reserved", C.c_int32 * 8),
    ]

    @classmethod
    def default(cls, **kw) -> "ExtConfig":
        e = cls()
        e.struct_size = C.sizeof(cls)
        e.version = EXT_VERSION
        e.seed = 1
        e.threshold = THRESHOLD
        e.chance = CHANCE
        e.max_events = 100_000_000
        e.nr, e.nz, e.na = 100, 100, 30
        e.dr, e.dz = 0.01, 0.01
        for k, v in kw.items():
            if not hasattr(e, k):
                raise AttributeError(f"nmc_ext_config has no field {k!r}")
            setattr(e, k, v)
        return e


TOTAL_FLOAT_FIELDS = ("specular", "rd", "rd_outside", "absorbed", "absorbed_outside", "tt",
                      "lateral", "top", "roulette_gain", "roulette_loss", "capped")
TOTAL_INT_FIELDS = ("photons", "detections", "surface_hits", "escapes_bottom", "escapes_side",
                    "escapes_top", "roulette_kills", "roulette_survivals", "capped_photons",
                    "substeps_same", "substeps_cross", "interactions")


class Totals(C.Structure):
    _fields_ = ([(n, C.c_double) for n in TOTAL_FLOAT_FIELDS]
                + [(n, C.c_uint64) for n in TOTAL_INT_FIELDS])

    def as_dict(self) -> dict:
        return {n: getattr(self, n) for n in TOTAL_FLOAT_FIELDS + TOTAL_INT_FIELDS}


class Results(C.Structure):
    _fields_ = [
        ("fluence", C.POINTER(C.c_double)),
        ("ryx", C.POINTER(C.c_double)),
        ("det2d", C.POINTER(C.c_uint32)),
        ("rd_ra", C.POINTER(C.c_double)),
        ("a_rz", C.POINTER(C.c_double)),
        ("totals", Totals),
    ]


class Tuning(C.Structure):
    """nmc_tuning: launch shape (block, blocks/CU, refill batch, queue chunk), the
    mcxyz LDS fluence tile box (x0, y0, z0, nx, ny, nz) and `mode` (mcsub tally
    placement 0 auto / 1 LDS / 2 global; mcxyz tile element 0 FP64 / 1 FP32)."""

    _fields_ = [("block", C.c_int32), ("blocks_per_cu", C.c_int32), ("refill_min", C.c_int32),
                ("chunk", C.c_int32), ("tile", C.c_int32 * 6), ("mode", C.c_int32),
                ("split", C.c_int32), ("rec_cap_pct", C.c_int32), ("rec_fuse", C.c_int32),
                ("reserved", C.c_int32 * 2)]

    def as_dict(self) -> dict:
        return {"block": self.block, "blocks_per_cu": self.blocks_per_cu,
                "refill_min": self.refill_min, "chunk": self.chunk, "tile": list(self.tile),
                "mode": self.mode, "split": self.split, "rec_cap_pct": self.rec_cap_pct,
                "rec_fuse": self.rec_fuse}

    @classmethod
    def make(cls, block=0, blocks_per_cu=0, refill_min=0, chunk=0, tile=None,
             mode=0, split=0, rec_cap_pct=0, rec_fuse=0) -> "Tuning":
        t = cls()
        t.block, t.blocks_per_cu, t.refill_min, t.chunk = block, blocks_per_cu, refill_min, chunk
        t.mode, t.split, t.rec_cap_pct, t.rec_fuse = mode, split, rec_cap_pct, rec_fuse
        for i, v in enumerate(tile or (0,) * 6):
            t.tile[i] = v
        return t


class WorkItem(C.Structure):
    """nmc_work_item: one medium of a batched mcsub launch."""

    _fields_ = [("mua", C.c_float), ("mus", C.c_float), ("g", C.c_float), ("n1", C.c_float),
                ("n2", C.c_float), ("stream_id", C.c_uint32), ("n_photons", C.c_uint64)]


SIM_LAYERED = 3
MAX_LAYERS = 8


class Layer(C.Structure):
    """nmc_layer: one layer of a stack (thickness <= 0 on the last layer:
    semi-infinite)."""

    _fields_ = [("thickness", C.c_float), ("mua", C.c_float), ("mus", C.c_float), ("n", C.c_float),
                ("g", C.c_float), ("gf", C.c_float), ("af", C.c_float), ("gb", C.c_float),
                ("ab", C.c_float), ("cc", C.c_float), ("phase", C.c_int32),
                ("reserved", C.c_int32)]


class Stack(C.Structure):
    """nmc_stack: a layered medium with its Philox stream and photon budget."""

    _fields_ = [("n_layers", C.c_int32), ("n_above", C.c_float), ("n_below", C.c_float),
                ("stream_id", C.c_uint32), ("n_photons", C.c_uint64),
                ("layer", Layer * MAX_LAYERS)]


class LayeredConfig(C.Structure):
    """nmc_layered_config."""

    _fields_ = [("struct_size", C.c_uint32), ("version", C.c_uint32), ("seed", C.c_uint64),
                ("nr", C.c_int32), ("dr", C.c_float), ("threshold", C.c_float),
                ("chance", C.c_float), ("max_events", C.c_uint32), ("records", C.c_int32),
                ("record_capacity", C.c_uint64), ("absorbed_grad", C.c_int32),
                ("reserved", C.c_int32 * 7)]

    @classmethod
    def default(cls, **kw) -> "LayeredConfig":
        c = cls()
        c.struct_size = C.sizeof(cls)
        c.version = EXT_VERSION
        c.seed = 1
        c.nr, c.dr = 200, 0.005
        c.threshold, c.chance = THRESHOLD, CHANCE
        c.max_events = 100_000_000
        c.records = 1
        for k, v in kw.items():
            if not hasattr(c, k):
                raise AttributeError(f"nmc_layered_config has no field {k!r}")
            setattr(c, k, v)
        return c


assert C.sizeof(RunParams) == 172
assert C.sizeof(TissueParams) == 608
assert C.sizeof(WorkItem) == 32 and C.sizeof(Tuning) == 64
assert C.sizeof(Layer) == 48 and C.sizeof(Stack) == 408 and C.sizeof(LayeredConfig) == 80


def dptr(a: np.ndarray | None):
    if a is None:
        return C.POINTER(C.c_double)()
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_double))


def uptr(a: np.ndarray | None):
    if a is None:
        return C.POINTER(C.c_uint32)()
    assert a.dtype == np.uint32 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_uint32))


class TallyArrays:
    """Host arrays for one tally set plus the Results struct pointing at them."""

    def __init__(self, rp: RunParams, ext: ExtConfig, sim: int, fluence: bool = True):
        self.fluence = self.ryx = self.det2d = self.rd_ra = self.a_rz = None
        if sim in (SIM_MCXYZ, SIM_MCXYZ_TT):
            if fluence:
                self.fluence = np.zeros(rp.Nz * rp.Nx * rp.Ny, np.float64)
            self.ryx = np.zeros(rp.Nx * rp.Ny, np.float64)
            self.det2d = np.zeros(rp.Nx * rp.Ny, np.uint32)
        else:
            self.rd_ra = np.zeros((ext.nr + 1) * ext.na, np.float64)
            if fluence:
                self.a_rz = np.zeros((ext.nr + 1) * (ext.nz + 1), np.float64)
        self.res = Results()
        self.res.fluence = dptr(self.fluence)
        self.res.ryx = dptr(self.ryx)
        self.res.det2d = uptr(self.det2d)
        self.res.rd_ra = dptr(self.rd_ra)
        self.res.a_rz = dptr(self.a_rz)

    @property
    def totals(self) -> dict:
        return self.res.totals.as_dict()
