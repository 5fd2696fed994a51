"""ctypes mirror of the POD types declared in include/erp.h."""
import ctypes

import numpy as np

ERP_OK = 0
ERP_E_INVALID = -1
ERP_E_HIP = -2
ERP_E_NOMEM = -3
ERP_E_UNSUPPORTED = -4
ERP_E_NOTREADY = -5
ERP_E_INTERNAL = -6

ERP_MODE_CONVOLUTION = 0x01
ERP_MODE_CONV_LINEAR = 0x02
ERP_MODE_SPECULAR = 0x04
ERP_MODE_LOSSLESS = 0x08
ERP_MODE_INCOHERENT = 0x10
ERP_MODE_DISABLE_SURFACE = 0x20

ERP_RUN_DETERMINISTIC = 0x01
ERP_RUN_PHASE_F32 = 0x02
ERP_RUN_PHASE_DF = 0x04
ERP_RUN_RAW_SURFACE = 0x08
ERP_RUN_ACC_F32 = 0x10
ERP_RUN_ACC_COMP = 0x20

ERP_TUNE_SHOTS_PER_BATCH = 1
ERP_TUNE_PAIRS_PER_LAUNCH = 2
ERP_TUNE_TRACK_ROWS = 3
ERP_TUNE_DIRECT_LAUNCH = 4
ERP_TUNE_WINDOW_SHIFT = 5
ERP_TUNE_SINGLE_SHOT = 6

ERP_TERRAIN_GAUSSIAN = 0
ERP_TERRAIN_FBM = 1

_PARAM_FLOATS = ("lam", "k", "rng_res", "P", "fs", "ks", "eps_1", "eps_2", "c",
                 "c_1", "c_2", "nu_1", "nu_2", "alpha1", "alpha2", "aperture",
                 "smpl", "B", "f0", "dr")


class ErpParams(ctypes.Structure):
    _fields_ = ([(n, ctypes.c_float) for n in _PARAM_FLOATS] +
                [("nr", ctypes.c_int32), ("rerad_funct", ctypes.c_int32),
                 ("pol", ctypes.c_int32), ("mode", ctypes.c_uint32),
                 ("cos_ap", ctypes.c_double), ("seed", ctypes.c_uint64),
                 ("eps_3", ctypes.c_float), ("z_int", ctypes.c_float)])


class ErpShot(ctypes.Structure):
    _fields_ = [("sx", ctypes.c_double), ("sy", ctypes.c_double),
                ("sz", ctypes.c_double), ("snx", ctypes.c_float),
                ("sny", ctypes.c_float), ("snz", ctypes.c_float),
                ("rst", ctypes.c_float), ("grefl_lin", ctypes.c_float),
                ("grefr_lin", ctypes.c_float)]


class ErpTrack(ctypes.Structure):
    """erp_track: one job of erp_run_tracks (pointers as erp_run_dev takes them)."""
    _fields_ = [("par", ctypes.POINTER(ErpParams)), ("mesh", ctypes.c_void_p),
                ("targets", ctypes.c_void_p), ("shots", ctypes.c_void_p),
                ("ns", ctypes.c_int64), ("out_dev", ctypes.c_void_p),
                ("nvalid_dev", ctypes.c_void_p)]


class ErpTerrain(ctypes.Structure):
    """erp_terrain: a synthetic terrain for erp_mesh_from_terrain."""
    _fields_ = [("kind", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("nx", ctypes.c_int64), ("ny", ctypes.c_int64),
                ("fs", ctypes.c_double), ("ox", ctypes.c_double), ("oy", ctypes.c_double),
                ("rms", ctypes.c_double), ("corr_len", ctypes.c_double),
                ("hurst", ctypes.c_double), ("seed", ctypes.c_uint64),
                ("noise", ctypes.c_void_p)]


class ErpPrism(ctypes.Structure):
    _fields_ = [(n, ctypes.c_float) for n in
                ("alpha", "xmin", "ymin", "zmin", "xmax", "ymax", "zmax")]


SHOT_DTYPE = np.dtype([("sx", "<f8"), ("sy", "<f8"), ("sz", "<f8"),
                       ("snx", "<f4"), ("sny", "<f4"), ("snz", "<f4"),
                       ("rst", "<f4"), ("grefl_lin", "<f4"), ("grefr_lin", "<f4")])
PRISM_DTYPE = np.dtype([(n, "<f4") for n in
                        ("alpha", "xmin", "ymin", "zmin", "xmax", "ymax", "zmax")])

assert SHOT_DTYPE.itemsize == ctypes.sizeof(ErpShot) == 48
assert PRISM_DTYPE.itemsize == ctypes.sizeof(ErpPrism) == 28
assert ctypes.sizeof(ErpParams) == 120
