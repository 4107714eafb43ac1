"""ctypes binding of libmcx.so (include/mcx.h).

The library is built in-tree (mcmax_amd/libmcx.so) by ``mcmax_amd.build``.
There is no CPU fallback: if the library or a GPU is missing, loading fails
with an error.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmcx.so")

MCX_OK, MCX_EINVAL, MCX_EDEVICE, MCX_ENOMEM, MCX_ESTATE, MCX_EGEOM, MCX_ETIME = 0, -1, -2, -3, -4, -5, -6
TMAX = 10000
NANGLE = 30

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


class Tables(C.Structure):
    _fields_ = [
        ("nlam", C.c_int32), ("ng", C.c_int32), ("tmax", C.c_int32),
        ("lam", _dp), ("nu", _dp), ("BB", _dp), ("Kabs", _dp), ("Ksca", _dp),
        ("F", _dp), ("IF11", _dp), ("IF12", _dp), ("Kp", _dp), ("Fstar", _dp),
        ("Lstar", C.c_double), ("cos2phi", _dp), ("sin2phi", _dp),
        ("scat_how", C.c_int32), ("Kext", _dp), ("gsym", _dp),
    ]


class Grid(C.Structure):
    _fields_ = [
        ("nr", C.c_int32), ("nt", C.c_int32), ("R", _dp), ("Theta", _dp),
        ("dens", _dp), ("V", _dp), ("w", _dp), ("Rstar_cm", C.c_double), ("Tstar", C.c_double),
        ("thet", _dp),
    ]


class Run(C.Structure):
    _fields_ = [
        ("nphot", C.c_int64), ("seed", C.c_uint64), ("maxinteract", C.c_int32),
        ("nlanes", C.c_int32), ("stage0", C.c_int64), ("randomwalk", C.c_int32),
        ("factRW", C.c_double), ("coop_lanes", C.c_int32), ("stage_ratio", C.c_int32),
        ("precision", C.c_int32), ("max_ms", C.c_double), ("epoch_deposits", C.c_int64),
    ]


class Tallies(C.Structure):
    _fields_ = [
        ("EJv", _dp), ("EJv2", _dp), ("Eabs", _dp), ("Tbw", _dp), ("Ni", _ip),
        ("spec", _dp), ("scatspec", _dp), ("ispec", _ip), ("iscatspec", _ip),
        ("nemit", C.c_int64), ("ninteract", C.c_int64), ("nabs", C.c_int64),
        ("nscat", C.c_int64), ("nescape", C.c_int64), ("nstarhit", C.c_int64),
        ("nsteps", C.c_int64), ("nraar", C.c_int64), ("nmaxinteract", C.c_int64),
        ("ngeomfail", C.c_int64), ("tautot", C.c_double), ("Eescape", C.c_double),
        ("nrw", C.c_int64), ("noverflow", C.c_int64), ("thick", _ip),
        ("maxsteps", C.c_int64), ("t_done", C.c_double * 5), ("nvisit_overflow", C.c_int64), ("ntimeout", C.c_int64),
        ("ntallyatomic", C.c_int64),
    ]


class Temps(C.Structure):
    _fields_ = [("T", _dp), ("dT", _dp), ("dEJv", _dp)]


class Obs(C.Structure):
    _fields_ = [
        ("angle", C.c_double), ("nphi", C.c_int32), ("nj0", C.c_int32), ("nphot", C.c_int32),
        ("nphotstar", C.c_int32), ("nexits", C.c_int32), ("usepol", C.c_int32), ("tracestar", C.c_int32),
        ("traceemis", C.c_int32), ("tracescat", C.c_int32), ("pfstop", C.c_double), ("seed", C.c_uint64),
    ]


class Struct(C.Structure):
    _fields_ = [
        ("R_av", _dp), ("theta_av", _dp), ("Mstar", C.c_double), ("shscale", _dp), ("gshscale", _dp),
        ("settle", _ip), ("gascoupled", _ip), ("mpstr", C.c_int32), ("FradZ", _dp),
    ]


class Density(C.Structure):
    _fields_ = [("dens", _dp), ("dens0", _dp), ("gasdens", _dp), ("w", _dp), ("w0", _dp)]


class CellState(C.Structure):
    _fields_ = [("T", _dp), ("dT", _dp), ("dens", _dp), ("ddens", _dp), ("EJv", _dp), ("dEJv", _dp)]


_fp = C.POINTER(C.c_float)


class PartSpec(C.Structure):
    _fields_ = [("nm", C.c_int32), ("ns", C.c_int32), ("nf", C.c_int32), ("pad", C.c_int32),
                ("amin", C.c_double), ("amax", C.c_double), ("apow", C.c_double), ("fmax", C.c_double),
                ("rscale", C.c_double), ("mscale", C.c_double), ("e1", _fp), ("e2", _fp), ("frac", _fp),
                ("rho", _fp)]


class PartOut(C.Structure):
    _fields_ = [("Kabs", _dp), ("Ksca", _dp), ("Kext", _dp), ("F", _dp), ("rho", _dp)]


class PartStats(C.Structure):
    _fields_ = [("ndhs", C.c_int64), ("nhom", C.c_int64), ("nerr", C.c_int64), ("dhs_terms", C.c_int64),
                ("hom_terms", C.c_int64), ("scratch_bytes", C.c_int64), ("ms_total", C.c_double)]


class ImageOpts(C.Structure):
    _fields_ = [("npix", C.c_int32), ("nint", C.c_int32), ("Rmax", C.c_double), ("pa_deg", C.c_double),
                ("seed", C.c_uint64)]


EXPORTS = (
    "mcx_create", "mcx_destroy", "mcx_last_error", "mcx_set_tables", "mcx_set_model",
    "mcx_transport", "mcx_get_tallies", "mcx_temperature", "mcx_last_timing", "mcx_sed",
    "mcx_set_temperature", "mcx_diffuse", "mcx_temp_average", "mcx_disk_structure", "mcx_compare_cells",
    "mcx_compute_parts", "mcx_image", "mcx_last_image_timing", "mcx_batch_transport",
    "mcx_batch_transport_blocks", "mcx_mlp_create", "mcx_mlp_destroy", "mcx_mlp_last_error", "mcx_mlp_set_params", "mcx_mlp_get_params",
    "mcx_mlp_forward", "mcx_mlp_gradient", "mcx_mlp_set_data", "mcx_mlp_train", "mcx_mlp_last_timing",
    "mcx_mlp_forward_noalloc",
)

MLP_NIN, MLP_WIDTH, MLP_DEPTH, MLP_NPARAMS = 8, 256, 4, 199937
MLP_NOALLOC_MAXROWS = 65536


class MlpTrainOpts(C.Structure):
    _fields_ = [("batch", C.c_int32), ("nsteps", C.c_int32), ("lr", C.c_float), ("beta1", C.c_float),
                ("beta2", C.c_float), ("eps", C.c_float), ("seed", C.c_uint64)]

_lib = None


def load():
    """Load libmcx.so; raises RuntimeError when it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`")
    lib = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    lib.mcx_create.argtypes = [C.POINTER(vp), C.c_int]
    lib.mcx_create.restype = C.c_int
    lib.mcx_destroy.argtypes = [vp]
    lib.mcx_destroy.restype = None
    lib.mcx_last_error.argtypes = [vp]
    lib.mcx_last_error.restype = C.c_char_p
    lib.mcx_set_tables.argtypes = [vp, C.POINTER(Tables)]
    lib.mcx_set_model.argtypes = [vp, C.POINTER(Grid)]
    lib.mcx_transport.argtypes = [vp, C.POINTER(Run)]
    lib.mcx_batch_transport.argtypes = [C.POINTER(vp), C.c_int32, C.POINTER(Run), C.c_int32, C.c_int64, C.c_int32,
                                        C.POINTER(C.c_int32)]
    lib.mcx_batch_transport_blocks.argtypes = [C.POINTER(vp), C.c_int32, C.POINTER(Run), C.POINTER(C.c_int32),
                                               C.c_int64, C.c_int32, C.POINTER(C.c_int32)]
    lib.mcx_get_tallies.argtypes = [vp, C.POINTER(Tallies)]
    lib.mcx_temperature.argtypes = [vp, C.POINTER(Temps)]
    lib.mcx_last_timing.argtypes = [vp, _dp, _dp]
    lib.mcx_sed.argtypes = [vp, C.POINTER(Obs), _dp, _dp, C.c_int32, _dp, _dp, _dp, _dp]
    lib.mcx_set_temperature.argtypes = [vp, _dp, _dp, _ip]
    lib.mcx_diffuse.argtypes = [vp, C.c_int32, C.c_double, C.c_double, _dp, _dp, _dp, _ip, _dp,
                                C.POINTER(C.c_int32)]
    lib.mcx_temp_average.argtypes = [vp, C.c_double, C.c_int32, _dp]
    lib.mcx_disk_structure.argtypes = [vp, C.POINTER(Struct), _dp, C.POINTER(Density), C.POINTER(C.c_int64)]
    lib.mcx_compare_cells.argtypes = [vp, C.POINTER(CellState), C.POINTER(CellState), _dp, _dp, _dp]
    lib.mcx_compute_parts.argtypes = [vp, C.c_int32, _dp, C.c_int32, C.POINTER(PartSpec), C.POINTER(PartOut),
                                      C.POINTER(PartStats)]
    lib.mcx_last_image_timing.argtypes = [vp, _dp, _dp]
    lib.mcx_image.argtypes = [vp, C.POINTER(Obs), _dp, _dp, C.c_double, C.POINTER(ImageOpts), _dp, _dp, _dp, _dp,
                              C.POINTER(C.c_int32), C.c_int32, _dp, _dp]
    lib.mcx_mlp_create.argtypes = [C.POINTER(vp), C.c_int, C.c_int32, C.c_int32, C.c_int32]
    lib.mcx_mlp_destroy.argtypes = [vp]
    lib.mcx_mlp_destroy.restype = None
    lib.mcx_mlp_last_error.argtypes = [vp]
    lib.mcx_mlp_last_error.restype = C.c_char_p
    lib.mcx_mlp_set_params.argtypes = [vp, _fp]
    lib.mcx_mlp_get_params.argtypes = [vp, _fp]
    lib.mcx_mlp_forward.argtypes = [vp, C.c_int64, _fp, _fp]
    lib.mcx_mlp_forward_noalloc.argtypes = [vp, C.c_int64, _fp, _fp]
    lib.mcx_mlp_gradient.argtypes = [vp, C.c_int64, _fp, _fp, _dp, _fp]
    lib.mcx_mlp_set_data.argtypes = [vp, C.c_int64, _fp, _fp]
    lib.mcx_mlp_train.argtypes = [vp, C.POINTER(MlpTrainOpts), _fp]
    lib.mcx_mlp_last_timing.argtypes = [vp, _dp, C.POINTER(C.c_int64), _dp]
    for n in ("mcx_mlp_create", "mcx_mlp_set_params", "mcx_mlp_get_params", "mcx_mlp_forward",
              "mcx_mlp_forward_noalloc", "mcx_mlp_gradient",
              "mcx_mlp_set_data", "mcx_mlp_train", "mcx_mlp_last_timing"):
        getattr(lib, n).restype = C.c_int
    for n in ("mcx_set_tables", "mcx_set_model", "mcx_transport", "mcx_get_tallies",
              "mcx_temperature", "mcx_last_timing", "mcx_sed", "mcx_set_temperature", "mcx_diffuse",
              "mcx_temp_average", "mcx_disk_structure", "mcx_compare_cells", "mcx_compute_parts", "mcx_image",
              "mcx_last_image_timing", "mcx_batch_transport", "mcx_batch_transport_blocks"):
        getattr(lib, n).restype = C.c_int
    _lib = lib
    return lib


def ptr(a):
    """Pointer to a C-contiguous numpy array (None -> NULL)."""
    if a is None:
        return None
    if a.dtype == np.float64:
        return a.ctypes.data_as(_dp)
    if a.dtype == np.int32:
        return a.ctypes.data_as(_ip)
    if a.dtype == np.float32:
        return a.ctypes.data_as(_fp)
    raise TypeError(a.dtype)
