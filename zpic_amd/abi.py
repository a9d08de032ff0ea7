"""ctypes mirror of the em2d C structs (include/*.h).

The layouts follow the reference headers field by field
(em2d/particles.h:29-132, em2d/emf.h:30-120, em2d/current.h:30-70,
em2d/simulation.h:13-29, em2d/zpic.h:18-22).  This library appends one
trailing `dev` pointer to t_species, t_emf and t_current; `make_abi(False)`
builds the reference layout without it (used to drive a library built from
the reference sources).
"""
from __future__ import annotations

import ctypes as C
from types import SimpleNamespace

import numpy as np

MAX_SPNAME_LEN = 32

# enum values (em2d/particles.h:43-49, em2d/current.h:19-23, em2d/emf.h:20-24,126-129)
DENSITY = {"uniform": 0, "empty": 1, "step": 2, "slab": 3, "custom": 4}
SMOOTH = {"none": 0, "binomial": 1, "compensated": 2}
FLD_TYPE = {"none": 0, "uniform": 1, "custom": 2}
LASER_TYPE = {"plane": 0, "gaussian": 1}

PART_DTYPE = np.dtype([("ix", "<i4"), ("iy", "<i4"), ("x", "<f4"), ("y", "<f4"),
                       ("ux", "<f4"), ("uy", "<f4"), ("uz", "<f4")])


class float3(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("z", C.c_float)]


class t_part(C.Structure):
    _fields_ = [("ix", C.c_int), ("iy", C.c_int), ("x", C.c_float), ("y", C.c_float),
                ("ux", C.c_float), ("uy", C.c_float), ("uz", C.c_float)]


DENSITY_FN = C.CFUNCTYPE(C.c_float, C.c_float, C.c_void_p)
# float3 (*)(int, float, int, float, void*): only C code can implement it
FIELD_FN = C.CFUNCTYPE(float3, C.c_int, C.c_float, C.c_int, C.c_float, C.c_void_p)
# what zpic_field_callback calls: writes the 3 components through a pointer
FIELD_OUT_FN = C.CFUNCTYPE(None, C.c_int, C.c_float, C.c_int, C.c_float, C.POINTER(C.c_float), C.c_void_p)


class zpic_field_cb(C.Structure):
    _fields_ = [("fn", FIELD_OUT_FN), ("user", C.c_void_p)]


def field_callback(adapter_lib: C.CDLL, fn, keep: list):
    """(function pointer, data pointer) for a Python field function
    fn(ix, dx, iy, dy) -> 3 floats, through the library's
    zpic_field_callback adapter.  `keep` holds the ctypes objects alive."""

    def body(i, dx, j, dy, out, user):
        v = fn(i, dx, j, dy)
        out[0], out[1], out[2] = float(v[0]), float(v[1]), float(v[2])

    cfn = FIELD_OUT_FN(body)
    cb = zpic_field_cb(cfn, None)
    keep.extend([cfn, cb])
    return C.cast(adapter_lib.zpic_field_callback, FIELD_FN), C.cast(C.pointer(cb), C.c_void_p)


class t_density(C.Structure):
    _fields_ = [("n", C.c_float), ("type", C.c_int), ("start", C.c_float), ("end", C.c_float),
                ("custom_x", DENSITY_FN), ("custom_data_x", C.c_void_p),
                ("custom_y", DENSITY_FN), ("custom_data_y", C.c_void_p),
                ("custom_x_total_part", C.c_ulong), ("custom_x_total_q", C.c_double)]


class t_smooth(C.Structure):
    _fields_ = [("xtype", C.c_int), ("ytype", C.c_int), ("xlevel", C.c_int), ("ylevel", C.c_int)]


class t_emf_ext_fld(C.Structure):
    _fields_ = [("E_type", C.c_int), ("B_type", C.c_int), ("E_0", float3), ("B_0", float3),
                ("E_custom", FIELD_FN), ("B_custom", FIELD_FN),
                ("E_custom_data", C.c_void_p), ("B_custom_data", C.c_void_p),
                ("E_part_buf", C.POINTER(float3)), ("B_part_buf", C.POINTER(float3))]


class t_emf_init_fld(C.Structure):
    _fields_ = [("E_type", C.c_int), ("B_type", C.c_int), ("E_0", float3), ("B_0", float3),
                ("E_custom", FIELD_FN), ("B_custom", FIELD_FN),
                ("E_custom_data", C.c_void_p), ("B_custom_data", C.c_void_p)]


class t_emf_laser(C.Structure):
    _fields_ = [("type", C.c_int), ("start", C.c_float), ("fwhm", C.c_float), ("rise", C.c_float),
                ("flat", C.c_float), ("fall", C.c_float), ("a0", C.c_float), ("omega0", C.c_float),
                ("polarization", C.c_float), ("W0", C.c_float), ("focus", C.c_float),
                ("axis", C.c_float)]


_ABI_CACHE: dict = {}


def make_abi(with_dev: bool = True) -> SimpleNamespace:
    """Struct classes for one ABI flavour (one set of classes per flavour).

    with_dev=True is this library's layout (trailing `dev` pointer);
    with_dev=False is the reference's layout.
    """
    if with_dev in _ABI_CACHE:
        return _ABI_CACHE[with_dev]
    tail = [("dev", C.c_void_p)] if with_dev else []

    class t_species(C.Structure):
        _fields_ = [("name", C.c_char * (MAX_SPNAME_LEN + 1)), ("part", C.POINTER(t_part)),
                    ("np", C.c_int), ("np_max", C.c_int), ("m_q", C.c_float),
                    ("energy", C.c_double), ("q", C.c_float), ("ppc", C.c_int * 2),
                    ("density", t_density), ("ufl", C.c_float * 3), ("uth", C.c_float * 3),
                    ("nx", C.c_int * 2), ("dx", C.c_float * 2), ("box", C.c_float * 2),
                    ("dt", C.c_float), ("iter", C.c_int), ("moving_window", C.c_int),
                    ("n_move", C.c_int), ("n_sort", C.c_int)] + tail

    class t_current(C.Structure):
        _fields_ = [("J", C.POINTER(float3)), ("J_buf", C.POINTER(float3)), ("nx", C.c_int * 2),
                    ("nrow", C.c_int), ("gc", (C.c_int * 2) * 2), ("box", C.c_float * 2),
                    ("dx", C.c_float * 2), ("smooth", t_smooth), ("dt", C.c_float),
                    ("iter", C.c_int), ("moving_window", C.c_int)] + tail

    class t_emf(C.Structure):
        _fields_ = [("E", C.POINTER(float3)), ("B", C.POINTER(float3)),
                    ("E_buf", C.POINTER(float3)), ("B_buf", C.POINTER(float3)),
                    ("E_part", C.POINTER(float3)), ("B_part", C.POINTER(float3)),
                    ("nx", C.c_int * 2), ("nrow", C.c_int), ("gc", (C.c_int * 2) * 2),
                    ("box", C.c_float * 2), ("dx", C.c_float * 2), ("dt", C.c_float),
                    ("iter", C.c_int), ("moving_window", C.c_int), ("n_move", C.c_int),
                    ("ext_fld", t_emf_ext_fld)] + tail

    class t_simulation(C.Structure):
        _fields_ = [("dt", C.c_float), ("tmax", C.c_float), ("ndump", C.c_int),
                    ("n_species", C.c_int), ("species", C.POINTER(t_species)),
                    ("emf", t_emf), ("current", t_current), ("moving_window", C.c_int)]

    ns = SimpleNamespace(t_species=t_species, t_current=t_current, t_emf=t_emf,
                         t_simulation=t_simulation, with_dev=with_dev)
    _ABI_CACHE[with_dev] = ns
    return ns


def bind_api(lib: C.CDLL, abi: SimpleNamespace) -> C.CDLL:
    """Declare argument/return types of the em2d C API on `lib`."""
    P = C.POINTER
    sp, em, cu, sim = P(abi.t_species), P(abi.t_emf), P(abi.t_current), P(abi.t_simulation)
    sig = {
        "set_rand_seed": (None, [C.c_uint32, C.c_uint32]),
        "rand_norm": (C.c_double, []),
        "spec_new": (None, [sp, C.c_char_p, C.c_float, P(C.c_int), P(C.c_float), P(C.c_float),
                            P(C.c_int), P(C.c_float), C.c_float, P(t_density)]),
        "spec_delete": (None, [sp]),
        "spec_grow_buffer": (None, [sp, C.c_int]),
        "spec_advance": (None, [sp, em, cu]),
        "spec_deposit_charge": (None, [sp, P(C.c_float)]),
        "spec_deposit_pha": (None, [sp, C.c_int, P(C.c_int), P(C.c_float * 2), P(C.c_float)]),
        "spec_npush": (C.c_uint64, []),
        "spec_time": (C.c_double, []),
        "spec_perf": (C.c_double, []),
        "current_zero": (None, [cu]),
        "current_update": (None, [cu]),
        "emf_advance": (None, [em, cu]),
        "emf_get_energy": (None, [em, P(C.c_double)]),
        "emf_init_fld": (None, [em, P(t_emf_init_fld)]),
        "emf_set_ext_fld": (None, [em, P(t_emf_ext_fld)]),
        "emf_report": (None, [em, C.c_byte, C.c_int]),
        "current_report": (None, [cu, C.c_int]),
        "spec_report": (None, [sp, C.c_int, P(C.c_int), P(C.c_float * 2)]),
        "emf_time": (C.c_double, []),
        "sim_new": (None, [sim, P(C.c_int), P(C.c_float), C.c_float, C.c_float, C.c_int, sp, C.c_int]),
        "sim_iter": (None, [sim]),
        "sim_delete": (None, [sim]),
        "sim_set_smooth": (None, [sim, P(t_smooth)]),
        "sim_set_moving_window": (None, [sim]),
        "sim_set_ext_fld": (None, [sim, P(t_emf_ext_fld)]),
        "sim_add_laser": (None, [sim, P(t_emf_laser)]),
        "sim_report_energy": (None, [sim]),
        "report": (C.c_int, [C.c_int, C.c_int]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


def grid_view(buf_ptr, nx: tuple[int, int], gc=((1, 2), (1, 2))) -> np.ndarray:
    """Guard-inclusive float3 grid as a numpy array [rows, cols, 3] (no copy)."""
    cols = gc[0][0] + nx[0] + gc[0][1]
    rows = gc[1][0] + nx[1] + gc[1][1]
    arr = np.ctypeslib.as_array(C.cast(buf_ptr, C.POINTER(C.c_float)), shape=(rows * cols * 3,))
    return arr.reshape(rows, cols, 3)


def particles_view(spec) -> np.ndarray:
    """The species host buffer part[0:np] as a structured array (no copy)."""
    if spec.np <= 0 or not spec.part:
        return np.zeros(0, dtype=PART_DTYPE)
    raw = np.ctypeslib.as_array(C.cast(spec.part, C.POINTER(C.c_byte)), shape=(spec.np * 28,))
    return raw.view(PART_DTYPE)
