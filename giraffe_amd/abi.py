"""ctypes mirror of include/giraffe_amd.h (structs and constants).

Field order and types must match the C header exactly.
"""
import ctypes as C

GF_OK = 0
GF_E_ARG = -1
GF_E_DEVICE = -2
GF_E_NOMEM = -3
GF_E_SINGULAR = -4
GF_E_STATE = -5
GF_E_UNSUPPORTED = -6
GF_E_CAPACITY = -7

GF_PREC_FP64 = 0
GF_PREC_MIXED = 1
GF_GRAVITY_OFF, GF_GRAVITY_RAMP_UP, GF_GRAVITY_ON, GF_GRAVITY_RAMP_DOWN = 0, 1, 2, 3
GF_SOLVER_AUTO, GF_SOLVER_BAND, GF_SOLVER_FRONTS = 0, 1, 2

GF_LOAD_NODAL = 0
GF_LOAD_NODAL_FOLLOWER = 1
GF_LOAD_SHELL_PRESSURE = 2

GF_BEAM_NPROP = 9
GF_BEAM_KIND_BEAM1 = 0
GF_BEAM_KIND_PIPE1 = 1
GF_SHELL_NPROP = 4

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_up = C.POINTER(C.c_ubyte)


class GfModelDesc(C.Structure):
    _fields_ = [
        ("n_nodes", C.c_int),
        ("node_xyz", _dp),
        ("node_fix", _up),
        ("n_beams", C.c_int),
        ("beam_conn", _ip),
        ("beam_cs", _dp),
        ("beam_prop", _dp),
        ("n_shells", C.c_int),
        ("shell_conn", _ip),
        ("shell_prop", _dp),
        ("n_loads", C.c_int),
        ("load_kind", _ip),
        ("load_set_off", _ip),
        ("load_set_items", _ip),
        ("load_cs", _dp),
        ("load_flags", _ip),
        ("load_tab_off", _ip),
        ("tab_time", _dp),
        ("tab_val", _dp),
        ("gravity", C.c_double * 3),
        ("gravity_on", C.c_int),
        ("beam_number", _ip),
        ("shell_number", _ip),
        ("beam_kind", _ip),
        ("beam_buoy", _dp),
        ("ocean_on", C.c_int),
        ("ocean_surface", C.c_double * 3),
        ("ocean_rho_fluid", C.c_double),
    ]


class GfStaticParams(C.Structure):
    _fields_ = [
        ("start_time", C.c_double),
        ("end_time", C.c_double),
        ("time_step", C.c_double),
        ("max_time_step", C.c_double),
        ("min_time_step", C.c_double),
        ("max_it", C.c_int),
        ("min_it", C.c_int),
        ("conv_increase", C.c_int),
        ("inc_factor", C.c_double),
        ("sample", C.c_int),
        ("force_tol", C.c_double),
        ("moment_tol", C.c_double),
        ("force_min", C.c_double),
        ("moment_min", C.c_double),
        ("constraint_min", C.c_double),
        ("disp_tol", C.c_double),
        ("rot_tol", C.c_double),
        ("lag_tol", C.c_double),
        ("disp_min", C.c_double),
        ("rot_min", C.c_double),
        ("lag_min", C.c_double),
        ("divergence_ref", C.c_double),
    ]


class GfStaticReport(C.Structure):
    _fields_ = [
        ("aborted", C.c_int),
        ("converged_steps", C.c_int),
        ("diverged_steps", C.c_int),
        ("total_iterations", C.c_int),
        ("total_solves", C.c_int),
        ("final_time", C.c_double),
        ("max_rel_residual", C.c_double),
        ("step_iters", _ip),
        ("step_iters_cap", C.c_int),
        ("n_step_iters", C.c_int),
        ("t_assemble_ms", C.c_double),
        ("t_solve_ms", C.c_double),
        ("t_total_ms", C.c_double),
    ]


class GfOptions(C.Structure):
    _fields_ = [
        ("precision", C.c_int),
        ("device", C.c_int),
        ("stream", C.c_void_p),
        ("refine_max", C.c_int),
        ("refine_tol", C.c_double),
        ("refine_reuse", C.c_int),
        ("solver", C.c_int),
    ]


class GfServerOptions(C.Structure):
    _fields_ = [
        ("precision", C.c_int),
        ("device", C.c_int),
        ("max_inflight", C.c_int),
        ("mem_cap", C.c_ulonglong),
    ]


def ptr(a, ctype):
    """Pointer to a contiguous numpy array (the caller keeps `a` alive)."""
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(ctype))
