"""Loader and thin Python binding of libgiraffe_amd.so (the HIP compute path).

There is no CPU fallback: if the library is missing, or no GPU is visible,
every entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os
import re

import numpy as np

from . import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgiraffe_amd.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "giraffe_amd.h")
_lib = None


class GiraffeError(RuntimeError):
    pass


def header_symbols():
    """Function names declared in include/giraffe_amd.h."""
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:int|const char \*|void)\s*\*?\s*(gf_\w+)\s*\(", txt, flags=re.M)))


def load():
    """Load libgiraffe_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise GiraffeError(f"{LIB_PATH} is missing: run __graft_entry__.build() (make -C giraffe_amd/csrc)")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    dp = C.POINTER(C.c_double)
    L.gf_last_error.restype = C.c_char_p
    L.gf_version.restype = C.c_int
    L.gf_default_options.argtypes = [C.POINTER(abi.GfOptions)]
    L.gf_model_create.argtypes = [C.POINTER(abi.GfModelDesc), C.POINTER(abi.GfOptions), C.POINTER(vp)]
    L.gf_model_destroy.argtypes = [vp]
    L.gf_model_info.argtypes = [vp, C.POINTER(C.c_longlong)]
    L.gf_model_stats.argtypes = [vp, C.POINTER(C.c_longlong)]
    L.gf_set_solution_window.argtypes = [vp, C.c_double, C.c_double]
    L.gf_begin_step.argtypes = [vp, C.c_double, C.c_double]
    L.gf_assemble.argtypes = [vp, dp]
    L.gf_solve.argtypes = [vp, C.POINTER(C.c_int)]
    L.gf_update.argtypes = [vp, dp]
    L.gf_commit.argtypes = [vp]
    L.gf_rollback.argtypes = [vp]
    L.gf_static_solve.argtypes = [vp, C.POINTER(abi.GfStaticParams), C.POINTER(abi.GfStaticReport)]
    L.gf_set_gravity_mode.argtypes = [vp, C.c_int]
    L.gf_static_solve_steps.argtypes = [vp, C.c_int, C.POINTER(abi.GfStaticParams), C.POINTER(C.c_int),
                                        C.POINTER(abi.GfStaticReport), C.POINTER(C.c_int)]
    L.gf_deck_steps.argtypes = [vp, C.POINTER(C.c_int)]
    L.gf_deck_step.argtypes = [vp, C.c_int, C.POINTER(abi.GfStaticParams), C.POINTER(C.c_int)]
    for f in ("gf_read_nodes", "gf_read_disp", "gf_read_residual", "gf_read_increment", "gf_write_disp"):
        getattr(L, f).argtypes = [vp, dp]
    L.gf_read_tangent_dense.argtypes = [vp, dp, C.c_longlong]
    L.gf_write_results.argtypes = [vp, C.c_char_p]
    L.gf_read_element_results.argtypes = [vp, dp, dp]
    L.gf_beam_eval.argtypes = [C.c_int, dp, dp, dp, dp, dp, dp, dp]
    L.gf_shell_eval.argtypes = [C.c_int, dp, dp, dp, dp, dp, dp]
    L.gf_bandsystem.argtypes = [C.c_int, C.c_int, dp, dp, dp, C.POINTER(C.c_int)]
    L.gf_btdsystem.argtypes = [C.c_int, C.c_int, dp, dp, dp, dp, dp, C.POINTER(C.c_int)]
    L.gf_reset_state.argtypes = [vp]
    L.gf_timing_enable.argtypes = [vp, C.c_int]
    L.gf_timing_read.argtypes = [vp, dp, C.POINTER(C.c_longlong), C.c_int]
    L.gf_kernel_probe.argtypes = [vp, C.c_uint]
    L.gf_kernel_probe_read.argtypes = [vp, dp, C.POINTER(C.c_longlong), dp, C.c_int]
    L.gf_model_footprint.argtypes = [C.POINTER(abi.GfModelDesc), C.POINTER(abi.GfOptions), C.POINTER(C.c_ulonglong)]
    L.gf_model_plan.argtypes = [C.POINTER(abi.GfModelDesc), C.POINTER(C.c_longlong)]
    ip = C.POINTER(C.c_int)
    L.gf_front_plan.argtypes = [C.POINTER(abi.GfModelDesc), C.POINTER(C.c_longlong), ip, ip, ip]
    L.gf_default_server_options.argtypes = [C.POINTER(abi.GfServerOptions)]
    L.gf_server_create.argtypes = [C.POINTER(abi.GfServerOptions), C.POINTER(vp)]
    L.gf_server_destroy.argtypes = [vp]
    L.gf_server_submit.argtypes = [vp, C.POINTER(abi.GfModelDesc), C.POINTER(abi.GfStaticParams),
                                   C.POINTER(C.c_longlong)]
    L.gf_server_pump.argtypes = [vp, C.c_int]
    L.gf_server_counts.argtypes = [vp, C.POINTER(C.c_longlong)]
    L.gf_server_take.argtypes = [vp, C.c_longlong, C.POINTER(abi.GfStaticReport), dp, dp]
    L.gf_deck_read.argtypes = [C.c_char_p, C.POINTER(vp)]
    L.gf_deck_desc.argtypes = [vp, C.POINTER(abi.GfModelDesc), C.POINTER(abi.GfStaticParams)]
    L.gf_deck_destroy.argtypes = [vp]
    L.gf_deck_write.argtypes = [C.c_char_p, C.POINTER(abi.GfModelDesc), C.POINTER(abi.GfStaticParams)]
    _lib = L
    return L


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def check(status, what=""):
    if status != abi.GF_OK:
        msg = load().gf_last_error().decode(errors="replace")
        raise GiraffeError(f"{what} failed with status {status}: {msg}")


def options(precision=abi.GF_PREC_FP64, device=0, refine_max=None, refine_tol=None, solver=None):
    """gf_options from the library's defaults (gf_default_options), with the
    given fields replaced."""
    o = abi.GfOptions()
    load().gf_default_options(C.byref(o))
    o.precision = precision
    o.device = device
    if solver is not None:
        o.solver = solver
    if refine_max is not None:
        o.refine_max = refine_max
    if refine_tol is not None:
        o.refine_tol = refine_tol
    return o


class DeviceModel:
    """One model resident on the GPU; mirrors the gf_* entry points."""

    def __init__(self, model, precision=abi.GF_PREC_FP64, device=0, refine_reuse=True, solver=abi.GF_SOLVER_AUTO):
        L = load()
        self.model = model
        self._desc = model.desc()
        self._opt = options(precision, device, solver=solver)
        self._opt.refine_reuse = 1 if refine_reuse else 0
        h = C.c_void_p()
        check(L.gf_model_create(C.byref(self._desc), C.byref(self._opt), C.byref(h)), "gf_model_create")
        self.h = h
        self.nn = model.n_nodes
        self.nb = model.n_beams
        self.ns = model.n_shells

    def close(self):
        if getattr(self, "h", None):
            load().gf_model_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self):
        o = (C.c_longlong * 8)()
        check(load().gf_model_info(self.h, o), "gf_model_info")
        keys = ["n_free", "half_band", "block", "n_blocks", "colors", "factor_bytes", "layout", "precision"]
        return dict(zip(keys, list(o)))

    def stats(self):
        o = (C.c_longlong * 5)()
        check(load().gf_model_stats(self.h, o), "gf_model_stats")
        return {"solves": o[0], "refine_sweeps": o[1], "fallbacks": o[2], "fp64_direct": o[3],
                "fp32_factorizations": o[4]}

    def set_window(self, start, end):
        check(load().gf_set_solution_window(self.h, start, end), "gf_set_solution_window")

    def begin_step(self, t_last, dt):
        check(load().gf_begin_step(self.h, t_last, dt), "gf_begin_step")

    def assemble(self):
        r = np.zeros(3)
        check(load().gf_assemble(self.h, _dp(r)), "gf_assemble")
        return r

    def solve(self):
        f = C.c_int(0)
        check(load().gf_solve(self.h, C.byref(f)), "gf_solve")
        return f.value

    def update(self):
        r = np.zeros(5)
        check(load().gf_update(self.h, _dp(r)), "gf_update")
        return r

    def commit(self):
        check(load().gf_commit(self.h), "gf_commit")

    def rollback(self):
        check(load().gf_rollback(self.h), "gf_rollback")

    def _read(self, fn, shape):
        out = np.zeros(shape)
        check(getattr(load(), fn)(self.h, _dp(out)), fn)
        return out

    def nodes(self):
        return self._read("gf_read_nodes", (self.nn, 6))

    def write_results(self, path):
        """Configuration results file (nodes and elements) of the last converged configuration."""
        check(load().gf_write_results(self.h, os.fsencode(path)), "gf_write_results")

    def element_results(self):
        """Gauss-point means of the committed state: Beam_1 [nb, 12] (sigma_r, epsilon_r) and
        Shell_1 [ns, 24] (n_r1 m_r1 n_r2 m_r2 eta_r1 kappa_r1 eta_r2 kappa_r2)."""
        b = np.zeros((self.nb, 12))
        sh = np.zeros((self.ns, 24))
        check(load().gf_read_element_results(self.h, _dp(b), _dp(sh)), "gf_read_element_results")
        return b, sh

    def disp(self):
        return self._read("gf_read_disp", (self.nn, 6))

    def residual(self):
        return self._read("gf_read_residual", (self.nn, 6))

    def increment(self):
        return self._read("gf_read_increment", (self.nn, 6))

    def write_disp(self, d):
        d = np.ascontiguousarray(d, float).reshape(self.nn, 6)
        check(load().gf_write_disp(self.h, _dp(d)), "gf_write_disp")

    def tangent_dense(self):
        n6 = 6 * self.nn
        out = np.zeros((n6, n6))
        check(load().gf_read_tangent_dense(self.h, _dp(out), out.size), "gf_read_tangent_dense")
        return out

    def reset(self):
        check(load().gf_reset_state(self.h), "gf_reset_state")

    def timing(self, on=True):
        check(load().gf_timing_enable(self.h, 1 if on else 0), "gf_timing_enable")

    def timing_read(self, reset=True):
        ms = np.zeros(3)
        n = (C.c_longlong * 3)()
        check(load().gf_timing_read(self.h, _dp(ms), n, 1 if reset else 0), "gf_timing_read")
        names = ["assemble", "factor", "solve"]
        return {k: (float(ms[i]), int(n[i])) for i, k in enumerate(names)}

    KERNEL_CLASSES = ("btd_level", "btd_tail", "band_panel", "band_update", "band_sweep", "beam_assemble",
                      "shell_assemble", "front_step", "front_assemble", "front_solve")

    def kernel_probe(self, classes):
        """Bracket every launch of the named kernel classes with HIP events
        while timing is on (KERNEL_CLASSES; empty to stop)."""
        mask = 0
        for c in classes:
            mask |= 1 << self.KERNEL_CLASSES.index(c)
        check(load().gf_kernel_probe(self.h, mask), "gf_kernel_probe")

    def kernel_probe_read(self, reset=True):
        """{class: (ms, launches, units)} for the classes that ran."""
        k = len(self.KERNEL_CLASSES)
        ms, units = np.zeros(k), np.zeros(k)
        n = (C.c_longlong * k)()
        check(load().gf_kernel_probe_read(self.h, _dp(ms), n, _dp(units), 1 if reset else 0), "gf_kernel_probe_read")
        return {c: (float(ms[i]), int(n[i]), float(units[i])) for i, c in enumerate(self.KERNEL_CLASSES) if n[i]}

    def static_solve(self, params, cap=4096):
        iters = np.zeros(cap, np.int32)
        rep = abi.GfStaticReport()
        rep.step_iters = iters.ctypes.data_as(C.POINTER(C.c_int))
        rep.step_iters_cap = cap
        pc = params.c()
        check(load().gf_static_solve(self.h, C.byref(pc), C.byref(rep)), "gf_static_solve")
        n = min(rep.n_step_iters, cap)
        return rep, iters[:n].copy()

    def set_gravity_mode(self, mode):
        check(load().gf_set_gravity_mode(self.h, int(mode)), "gf_set_gravity_mode")

    def static_solve_steps(self, steps, cap=4096):
        """gf_static_solve_steps: steps = [(params, gravity_mode)] in sequence,
        stopping after a step that aborts.  Returns [(report, iterations)] for
        the steps run."""
        n = len(steps)
        iters = [np.zeros(cap, np.int32) for _ in range(n)]
        ps = (abi.GfStaticParams * n)(*[q.c() for q, _ in steps])
        modes = (C.c_int * n)(*[int(m) for _, m in steps])
        reps = (abi.GfStaticReport * n)()
        for r, it in zip(reps, iters):
            r.step_iters = it.ctypes.data_as(C.POINTER(C.c_int))
            r.step_iters_cap = cap
        run = C.c_int(0)
        check(load().gf_static_solve_steps(self.h, n, ps, modes, reps, C.byref(run)), "gf_static_solve_steps")
        return [(reps[i], iters[i][:min(reps[i].n_step_iters, cap)].copy()) for i in range(run.value)]


def beam_eval(x, cs, prop, d, gp):
    """Element-level Beam_1::Mount on the device for independent elements.

    x [n,9], cs [n,9], prop [n,9], d [n,18], gp [n,2,15] -> K [n,18,18], f [n,18]."""
    x = np.ascontiguousarray(x, float)
    n = x.shape[0]
    cs = np.ascontiguousarray(cs, float)
    prop = np.ascontiguousarray(prop, float)
    d = np.ascontiguousarray(d, float)
    gp = np.ascontiguousarray(gp, float)
    K = np.zeros((n, 18, 18))
    f = np.zeros((n, 18))
    check(load().gf_beam_eval(n, _dp(x), _dp(cs), _dp(prop), _dp(d), _dp(gp), _dp(K), _dp(f)), "gf_beam_eval")
    return K, f


def shell_eval(x, prop, d, gp):
    """Element-level Shell_1::Mount on the device for independent elements.

    x [n,18], prop [n,4], d [n,36], gp [n,3,21] -> K [n,27,27], f [n,27]."""
    x = np.ascontiguousarray(x, float)
    n = x.shape[0]
    K = np.zeros((n, 27, 27))
    f = np.zeros((n, 27))
    check(load().gf_shell_eval(n, _dp(x), _dp(np.ascontiguousarray(prop, float)), _dp(np.ascontiguousarray(d, float)),
                               _dp(np.ascontiguousarray(gp, float)), _dp(K), _dp(f)), "gf_shell_eval")
    return K, f


def bandsystem(A_rows, hb, b):
    """Solve with the device band LU; A_rows [n, 2hb+1] (row i: columns i-hb..i+hb)."""
    A_rows = np.ascontiguousarray(A_rows, float)
    n = A_rows.shape[0]
    x = np.zeros(n)
    fail = C.c_int(0)
    check(load().gf_bandsystem(n, hb, _dp(A_rows), _dp(np.ascontiguousarray(b, float)), _dp(x), C.byref(fail)),
          "gf_bandsystem")
    return x, fail.value


def btdsystem(A_blocks, B_blocks, C_blocks, b):
    """Solve a block-tridiagonal system with the device cyclic reduction;
    blocks [nblk, m, m] (A[i] couples block row i to i-1, C[i] to i+1)."""
    A_blocks, B_blocks, C_blocks = (np.ascontiguousarray(v, float) for v in (A_blocks, B_blocks, C_blocks))
    nblk, m = B_blocks.shape[0], B_blocks.shape[1]
    x = np.zeros(nblk * m)
    fail = C.c_int(0)
    check(load().gf_btdsystem(m, nblk, _dp(A_blocks), _dp(B_blocks), _dp(C_blocks),
                              _dp(np.ascontiguousarray(b, float)), _dp(x), C.byref(fail)), "gf_btdsystem")
    return x, fail.value


def front_plan(model):
    """Host-side multifrontal plan of `model` (gf_front_plan): sizes, per-front
    arrays (k, kp, mp, level, parent, row-map offset), row maps and the DOF map."""
    desc = model.desc()
    info = (C.c_longlong * 8)()
    L = load()
    check(L.gf_front_plan(C.byref(desc), info, None, None, None), "gf_front_plan")
    nf, nrows = info[0], info[2]
    fronts = np.zeros((nf, 6), np.int32)
    rows = np.zeros(max(nrows, 1), np.int32)
    dof = np.zeros((model.n_nodes, 6), np.int32)
    ip = lambda a: a.ctypes.data_as(C.POINTER(C.c_int))
    check(L.gf_front_plan(C.byref(desc), info, ip(fronts), ip(rows), ip(dof)), "gf_front_plan")
    keys = ["fronts", "levels", "rows", "front_elems", "inv_elems", "flops", "factor_launches", "solve_launches"]
    out = dict(zip(keys, list(info)))
    out.update({"front": fronts, "row_map": rows[:nrows], "dof": dof})
    return out


def footprint(model, precision=abi.GF_PREC_FP64, solver=None):
    """Device bytes a DeviceModel of `model` would hold (gf_model_footprint)."""
    desc = model.desc()
    o = options(precision, solver=solver)
    b = C.c_ulonglong()
    check(load().gf_model_footprint(C.byref(desc), C.byref(o), C.byref(b)), "gf_model_footprint")
    return int(b.value)


def plan(model):
    """Host-side planning of `model` (gf_model_plan): n_free, half-band, BTD block
    size (0: band layout) and whether the DOFs follow reverse Cuthill-McKee order."""
    desc = model.desc()
    o = (C.c_longlong * 4)()
    check(load().gf_model_plan(C.byref(desc), o), "gf_model_plan")
    return {"n_free": o[0], "half_band": o[1], "block": o[2], "rcm": bool(o[3])}


class Server:
    """Request server (gf_server_*): independent static solves resident on one GPU,
    admitted FIFO by slot and device memory, advanced one Newton step at a time."""

    def __init__(self, precision=abi.GF_PREC_FP64, device=0, max_inflight=16, mem_cap=200 << 30):
        L = load()
        o = abi.GfServerOptions()
        L.gf_default_server_options(C.byref(o))
        o.precision, o.device, o.max_inflight, o.mem_cap = precision, device, max_inflight, int(mem_cap)
        h = C.c_void_p()
        check(L.gf_server_create(C.byref(o), C.byref(h)), "gf_server_create")
        self.h = h
        self.n_nodes = {}

    def submit(self, model, params):
        desc = model.desc()
        pc = params.c()
        rid = C.c_longlong()
        check(load().gf_server_submit(self.h, C.byref(desc), C.byref(pc), C.byref(rid)), "gf_server_submit")
        self.n_nodes[rid.value] = model.n_nodes
        return rid.value

    def pump(self, wait_us=0):
        n = load().gf_server_pump(self.h, int(wait_us))
        if n < 0:
            check(n, "gf_server_pump")
        return n

    def counts(self):
        o = (C.c_longlong * 4)()
        check(load().gf_server_counts(self.h, o), "gf_server_counts")
        return {"queued": o[0], "resident": o[1], "finished": o[2], "bytes": o[3]}

    def take(self, rid, cap=4096):
        """(report, step iterations, nodes [n][6], times_ms [submit->admit, submit->end, admit->end])."""
        iters = np.zeros(cap, np.int32)
        rep = abi.GfStaticReport()
        rep.step_iters = iters.ctypes.data_as(C.POINTER(C.c_int))
        rep.step_iters_cap = cap
        if rid not in self.n_nodes:
            raise GiraffeError(f"unknown request {rid}")
        nodes = np.zeros((self.n_nodes[rid], 6))
        times = np.zeros(3)
        check(load().gf_server_take(self.h, rid, C.byref(rep), _dp(nodes), _dp(times)), "gf_server_take")
        del self.n_nodes[rid]
        return rep, iters[:min(rep.n_step_iters, cap)].copy(), nodes, times

    def close(self):
        if getattr(self, "h", None):
            load().gf_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Deck:
    """A GIRAFFE input deck read by gf_deck_read (the static hot-path subset).
    Behaves like a Model for DeviceModel / Server (desc(), n_nodes); `params`
    holds its Static and ConvergenceCriteria parameters."""

    def __init__(self, path):
        from .model import StaticParams

        L = load()
        h = C.c_void_p()
        check(L.gf_deck_read(os.fsencode(path), C.byref(h)), "gf_deck_read")
        self.h = h
        d = abi.GfModelDesc()
        gp = abi.GfStaticParams()
        check(L.gf_deck_desc(self.h, C.byref(d), C.byref(gp)), "gf_deck_desc")
        self._desc = d
        self.params = StaticParams(**{f: getattr(gp, f) for f, _ in abi.GfStaticParams._fields_})
        n = C.c_int(0)
        check(L.gf_deck_steps(self.h, C.byref(n)), "gf_deck_steps")
        self.steps = []  # [(StaticParams, gravity mode)] of every solution step
        for i in range(n.value):
            q, mode = abi.GfStaticParams(), C.c_int(0)
            check(L.gf_deck_step(self.h, i, C.byref(q), C.byref(mode)), "gf_deck_step")
            self.steps.append((StaticParams(**{f: getattr(q, f) for f, _ in abi.GfStaticParams._fields_}),
                               mode.value))

    @property
    def n_nodes(self):
        return self._desc.n_nodes

    @property
    def n_beams(self):
        return self._desc.n_beams

    @property
    def n_shells(self):
        return self._desc.n_shells

    def desc(self):
        d = abi.GfModelDesc()
        C.pointer(d)[0] = self._desc
        d._keep = self  # the arrays belong to the deck
        return d

    def close(self):
        if getattr(self, "h", None):
            load().gf_deck_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def write_deck(path, model, params):
    """Writes `model` (Model or Deck) and `params` as a GIRAFFE deck (gf_deck_write)."""
    desc = model.desc()
    pc = params.c()
    check(load().gf_deck_write(os.fsencode(path), C.byref(desc), C.byref(pc)), "gf_deck_write")
