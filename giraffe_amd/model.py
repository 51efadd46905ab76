"""Model descriptions for the static Newton-step path and seeded generators for
the benchmark workloads.

A `Model` is the resolved hot-path subset of a GIRAFFE deck (node sets
expanded, sections and materials folded into per-element records); `desc()`
returns the `gf_model_desc` the C ABI takes.  Generators:

* `cantilever_rollup`  -- straight Beam_1 cantilever under a follower tip
  moment ramped to 2*pi*EI/L (rolls into a full circle; Simo & Vu-Quoc 1986).
* `catenary`           -- Beam_1 riser catenary between two pinned ends under
  self-weight and seeded N(0, sigma) lateral nodal loads.
* `shell_plate`        -- Shell_1 clamped square plate under follower pressure.

Deck grammar these correspond to: Node/Beam_1/Shell_1/Hooke/Tube/Homogeneous/CS/
NodalLoad/NodalFollowerLoad/ShellLoad/NodalConstraint/GravityData
(src/IO.cpp:290-660 dispatch, readers cited in giraffe_amd/deck.py).
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field, asdict

import numpy as np

from . import abi

PI = 3.1415926535897932384626433832795  # src/Geometry.h:4


def tube_section(De: float, Di: float):
    """SecTube::PreCalc src/SecTube.cpp:86-106: A, I11, I22, I12, It."""
    A = (PI / 4.0) * (De * De - Di * Di)
    I = (PI / 64.0) * (De ** 4 - Di ** 4)
    It = (PI / 32.0) * (De ** 4 - Di ** 4)
    return A, I, I, 0.0, It


def cs_rows(E1, E3):
    """CoordinateSystem::Read src/CoordinateSystem.cpp:23-92: rows E1, E2=E3xE1, E3 (normalised)."""
    E1 = np.asarray(E1, float)
    E3 = np.asarray(E3, float)
    E2 = np.cross(E3, E1)
    n1, n2, n3 = np.linalg.norm(E1), np.linalg.norm(E2), np.linalg.norm(E3)
    if n1 != 1.0:
        E1 = E1 / n1
    if n2 != 1.0:
        E2 = E2 / n2
    if n3 != 1.0:
        E3 = E3 / n3
    return np.concatenate([E1, E2, E3])


def cs_rows_many(E1, E3):
    """cs_rows for many E3 directions (rows [n][3], normalised first) and one E1."""
    E3 = np.asarray(E3, float)
    E3 = E3 / np.linalg.norm(E3, axis=1)[:, None]
    E1 = np.broadcast_to(np.asarray(E1, float), E3.shape)
    E2 = np.cross(E3, E1)
    n1 = np.linalg.norm(E1, axis=1)[:, None]
    n2 = np.linalg.norm(E2, axis=1)[:, None]
    n3 = np.linalg.norm(E3, axis=1)[:, None]
    E1 = np.where(n1 != 1.0, E1 / n1, E1)
    E2 = np.where(n2 != 1.0, E2 / n2, E2)
    E3 = np.where(n3 != 1.0, E3 / n3, E3)
    return np.concatenate([E1, E2, E3], 1)


GLOBAL_CS = cs_rows([1, 0, 0], [0, 0, 1])


@dataclass
class StaticParams:
    """Static (src/Static.cpp:43-130) + ConvergenceCriteria (src/ConvergenceCriteria.cpp:64-165)."""
    start_time: float = 0.0
    end_time: float = 1.0
    time_step: float = 0.1
    max_time_step: float = 0.1
    min_time_step: float = 1e-4
    max_it: int = 30
    min_it: int = 2
    conv_increase: int = 4
    inc_factor: float = 1.0
    sample: int = 1
    force_tol: float = 1e-4
    moment_tol: float = 1e-4
    force_min: float = 1e-5
    moment_min: float = 1e-5
    constraint_min: float = 1e-7
    disp_tol: float = 1e-4
    rot_tol: float = 1e-4
    lag_tol: float = 1e-4
    disp_min: float = 1e-6
    rot_min: float = 1e-6
    lag_min: float = 1e-6
    divergence_ref: float = 1e15

    def c(self) -> abi.GfStaticParams:
        return abi.GfStaticParams(**asdict(self))


class Model:
    """Resolved model; indices are 0-based."""

    def __init__(self, name: str = "model"):
        self.name = name
        self.xyz = []
        self.fix = []
        self.beam_conn, self.beam_cs, self.beam_prop = [], [], []
        self.beam_kind, self.beam_De = [], []   # GF_BEAM_KIND_*, Pipe_1 outer diameter
        self.shell_conn, self.shell_prop = [], []
        self.loads = []  # (kind, items, cs9, flags, times, vals[n][6])
        self.gravity = (0.0, 0.0, 0.0)
        self.gravity_on = 0
        self.ocean_on, self.ocean_rho_fluid, self.ocean_surface = 0, 0.0, [0.0, 0.0, 0.0]
        # optional deck element numbers (a permutation of 1..n_beams+n_shells)
        self.beam_number = None
        self.shell_number = None
        self._frozen = None

    # -- building ---------------------------------------------------------
    def add_nodes(self, xyz):
        xyz = np.asarray(xyz, float).reshape(-1, 3)
        i0 = len(self.xyz)
        self.xyz.extend(xyz.tolist())
        self.fix.extend([[0] * 6 for _ in range(len(xyz))])
        self._frozen = None
        return i0

    def fix_dofs(self, node: int, dofs):
        for k in dofs:
            self.fix[node][k] = 1
        self._frozen = None

    def add_beam(self, n1, n2, n3, cs9, E, nu, rho, A, I11, I22, I12, It, T0=0.0):
        self.beam_conn.append([n1, n2, n3])
        self.beam_cs.append(list(cs9))
        self.beam_prop.append([E, nu, rho, A, I11, I22, I12, It, T0])
        self.beam_kind.append(abi.GF_BEAM_KIND_BEAM1)
        self.beam_De.append(0.0)
        self._frozen = None

    def add_pipe(self, n1, n2, n3, cs9, EA, EI, GJ, GA, rho_line, De=0.0):
        """Pipe_1 with a PipeSection's EA, EI, GJ, GA, Rho (mass per length) and outer
        diameter De (buoyancy below the ocean surface, set_ocean)."""
        self.beam_conn.append([n1, n2, n3])
        self.beam_cs.append(list(cs9))
        self.beam_prop.append([EA, EI, GJ, GA, rho_line, 0.0, 0.0, 0.0, 0.0])
        self.beam_kind.append(abi.GF_BEAM_KIND_PIPE1)
        self.beam_De.append(float(De))
        self._frozen = None

    def set_ocean(self, rho_fluid, surface):
        """OceanData: fluid density and a point of the free surface (no sea current)."""
        self.ocean_on = 1
        self.ocean_rho_fluid = float(rho_fluid)
        self.ocean_surface = [float(v) for v in surface]
        self._frozen = None

    def add_shell(self, nodes6, E, nu, rho, t):
        self.shell_conn.append(list(nodes6))
        self.shell_prop.append([E, nu, rho, t])
        self._frozen = None

    def add_load(self, kind, items, times, vals, cs9=GLOBAL_CS, flags=0):
        vals = np.asarray(vals, float).reshape(len(times), 6)
        items = np.asarray(list(items), np.int32)
        self.loads.append({"kind": np.array([kind], np.int32), "set_off": np.array([0, len(items)], np.int32),
                           "items": items, "cs": np.asarray(cs9, float).reshape(1, 9),
                           "flags": np.array([int(flags)], np.int32),
                           "tab_off": np.array([0, len(times)], np.int32),
                           "tab_time": np.asarray(times, float), "tab_val": vals})
        self._frozen = None

    def add_nodal_loads(self, kind, nodes, times, vals, cs9=GLOBAL_CS):
        """One load per node (load i on nodes[i]) sharing the instants `times`;
        vals [len(nodes)][len(times)][6].  Same as add_load per node, in order."""
        nodes = np.asarray(nodes, np.int32)
        n, nt = len(nodes), len(times)
        vals = np.asarray(vals, float).reshape(n, nt, 6)
        self.loads.append({"kind": np.full(n, kind, np.int32), "set_off": np.arange(n + 1, dtype=np.int32),
                           "items": nodes, "cs": np.tile(np.asarray(cs9, float).reshape(1, 9), (n, 1)),
                           "flags": np.zeros(n, np.int32), "tab_off": (np.arange(n + 1) * nt).astype(np.int32),
                           "tab_time": np.tile(np.asarray(times, float), n), "tab_val": vals.reshape(n * nt, 6)})
        self._frozen = None

    def add_beams(self, conn, cs, prop):
        """Bulk add_beam: conn [n][3], cs [n][9], prop [n][GF_BEAM_NPROP]."""
        self.beam_conn.extend(np.asarray(conn, np.int64).reshape(-1, 3).tolist())
        self.beam_cs.extend(np.asarray(cs, float).reshape(-1, 9).tolist())
        self.beam_prop.extend(np.asarray(prop, float).reshape(-1, abi.GF_BEAM_NPROP).tolist())
        n = len(self.beam_conn) - len(self.beam_kind)
        self.beam_kind.extend([abi.GF_BEAM_KIND_BEAM1] * n)
        self.beam_De.extend([0.0] * n)
        self._frozen = None

    # -- sizes ------------------------------------------------------------
    @property
    def n_nodes(self):
        return len(self.xyz)

    @property
    def n_beams(self):
        return len(self.beam_conn)

    @property
    def n_shells(self):
        return len(self.shell_conn)

    @property
    def n_loads(self):
        return int(sum(len(c["kind"]) for c in self.loads))

    def freeze(self):
        if self._frozen is not None:
            return self._frozen
        f = {}
        f["xyz"] = np.ascontiguousarray(np.asarray(self.xyz, float).reshape(-1, 3))
        f["fix"] = np.ascontiguousarray(np.asarray(self.fix, np.uint8).reshape(-1, 6))
        f["beam_conn"] = np.ascontiguousarray(np.asarray(self.beam_conn, np.int32).reshape(-1, 3))
        f["beam_cs"] = np.ascontiguousarray(np.asarray(self.beam_cs, float).reshape(-1, 9))
        f["beam_prop"] = np.ascontiguousarray(np.asarray(self.beam_prop, float).reshape(-1, abi.GF_BEAM_NPROP))
        f["shell_conn"] = np.ascontiguousarray(np.asarray(self.shell_conn, np.int32).reshape(-1, 6))
        f["shell_prop"] = np.ascontiguousarray(np.asarray(self.shell_prop, float).reshape(-1, abi.GF_SHELL_NPROP))
        L = self.loads
        nl = self.n_loads
        if nl:
            f["load_kind"] = np.concatenate([c["kind"] for c in L])
            so, to, items_base, tab_base = [np.zeros(1, np.int32)], [np.zeros(1, np.int32)], 0, 0
            for c in L:
                so.append(c["set_off"][1:] + items_base)
                to.append(c["tab_off"][1:] + tab_base)
                items_base += int(c["set_off"][-1])
                tab_base += int(c["tab_off"][-1])
            f["load_set_off"] = np.ascontiguousarray(np.concatenate(so), np.int32)
            items = np.concatenate([c["items"] for c in L])
            f["load_set_items"] = np.ascontiguousarray(items if len(items) else np.zeros(1), np.int32)
            f["load_cs"] = np.ascontiguousarray(np.concatenate([c["cs"] for c in L]), float)
            f["load_flags"] = np.ascontiguousarray(np.concatenate([c["flags"] for c in L]), np.int32)
            f["load_tab_off"] = np.ascontiguousarray(np.concatenate(to), np.int32)
            f["tab_time"] = np.ascontiguousarray(np.concatenate([c["tab_time"] for c in L]), float)
            f["tab_val"] = np.ascontiguousarray(np.concatenate([c["tab_val"] for c in L]), float)
        else:
            f["load_kind"] = np.zeros(1, np.int32)
            f["load_set_off"] = np.zeros(1, np.int32)
            f["load_set_items"] = np.zeros(1, np.int32)
            f["load_cs"] = np.zeros((1, 9))
            f["load_flags"] = np.zeros(1, np.int32)
            f["load_tab_off"] = np.zeros(1, np.int32)
            f["tab_time"] = np.zeros(1)
            f["tab_val"] = np.zeros((1, 6))
        self._frozen = f
        return f

    def desc(self) -> abi.GfModelDesc:
        f = self.freeze()
        d = abi.GfModelDesc()
        d.n_nodes = self.n_nodes
        d.node_xyz = abi.ptr(f["xyz"], C.c_double)
        d.node_fix = abi.ptr(f["fix"], C.c_ubyte)
        d.n_beams = self.n_beams
        d.beam_conn = abi.ptr(f["beam_conn"], C.c_int)
        d.beam_cs = abi.ptr(f["beam_cs"], C.c_double)
        d.beam_prop = abi.ptr(f["beam_prop"], C.c_double)
        d.n_shells = self.n_shells
        d.shell_conn = abi.ptr(f["shell_conn"], C.c_int)
        d.shell_prop = abi.ptr(f["shell_prop"], C.c_double)
        d.n_loads = self.n_loads
        d.load_kind = abi.ptr(f["load_kind"], C.c_int)
        d.load_set_off = abi.ptr(f["load_set_off"], C.c_int)
        d.load_set_items = abi.ptr(f["load_set_items"], C.c_int)
        d.load_cs = abi.ptr(f["load_cs"], C.c_double)
        d.load_flags = abi.ptr(f["load_flags"], C.c_int)
        d.load_tab_off = abi.ptr(f["load_tab_off"], C.c_int)
        d.tab_time = abi.ptr(f["tab_time"], C.c_double)
        d.tab_val = abi.ptr(f["tab_val"], C.c_double)
        d.gravity = (C.c_double * 3)(*self.gravity)
        d.gravity_on = int(self.gravity_on)
        if any(k != abi.GF_BEAM_KIND_BEAM1 for k in self.beam_kind):
            f["beam_kind"] = np.ascontiguousarray(self.beam_kind, np.int32)
            d.beam_kind = abi.ptr(f["beam_kind"], C.c_int)
            # Pipe_1 buoyant mass per length rho_fluid pi De^2 / 4 (only with OceanData)
            De = np.asarray(self.beam_De, float)
            f["beam_buoy"] = np.ascontiguousarray(self.ocean_rho_fluid * (np.pi * De * De / 4.0) if self.ocean_on
                                                  else np.zeros_like(De))
            d.beam_buoy = abi.ptr(f["beam_buoy"], C.c_double)
        d.ocean_on = int(self.ocean_on)
        d.ocean_surface = (C.c_double * 3)(*self.ocean_surface)
        d.ocean_rho_fluid = self.ocean_rho_fluid
        if self.beam_number is not None:
            f["beam_number"] = np.ascontiguousarray(self.beam_number, np.int32)
            d.beam_number = abi.ptr(f["beam_number"], C.c_int)
        if self.shell_number is not None:
            f["shell_number"] = np.ascontiguousarray(self.shell_number, np.int32)
            d.shell_number = abi.ptr(f["shell_number"], C.c_int)
        d._keep = f  # keep arrays alive with the struct
        return d


# ---------------------------------------------------------------------------
# Generators
# ---------------------------------------------------------------------------

def cantilever_rollup(n_elem=100, L=10.0, De=0.2, Di=0.18, E=210e9, nu=0.3, rho=7850.0,
                      n_steps=20, turns=1.0, tol=1e-10):
    """Cantilever along x, clamped at node 0, follower moment about z at the tip
    ramped linearly to turns*2*pi*EI/L over n_steps load steps."""
    m = Model("cantilever_rollup")
    nn = 2 * n_elem + 1
    xs = np.linspace(0.0, L, nn)
    m.add_nodes(np.stack([xs, np.zeros(nn), np.zeros(nn)], 1))
    A, I11, I22, I12, It = tube_section(De, Di)
    cs = cs_rows([0, 1, 0], [1, 0, 0])
    for e in range(n_elem):
        m.add_beam(2 * e, 2 * e + 1, 2 * e + 2, cs, E, nu, rho, A, I11, I22, I12, It)
    m.fix_dofs(0, range(6))
    EI = E * I22
    M = turns * 2.0 * PI * EI / L
    m.add_load(abi.GF_LOAD_NODAL_FOLLOWER, [nn - 1], [0.0, 1.0],
               [[0, 0, 0, 0, 0, 0], [0, 0, 0, 0, 0, M]], GLOBAL_CS)
    dt = 1.0 / n_steps
    p = StaticParams(end_time=1.0, time_step=dt, max_time_step=dt, min_time_step=dt / 1024.0,
                     max_it=30, min_it=2, conv_increase=4, inc_factor=1.0,
                     force_tol=tol, moment_tol=tol, disp_tol=tol, rot_tol=tol,
                     force_min=M / L, moment_min=M, disp_min=1e-6, rot_min=1e-6)
    info = {"EI": EI, "M": M, "L": L, "tip": nn - 1}
    return m, p, info


def catenary_shape(span_h, span_v, arc_ratio):
    """Catenary z = a cosh((x - x0)/a) + c through (0,0) and (span_h, span_v) with
    arc length arc_ratio * chord.  Returns (a, x0, c, S)."""
    chord = math.hypot(span_h, span_v)
    S = arc_ratio * chord
    target = math.sqrt(S * S - span_v * span_v)
    lo, hi = 1e-3 * span_h, 1e6 * span_h
    for _ in range(200):  # 2 a sinh(span_h / 2a) = target, decreasing in a
        mid = math.sqrt(lo * hi)
        val = 2 * mid * math.sinh(min(span_h / (2 * mid), 700.0))
        if val > target:
            lo = mid
        else:
            hi = mid
    a = math.sqrt(lo * hi)
    # z(span_h) - z(0) = span_v fixes x0
    # a[cosh((D-x0)/a) - cosh(-x0/a)] = span_v -> 2a sinh(D/2a) sinh((D/2 - x0)/a) = span_v
    s = span_v / (2 * a * math.sinh(span_h / (2 * a)))
    x0 = span_h / 2 - a * math.asinh(s)
    c = -a * math.cosh(-x0 / a)
    return a, x0, c, S


def catenary(n_elem=10000, span_h=1500.0, span_v=1000.0, arc_ratio=1.15, De=0.3, Di=0.26,
             E=210e9, nu=0.3, w=1500.0, g=9.81, sigma=500.0, seed=1002, n_steps=10,
             lateral_start=0.3, tol=1e-10):
    """Beam_1 riser catenary hung between pins at (0,0,0) and (span_h,0,span_v).

    Geometry: nodes equally spaced in arc length on the catenary, each element
    straight (its middle node at the midpoint of nodes 1 and 3).  Self-weight w
    [N/m] through rho = w/(g A), ramped 0->1 over the solution (GravityData
    BoolTable).  Lateral (global y) nodal loads N(0, sigma) per node, seeded,
    ramped from t = lateral_start to 1 (the stress-free first steps carry
    gravity only).  Pins: UX UY UZ at both ends, plus ROTX (rotation about the
    horizontal axis of the catenary plane) at both ends, which removes the
    zero-stiffness rigid swing about the chord and the free end twist.
    """
    m = Model("catenary")
    a, x0, c, S = catenary_shape(span_h, span_v, arc_ratio)
    # arc length s(x) = a [sinh((x-x0)/a) - sinh(-x0/a)] -> x(s)
    sh0 = math.sinh(-x0 / a)
    ne = n_elem
    s_end = np.linspace(0.0, S, ne + 1)
    xe = x0 + a * np.arcsinh(s_end / a + sh0)
    ze = a * np.cosh((xe - x0) / a) + c
    ze[0], ze[-1] = 0.0, span_v
    xe[0], xe[-1] = 0.0, span_h
    nn = 2 * ne + 1
    xyz = np.zeros((nn, 3))
    xyz[0::2, 0] = xe
    xyz[0::2, 2] = ze
    xyz[1::2, 0] = 0.5 * (xe[:-1] + xe[1:])
    xyz[1::2, 2] = 0.5 * (ze[:-1] + ze[1:])
    m.add_nodes(xyz)
    A, I11, I22, I12, It = tube_section(De, Di)
    rho = w / (g * A)
    e = np.arange(ne)
    m.add_beams(np.stack([2 * e, 2 * e + 1, 2 * e + 2], 1), cs_rows_many([0.0, 1.0, 0.0], xyz[2::2] - xyz[:-2:2]),
                np.tile([E, nu, rho, A, I11, I22, I12, It, 0.0], (ne, 1)))
    m.fix_dofs(0, [0, 1, 2, 3])
    m.fix_dofs(nn - 1, [0, 1, 2, 3])
    m.gravity = (0.0, 0.0, -g)
    m.gravity_on = 1
    rng = np.random.default_rng(seed)
    fy = rng.normal(0.0, sigma, nn)
    vals = np.zeros((nn, 3, 6))
    vals[:, 2, 1] = fy
    m.add_nodal_loads(abi.GF_LOAD_NODAL, np.arange(nn), [0.0, lateral_start, 1.0], vals, GLOBAL_CS)
    dt = 1.0 / n_steps
    Wtot = w * S
    p = StaticParams(end_time=1.0, time_step=dt, max_time_step=dt, min_time_step=dt / 1024.0,
                     max_it=40, min_it=2, conv_increase=4, inc_factor=1.0,
                     force_tol=tol, moment_tol=tol, disp_tol=tol, rot_tol=tol,
                     force_min=Wtot / 10.0, moment_min=Wtot, disp_min=1e-6, rot_min=1e-6)
    info = {"a": a, "S": S, "rho": rho, "A": A, "n_free_expected": 6 * nn - 8}
    return m, p, info


def shell_plate(nq=64, L=1.0, t=0.005, E=70e9, nu=0.33, rho=2700.0, p_max=1.7e7, n_steps=15,
                imperf=1e-4, seed=1003, area_update=1, tol=1e-10, ramp="cubic"):
    """Clamped square Shell_1 plate: nq x nq quads, each split into two 6-node
    triangles ((2nq+1)^2 nodes on a grid of pitch L/(2nq)).  Corner nodes of the
    triangles get seeded N(0, imperf) out-of-plane offsets (interior only);
    mid-side nodes sit at the midpoints of their edges, so every element stays
    planar.  All boundary nodes are clamped.  Follower pressure (ShellLoad,
    AreaUpdate area_update) ramped linearly to p_max over t in [0, 1]."""
    m = Model("shell_plate")
    ng = 2 * nq + 1
    h = L / (2 * nq)
    ii, jj = np.meshgrid(np.arange(ng), np.arange(ng), indexing="xy")
    xyz = np.zeros((ng * ng, 3))
    xyz[:, 0] = (ii * h).ravel()
    xyz[:, 1] = (jj * h).ravel()
    rng = np.random.default_rng(seed)
    z = np.zeros((ng, ng))  # z[j, i]
    zc = rng.normal(0.0, imperf, (nq + 1, nq + 1)) if imperf > 0 else np.zeros((nq + 1, nq + 1))
    zc[0, :] = zc[-1, :] = zc[:, 0] = zc[:, -1] = 0.0
    z[0::2, 0::2] = zc
    z[1::2, 0::2] = 0.5 * (zc[:-1, :] + zc[1:, :])
    z[0::2, 1::2] = 0.5 * (zc[:, :-1] + zc[:, 1:])
    z[1::2, 1::2] = 0.5 * (zc[:-1, :-1] + zc[1:, 1:])  # on the diagonal (2p,2q)-(2p+2,2q+2)
    xyz[:, 2] = z.ravel()
    m.add_nodes(xyz)

    def nid(i, j):
        return j * ng + i

    elems = []
    for q in range(nq):
        for p in range(nq):
            i0, j0 = 2 * p, 2 * q
            a = [nid(i0, j0), nid(i0 + 2, j0), nid(i0 + 2, j0 + 2), nid(i0 + 1, j0), nid(i0 + 2, j0 + 1),
                 nid(i0 + 1, j0 + 1)]
            b = [nid(i0, j0), nid(i0 + 2, j0 + 2), nid(i0, j0 + 2), nid(i0 + 1, j0 + 1), nid(i0 + 1, j0 + 2),
                 nid(i0, j0 + 1)]
            for c6 in (a, b):
                m.add_shell(c6, E, nu, rho, t)
                elems.append(len(elems))
    for j in range(ng):
        for i in range(ng):
            if i == 0 or j == 0 or i == ng - 1 or j == ng - 1:
                m.fix_dofs(nid(i, j), range(6))
    if ramp == "cubic":
        ts = np.linspace(0.0, 1.0, n_steps + 1)
        vals = np.zeros((n_steps + 1, 6))
        vals[:, 0] = p_max * ts ** 3
    else:
        ts = np.array([0.0, 1.0])
        vals = np.array([[0.0] * 6, [p_max, 0, 0, 0, 0, 0]])
    m.add_load(abi.GF_LOAD_SHELL_PRESSURE, elems, ts, vals, flags=area_update)
    dt = 1.0 / n_steps
    Dflex = E * t ** 3 / (12 * (1 - nu * nu))
    # residuals judged against 1% of the total pressure resultant (above the FP64
    # noise floor of the membrane forces)
    fref = 0.01 * abs(p_max) * L * L
    p = StaticParams(end_time=1.0, time_step=dt, max_time_step=dt, min_time_step=dt / 1024.0,
                     max_it=40, min_it=2, conv_increase=4, inc_factor=1.0,
                     force_tol=tol, moment_tol=tol, disp_tol=tol, rot_tol=tol,
                     force_min=max(fref, 1e-6), moment_min=max(fref * L, 1e-6), disp_min=1e-6, rot_min=1e-6)
    center = nid(nq, nq)
    info = {"center": center, "D": Dflex, "ng": ng, "n_free_expected": None}
    return m, p, info


def plate_request(rng, steps=15):
    """One plate request of the mixed serving stream: clamped Shell_1 plate with a
    mesh of nq x nq quads, nq uniform over {16, 24, 32, 48, 64}, thickness
    U[2, 10] mm, follower pressure ramped so that the centre deflection reaches
    about U[5, 25] thicknesses (the pressure of the 64x64, 5 mm plate that gives
    20 thicknesses, scaled with the membrane law p ~ E t^4 (w/t)^3 / L^4), seeded
    imperfections.  Returns (model, params, info)."""
    nq = int(rng.choice([16, 24, 32, 48, 64]))
    t = rng.uniform(0.002, 0.010)
    k = rng.uniform(5.0, 25.0)
    p_max = 1.7e7 * (k / 20.0) ** 3 * (t / 0.005) ** 4
    seed = int(rng.integers(1 << 30))
    m, p, info = shell_plate(nq=nq, t=t, p_max=p_max, n_steps=steps, seed=seed)
    info = dict(info, nq=nq, t=t, target_ratio=k)
    return m, p, info


def riser_request(rng, n_elem=None, steps=10):
    """One request of the catenary serving stream: Beam_1 riser with E ~ log-U[50, 5000]
    elements, horizontal span U[200, 2000] m (vertical 2/3 of it), sag ratio U[0.2, 0.6]
    (arc length ~ chord (1 + 8/3 sag^2)), OD U[0.1, 0.4] m, wall U[0.05, 0.1] OD, steel
    rho = 7,850 kg/m^3 self-weight, N(0, 500 N) lateral nodal loads, `steps` load steps."""
    if n_elem is None:
        n_elem = int(round(math.exp(rng.uniform(math.log(50.0), math.log(5000.0)))))
    span = rng.uniform(200.0, 2000.0)
    sag = rng.uniform(0.2, 0.6)
    od = rng.uniform(0.1, 0.4)
    wall = rng.uniform(0.05, 0.1) * od
    A = tube_section(od, od - 2.0 * wall)[0]
    w = 7850.0 * 9.81 * A
    seed = int(rng.integers(0, 2**31 - 1))
    return catenary(n_elem=n_elem, span_h=span, span_v=2.0 * span / 3.0, arc_ratio=1.0 + 8.0 / 3.0 * sag * sag,
                    De=od, Di=od - 2.0 * wall, w=w, seed=seed, n_steps=steps)
