"""Node renumbering helpers for the ordering tests."""
import copy

import numpy as np


def renumber(m, new_of_old):
    """The same model with node i renamed new_of_old[i] (connectivity, loads'
    node sets and constraints follow the nodes)."""
    r = copy.deepcopy(m)
    new_of_old = np.asarray(new_of_old)
    old_of_new = np.empty(m.n_nodes, int)
    old_of_new[new_of_old] = np.arange(m.n_nodes)
    r.xyz = [m.xyz[o] for o in old_of_new]
    r.fix = [m.fix[o] for o in old_of_new]
    r.beam_conn = [[int(new_of_old[n]) for n in c] for c in m.beam_conn]
    r.shell_conn = [[int(new_of_old[n]) for n in c] for c in m.shell_conn]
    for ld in r.loads:
        if int(ld["kind"][0]) != 2:  # nodal loads carry node ids; pressure loads element ids
            ld["items"] = new_of_old[ld["items"]].astype(np.int32)
    r._frozen = None
    return r


def corners_first(nq):
    """new_of_old for a shell_plate(nq) mesh numbered like inputs/tutorial05: every
    corner (vertex) node first, then every mid-side node."""
    ng = 2 * nq + 1
    ii, jj = np.meshgrid(np.arange(ng), np.arange(ng), indexing="xy")
    corner = ((ii % 2 == 0) & (jj % 2 == 0)).ravel()
    order = np.concatenate([np.nonzero(corner)[0], np.nonzero(~corner)[0]])
    new_of_old = np.empty(ng * ng, int)
    new_of_old[order] = np.arange(ng * ng)
    return new_of_old
