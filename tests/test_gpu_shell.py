"""Shell_1 and the wide-band solver on the GPU against the oracle / numpy."""
import numpy as np
import pytest

from giraffe_amd import abi, model as M, runtime

pytestmark = pytest.mark.gpu

PROP = [70e9, 0.33, 2700.0, 0.005]
DEV_GP = list(range(0, 9)) + list(range(15, 27))  # oracle GP layout -> device layout


def _random_shells(oracle_lib, n, seed):
    rng = np.random.default_rng(seed)
    xs, props, ds, gps, Ks, fs = [], [], [], [], [], []
    for i in range(n):
        size = rng.uniform(0.01, 0.5)
        x = np.array([0, 0, 0, 1, 0, 0, 0, 1, 0, 0.5, 0, 0, 0.5, 0.5, 0, 0, 0.5, 0.0]) * size
        x[3:5] += rng.normal(size=2) * 0.1 * size
        x[6:8] += rng.normal(size=2) * 0.1 * size
        x[9:12] = 0.5 * (x[0:3] + x[3:6])
        x[12:15] = 0.5 * (x[3:6] + x[6:9])
        x[15:18] = 0.5 * (x[6:9] + x[0:3])
        R = np.linalg.qr(rng.normal(size=(3, 3)))[0]
        x = (x.reshape(6, 3) @ R.T + rng.normal(size=3)).ravel()
        prop = [rng.uniform(50e9, 210e9), rng.uniform(0.2, 0.45), 2700.0, rng.uniform(0.001, 0.02)]
        el = oracle_lib.ShellElement(x, prop)
        for _ in range(i % 3):
            d = np.zeros(36)
            d[el.DOF] = rng.normal(size=27) * 2e-3 * size
            el.save(d)
        d = np.zeros(36)
        d[el.DOF] = rng.normal(size=27) * 2e-3 * size
        K, f = el.mount(d)
        xs.append(x); props.append(prop); ds.append(d); gps.append(el.gp[:, DEV_GP].copy()); Ks.append(K); fs.append(f)
    return [np.array(v) for v in (xs, props, ds, gps, Ks, fs)]


def test_shell_eval_matches_oracle(oracle_lib):
    x, prop, d, gp, Kr, fr = _random_shells(oracle_lib, 600, 5)
    K, f = runtime.shell_eval(x, prop, d, gp)
    sK = np.abs(Kr).reshape(len(Kr), -1).max(1)
    errK = (np.abs(K - Kr).reshape(len(Kr), -1).max(1) / sK).max()
    errf = (np.abs(f - fr).max(1) / np.maximum(np.abs(fr).max(1), 1e-12 * sK)).max()
    assert errK < 1e-11, errK
    assert errf < 1e-9, errf


@pytest.mark.parametrize("n,hb", [(50, 7), (300, 40), (1000, 100), (4000, 300), (1500, 63), (1500, 64),
                                  (3000, 1346), (2100, 1500)])
def test_bandsystem_matches_dense(n, hb):
    rng = np.random.default_rng(n + hb)
    w = 2 * hb + 1
    rows = rng.normal(size=(n, w))
    for i in range(n):
        rows[i, hb] = 2.0 * (2 * hb + 1) + rng.uniform()  # diagonal dominance: no pivoting needed
    A = np.zeros((n, n))
    for i in range(n):
        for k in range(w):
            c = i - hb + k
            if 0 <= c < n:
                A[i, c] = rows[i, k]
    b = rng.normal(size=n)
    x, fail = runtime.bandsystem(rows, hb, b)
    assert fail == 0
    xr = np.linalg.solve(A, b)
    assert np.abs(x - xr).max() <= 1e-12 * np.abs(xr).max() * n


def test_bandsystem_reports_singular_block():
    """A diagonal block that is entirely zero, or a non-finite entry, fails the solve."""
    rows = np.zeros((200, 3))
    rows[:, 1] = 1.0
    rows[64:128, :] = 0.0
    x, fail = runtime.bandsystem(rows, 1, np.ones(200))
    assert fail == 1
    rows[64:128, 1] = 1.0
    rows[150, 1] = np.nan
    x, fail = runtime.bandsystem(rows, 1, np.ones(200))
    assert fail == 1


def test_bandsystem_perturbs_tiny_pivot():
    """A pivot below 1e-13 of its row's original diagonal (of the block's largest
    entry where that diagonal is zero) is replaced by that size with its sign and
    the solve goes on (PARDISO's static-pivoting default for mtype 11).  A zero
    diagonal entry in a block whose largest entry is 2 becomes 2e-13; a tiny
    but non-zero diagonal entry is a true pivot and stays."""
    n = 130
    rows = np.zeros((n, 3))
    rows[:, 1] = 2.0
    rows[10, 1] = 0.0
    rows[40, 1] = 1e-20
    # rows 100, 101: [[1, 1], [1, 1 + 1e-15]], whose second pivot cancels to ~1e-15
    rows[100] = [0.0, 1.0, 1.0]
    rows[101] = [1.0, 1.0 + 1e-15, 0.0]
    b = np.ones(n)
    x, fail = runtime.bandsystem(rows, 1, b)
    assert fail == 0 and np.isfinite(x).all()
    exp = np.full(n, 0.5)
    exp[10] = 1.0 / 2e-13
    exp[40] = 1e20
    np.testing.assert_allclose(np.delete(x, [100, 101]), np.delete(exp, [100, 101]), rtol=1e-12)
    # the perturbed pivot (1e-13 instead of ~1e-15) still satisfies the first row
    assert abs(x[100] + x[101] - 1.0) <= 1e-12 * max(1.0, abs(x[100]))


def test_bandsystem_with_interchange_free_blocks_of_mixed_scale():
    """Rows scaled over 12 orders of magnitude (translations and rotations of a
    stiffness matrix): the block inverses keep the solve accurate."""
    rng = np.random.default_rng(5)
    n, hb = 700, 90
    w = 2 * hb + 1
    rows = rng.normal(size=(n, w))
    for i in range(n):
        rows[i, hb] = 4.0 * w
    A = np.zeros((n, n))
    for i in range(n):
        for k in range(w):
            c = i - hb + k
            if 0 <= c < n:
                A[i, c] = rows[i, k]
    s = 10.0 ** rng.uniform(-6, 6, n)
    A = s[:, None] * A * s[None, :]
    rows2 = np.array([[A[i, i - hb + k] if 0 <= i - hb + k < n else 0.0 for k in range(w)] for i in range(n)])
    b = rng.normal(size=n)
    x, fail = runtime.bandsystem(rows2, hb, b)
    assert fail == 0
    xr = np.linalg.solve(A, b)

    def backward_error(v):  # componentwise: |A v - b| / (|A| |v| + |b|), worst row
        return (np.abs(A @ v - b) / (np.abs(A) @ np.abs(v) + np.abs(b))).max()

    # as accurate as LAPACK's partially pivoted solve of the same matrix
    assert backward_error(x) <= max(1e-14, 100.0 * backward_error(xr)), (backward_error(x), backward_error(xr))


def test_plate_assembly_matches_oracle(oracle_lib):
    m, p, info = M.shell_plate(nq=4, p_max=2e5)
    m.gravity = (0.0, 0.0, -9.81)
    m.gravity_on = 1
    om = oracle_lib.OracleModel(m)
    om.begin_solution(0.0, 1.0)
    dm = runtime.DeviceModel(m)
    dm.set_window(0.0, 1.0)
    rng = np.random.default_rng(3)
    for t_last, scale in [(0.0, 0.0), (0.4, 1e-4), (0.7, 1e-3)]:
        om.begin_step(t_last, 0.1)
        dm.begin_step(t_last, 0.1)
        d = rng.normal(size=(m.n_nodes, 6)) * scale
        om.write_disp(d)
        dm.write_disp(d)
        ro, rg = om.assemble(), dm.assemble()
        Ko, Kg = om.tangent_dense(), dm.tangent_dense()
        assert np.abs(Kg - Ko).max() <= 1e-12 * np.abs(Ko).max()
        Ro, Rg = om.residual(), dm.residual()
        assert np.abs(Rg - Ro).max() <= 1e-11 * max(np.abs(Ro).max(), 1e-12 * np.abs(Ko).max())


@pytest.mark.parametrize("nq", [4, 8])
def test_plate_static_matches_oracle(oracle_lib, nq):
    m, p, info = M.shell_plate(nq=nq)
    om = oracle_lib.OracleModel(m)
    ro, io = om.static_solve(p)
    dm = runtime.DeviceModel(m)
    rg, ig = dm.static_solve(p)
    assert rg.aborted == ro.aborted == 0
    assert len(ig) == len(io) and np.abs(ig - io).max() <= 1
    no, ng = om.nodes(), dm.nodes()
    x0 = m.freeze()["xyz"]
    scale = np.abs(no[:, :3] - x0).max()
    assert np.abs(ng[:, :3] - no[:, :3]).max() <= 1e-8 * scale
    assert np.abs(ng[:, 3:] - no[:, 3:]).max() <= 1e-8 * max(np.abs(no[:, 3:]).max(), 1e-12)
    assert rg.max_rel_residual <= 1e-10
    w = ng[info["center"], 2]
    assert -w / 0.005 > 15.0  # about 20 thicknesses at the centre
