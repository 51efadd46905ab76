"""Beam_1::Mount on the GPU (gf_beam_eval) against the oracle, element by element."""
import numpy as np
import pytest

from giraffe_amd import model as M, runtime

pytestmark = pytest.mark.gpu


def _random_elements(oracle_lib, n, seed):
    rng = np.random.default_rng(seed)
    xs, css, props, ds, gps, ref_K, ref_f = [], [], [], [], [], [], []
    for i in range(n):
        L = rng.uniform(0.05, 5.0)
        p1 = rng.normal(size=3) * 10
        e3 = rng.normal(size=3)
        e3 /= np.linalg.norm(e3)
        x = np.concatenate([p1, p1 + 0.5 * L * e3, p1 + L * e3])
        e1 = np.cross(e3, rng.normal(size=3))
        cs = M.cs_rows(e1 / np.linalg.norm(e1), e3)
        De = rng.uniform(0.05, 0.5)
        A, I1, I2, I12, It = M.tube_section(De, De * rng.uniform(0.5, 0.95))
        prop = np.array([rng.uniform(50e9, 250e9), rng.uniform(0.2, 0.45), 7850.0, A, I1,
                         I2 * rng.uniform(0.5, 1.5), I1 * rng.uniform(-0.2, 0.2), It, rng.uniform(0, 1e5)])
        c, gp = oracle_lib.beam_element(x, cs, prop)
        for _ in range(i % 3):
            gp = oracle_lib.beam_save(c, gp, rng.normal(size=18) * 0.05 * L)
        d = rng.normal(size=18) * 0.05 * rng.uniform(0.0, 1.0)
        K, f = oracle_lib.beam_mount(c, gp, d)
        xs.append(x); css.append(cs); props.append(prop); ds.append(d)
        gps.append(gp[:, :15]); ref_K.append(K); ref_f.append(f)
    return (np.array(xs), np.array(css), np.array(props), np.array(ds), np.array(gps),
            np.array(ref_K), np.array(ref_f))


def test_beam_eval_matches_oracle(oracle_lib):
    x, cs, prop, d, gp, Kr, fr = _random_elements(oracle_lib, 2000, 11)
    K, f = runtime.beam_eval(x, cs, prop, d, gp)
    sK = np.abs(Kr).reshape(len(Kr), -1).max(1)
    sf = np.maximum(np.abs(fr).max(1), 1e-30)
    errK = (np.abs(K - Kr).reshape(len(Kr), -1).max(1) / sK).max()
    errf = (np.abs(f - fr).max(1) / np.maximum(sf, 1e-12 * sK)).max()
    assert errK < 1e-11, errK
    assert errf < 1e-9, errf


def test_beam_eval_reference_state_is_stress_free(oracle_lib):
    x, cs, prop, d, gp, Kr, fr = _random_elements(oracle_lib, 64, 12)
    K, f = runtime.beam_eval(x, cs, prop, np.zeros_like(d), gp)
    # zero increment on a committed state reproduces the oracle's internal force
    for i in range(0, 64, 7):
        c, g0 = oracle_lib.beam_element(x[i], cs[i], prop[i])
        g0[:, :15] = gp[i]
        Ko, fo = oracle_lib.beam_mount(c, g0, np.zeros(18))
        assert np.abs(K[i] - Ko).max() <= 1e-11 * np.abs(Ko).max()
        assert np.abs(f[i] - fo).max() <= 1e-11 * np.abs(Ko).max()
