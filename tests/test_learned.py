"""Learned troubled-cell sensor, host side: the modal features against the oracle's Persson
sensor, bfloat16 rounding, the packed weight layout the device reads, and the shipped weights
against the Persson sigma on states the oracle integrates."""
import math

import numpy as np
import pytest

from frforge_amd.capturing import get_capturing_method, list_methods
from frforge_amd.learned import (DEFAULT_WEIGHTS, MLPWeights, PACKED_BYTES, LearnedSensorMethod,
                                 bf16_round, element_modes, features, persson_sigma_from_modes)

from helpers import EQ, make_case


def test_modes_reproduce_oracle_persson_sensor():
    """The host modal coefficients (k = a + 4 b) give the oracle's Persson sigma: pins the mode
    ordering and the high-mode set the features are built from."""
    st, _, orc = make_case(3, 12, 10, bc="transmissive", capturing="persson_br0", ic="riemann", seed=3,
                           noise=0.01)
    s = persson_sigma_from_modes(element_modes(st.u, st.ops.V_legendre))
    np.testing.assert_allclose(s, orc.sense(st.u), rtol=0, atol=1e-13)
    assert s.max() > 0.05


def test_modes_of_pure_legendre_fields():
    """u = P_a(xi) P_b(eta) (normalised as V) has the single mode (a, b) = 1."""
    st, _, _ = make_case(3, 2, 2, bc="periodic", ic="vortex")
    V = st.ops.V_legendre
    for a, b in [(0, 0), (3, 0), (1, 2), (3, 3)]:
        u = np.zeros_like(st.u, order="F")
        u[:, :, :, 0] = np.einsum("i,j->ij", V[:, a], V[:, b])[:, :, None]
        m = element_modes(u, V)
        want = np.zeros(16)
        want[a + 4 * b] = 1.0
        np.testing.assert_allclose(m[0], want, atol=1e-12)


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -9, -2.5, 0.0, 1e-30], np.float32)
    got = bf16_round(x)
    np.testing.assert_array_equal(got[:5], np.array([1.0, 1.0, 1.0 + 2.0 ** -7, -2.5, 0.0], np.float32))
    assert abs(got[5] - 1e-30) / 1e-30 < 2.0 ** -8


def test_packed_layout():
    w = MLPWeights.init(3)
    raw = w.pack()
    assert len(raw) == PACKED_BYTES == 4368
    blocks = np.frombuffer(raw[:4096], dtype=np.uint16).reshape(8, 64, 4)
    f = lambda bits: (bits.astype(np.uint32) << 16).view(np.float32)
    spec = [(w.W1, 0, 0), (w.W1, 16, 0), (w.W2, 0, 0), (w.W2, 16, 0), (w.W2, 0, 16), (w.W2, 16, 16),
            (w.W3, 0, 0), (w.W3, 0, 16)]
    for m, (W, n0, k0) in enumerate(spec):
        for lane in (0, 5, 17, 38, 63):
            for r in range(4):
                k, n = 4 * (lane // 16) + r, lane % 16
                want = W[n0 + n, k0 + k] if n0 + n < W.shape[0] else 0.0
                assert f(blocks[m, lane, r]) == bf16_round(np.float32(want)), (m, lane, r)
    tail = np.frombuffer(raw[4096:], dtype=np.float32)
    np.testing.assert_array_equal(tail[:32], w.b1)
    np.testing.assert_array_equal(tail[64:65], w.b3)


def test_bf16_inference_tracks_fp32():
    w = MLPWeights.load()
    x = features(np.random.default_rng(1).standard_normal((2000, 16)))
    full = np.clip(w.forward_fp32(x)[4], 0, 1)
    assert np.max(np.abs(w.predict_bf16(x) - full)) < 0.03


def test_shipped_weights_against_persson():
    """On oracle-integrated four-quadrant states the network's sigma stays close to the Persson
    sigma it was trained on (trained on device runs of other seeds and sizes)."""
    w = MLPWeights.load(DEFAULT_WEIGHTS)
    assert w.meta["val_mae"] < 0.02
    st, m, orc = make_case(3, 24, 24, bc="transmissive", capturing="persson_br0", ic="riemann", seed=9)
    _, _, _, u = orc.run(st.u, 0.0, 0.02, cfl=0.08)
    modes = element_modes(u, st.ops.V_legendre)
    s_p = persson_sigma_from_modes(modes)
    s_l = w.predict_bf16(features(modes))
    err = np.abs(s_l - s_p)
    print(f"learned vs Persson sigma on 24^2 seed 9: MAE {err.mean():.3e}, max {err.max():.3e}, "
          f"Persson max {s_p.max():.3f}")
    assert err.mean() < 0.03
    hot = s_p > 0.5
    assert not hot.any() or np.mean(s_l[hot] > 0.25) > 0.8


def test_registry_entry():
    assert "learned_persson" in list_methods()
    m = get_capturing_method("learned_persson")
    assert isinstance(m, LearnedSensorMethod)
    assert m.device_capturing()[0] == "persson_br0"
    assert m.method_params()["method"] == "learned_persson"
