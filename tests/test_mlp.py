"""The warm-start network: the CPU oracle pinned by finite differences, and the
gfx950 MFMA forward/backward checked against it."""
import numpy as np
import pytest

from oracle import mlp_oracle as M


def test_bf16_rounding_known_values():
    x = np.array([1.0, 1.00390625, 1.005859375, -2.5, 3.0e-5, 65504.0], np.float32)
    # 1.00390625 = 1 + 2^-8 is a tie at bf16 precision (8 mantissa bits kept with
    # the implicit one: spacing 2^-7) -> rounds to even (1.0); 1.005859375 rounds up
    np.testing.assert_array_equal(M.bf16(x)[:3], [1.0, 1.0, 1.0078125])
    assert M.bf16(x)[3] == -2.5
    assert abs(M.bf16(x)[4] - 3.0e-5) / 3.0e-5 < 2 ** -8
    assert M.bf16(x)[5] == 65536.0


def test_oracle_gradients_match_finite_differences():
    rng = np.random.default_rng(0)
    p = M.init_params(1)
    x = rng.normal(size=(37, 8))
    t = rng.normal(size=37)
    g, loss = M.gradients(p, x, t, bf=False)
    y, _ = M.forward(p, x, bf=False)
    assert loss == pytest.approx(0.5 * np.sum((y - t) ** 2))
    idx = rng.choice(M.N_PARAMS, 60, replace=False)
    idx = np.concatenate([idx, [M.N_PARAMS - 1, 0, 512, 576]])  # b3, W1, b1, W2 starts
    eps = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += eps
        pm[i] -= eps
        fp = M.gradients(pp, x, t, bf=False)[1] / len(t)
        fm = M.gradients(pm, x, t, bf=False)[1] / len(t)
        num = (fp - fm) / (2 * eps)
        assert g[i] == pytest.approx(num, rel=1e-4, abs=1e-9), i


@pytest.mark.gpu
def test_mfma_forward_matches_oracle():
    import pcisph_amd
    rng = np.random.default_rng(3)
    f = pcisph_amd.Fluid()
    p = M.init_params(7)
    f.mlp_set_params(p.astype(np.float32))
    for n in (1, 16, 37, 1000):
        print(f"case mlp_forward n={n}", flush=True)
        x = rng.normal(size=(n, 8)).astype(np.float32)
        y = f.mlp_forward(x)
        ref, _ = M.forward(p.astype(np.float32).astype(np.float64), x)
        np.testing.assert_allclose(y, ref, rtol=2e-3, atol=2e-3 * np.abs(ref).max())


@pytest.mark.gpu
def test_mfma_gradients_match_oracle():
    import pcisph_amd
    rng = np.random.default_rng(4)
    f = pcisph_amd.Fluid()
    p = M.init_params(9).astype(np.float32)
    f.mlp_set_params(p)
    for n in (16, 301, 4096):
        print(f"case mlp_gradients n={n}", flush=True)
        x = rng.normal(size=(n, 8)).astype(np.float32)
        t = rng.normal(size=n).astype(np.float32)
        g, loss = f.mlp_gradients(x, t)
        gr, lr = M.gradients(p.astype(np.float64), x, t)
        assert loss == pytest.approx(lr, rel=2e-3)
        for name, (a, b) in zip([s[0] for s in M.SIZES], zip(np.split(g, np.cumsum([np.prod(s[1]) for s in M.SIZES])[:-1]),
                                                           np.split(gr, np.cumsum([np.prod(s[1]) for s in M.SIZES])[:-1]))):
            scale = np.abs(b).max()
            np.testing.assert_allclose(a, b, atol=1e-2 * scale, err_msg=name)


@pytest.mark.gpu
def test_training_reduces_loss():
    import pcisph_amd
    rng = np.random.default_rng(5)
    f = pcisph_amd.Fluid()
    f.mlp_set_params(M.init_params(2).astype(np.float32))
    x = rng.normal(size=(20000, 8)).astype(np.float32)
    t = (np.tanh(x[:, 0]) + 0.5 * x[:, 2] * x[:, 3]).astype(np.float32)
    f.samples_reserve(len(t))
    f.samples_upload(x, t)
    first = f.mlp_train(1, batch=2048, lr=3e-3)
    last = f.mlp_train(400, batch=2048, lr=3e-3, seed=1)
    print(f"loss {first:.4f} -> {last:.4f}", flush=True)
    assert last < 0.25 * first
