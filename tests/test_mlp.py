"""Schedule-selector network: the numpy oracle pinned by finite differences,
then the GPU forward/backward kernels against it, training, and the selector."""
import numpy as np
import pytest

import mlp_oracle as M

rng = np.random.default_rng(0)


def rand_params(scale=0.3):
    return (rng.standard_normal(M.N_PARAMS) * scale).astype(np.float32)


def bf16_exact(p):
    return M.bf16_round(p)


# ---------------------------------------------------------------- oracle pins
def test_bf16_rounding_known_values():
    x = np.array([1.0, 1.00390625, 1.005859375, 1.01171875, -2.5, 3.0e38, np.inf], np.float32)
    # 1 + 2^-8 is halfway between BF16 1.0 and 1.0078125: ties to even -> 1.0
    # 1 + 1.5*2^-8 rounds up; 1 + 3*2^-8 is halfway between 1.0078125 and
    # 1.015625 -> even mantissa 1.015625
    np.testing.assert_array_equal(M.bf16_round(x)[:5], [1.0, 1.0, 1.0078125, 1.015625, -2.5])
    assert np.isinf(M.bf16_round(x)[6])
    v = M.bf16_round(rng.standard_normal(1000).astype(np.float32))
    assert np.all(v.view(np.uint32) & 0xFFFF == 0)


def finite_diff(p, x, y, idx, plant=0, h=1e-6):
    out = []
    for k in idx:
        a = p.astype(np.float64).copy()
        b = a.copy()
        a[k] += h
        b[k] -= h
        out.append((M.loss_and_grad(a, x, y, bf16=False, plant=plant)[0] -
                    M.loss_and_grad(b, x, y, bf16=False, plant=plant)[0]) / (2 * h))
    return np.array(out)


def test_oracle_gradient_matches_finite_differences():
    p = rand_params()
    x = rng.standard_normal((40, 16))
    y = rng.integers(0, 8, 40)
    _, g = M.loss_and_grad(p.astype(np.float64), x, y, bf16=False)
    idx = rng.choice(M.N_PARAMS, 60, replace=False)
    fd = finite_diff(p, x, y, idx)
    np.testing.assert_allclose(g[idx], fd, rtol=1e-5, atol=1e-8)


def test_oracle_planted_mistake_caught():
    p = rand_params()
    x = rng.standard_normal((40, 16))
    y = rng.integers(0, 8, 40)
    _, g_bad = M.loss_and_grad(p.astype(np.float64), x, y, bf16=False, plant=1)
    # W2 block: the dropped ReLU derivative shows up there
    idx = np.arange(16 * 64 + 64, 16 * 64 + 64 + 4096, 37)
    fd = finite_diff(p, x, y, idx)
    assert not np.allclose(g_bad[idx], fd, rtol=1e-4, atol=1e-7)


def test_oracle_softmax_closed_form():
    p = np.zeros(M.N_PARAMS, np.float32)
    x = rng.standard_normal((5, 16))
    loss, g = M.loss_and_grad(p, x, np.zeros(5, int))
    assert loss == pytest.approx(np.log(8))  # all-zero network: uniform over 8 classes
    # b3 gradient = mean(softmax - onehot) = 1/8 - [1,0..0]
    np.testing.assert_allclose(g[-8:], np.full(8, 1 / 8) - np.eye(8)[0])


# ---------------------------------------------------------------- GPU kernels
@pytest.mark.gpu
def test_gpu_forward_and_grad_match_oracle():
    from starplat_amd.schedule import MLP

    m = MLP(seed=1)
    p = rand_params()
    m.set_params(p)
    np.testing.assert_array_equal(m.params(), p)
    for n in (1, 37, 64, 200):
        x = rng.standard_normal((n, 16)).astype(np.float32)
        y = rng.integers(0, 8, n).astype(np.int32)
        ref_logits, _ = M.forward(p, x)
        got = m.forward(x)
        assert np.max(np.abs(got - ref_logits)) <= 1e-4 * max(1.0, np.max(np.abs(ref_logits)))
        loss, g = m.grad(x, y)
        rl, rg = M.loss_and_grad(p, x, y)
        assert loss == pytest.approx(rl, rel=1e-5)
        err = np.max(np.abs(g - rg)) / np.max(np.abs(rg))
        print(f"n={n}: grad rel Linf {err:.2e}")
        assert err <= 1e-4


@pytest.mark.gpu
def test_gpu_training_learns_separable_classes():
    from starplat_amd.schedule import MLP

    n = 1600
    centers = rng.standard_normal((8, 16)) * 3
    y = rng.integers(0, 8, n).astype(np.int32)
    x = (centers[y] + rng.standard_normal((n, 16)) * 0.5).astype(np.float32)
    m = MLP(seed=2)
    l0, _ = m.grad(x, y)
    lf = m.train(x, y, epochs=30, batch=64, lr=3e-3)
    assert lf < 0.2 * l0
    assert np.mean(m.predict(x) == y) > 0.95


@pytest.mark.gpu
def test_selector_beats_fixed_on_synthetic_costs():
    """A schedule-selection task with known costs: the best class depends on
    two features. The trained selector must cut the total cost well below
    always using the fixed class."""
    from starplat_amd.schedule import Selector

    n = 800
    f = rng.standard_normal((n, 16)).astype(np.float32)
    best = (f[:, 0] > 0).astype(int) * 4 + (f[:, 1] > 0.5).astype(int) * 2 + (f[:, 2] > -0.5).astype(int)
    times = np.ones((n, 8)) * 3.0
    times[np.arange(n), best] = 1.0
    sel = Selector()
    sel.fit(f[:600], times[:600], epochs=100)
    pick = sel.choose(f[600:])
    learned = times[600:][np.arange(200), pick].sum()
    fixed = times[600:, 0].sum()
    assert learned < 0.6 * fixed


def test_tie_aware_labels():
    from starplat_amd.schedule import tie_aware_labels

    # class 0 is within 3% of the fastest on three jobs out of four, so it
    # labels every job where it is near-fastest; job 2 has one clear winner
    t = np.array([[1.0, 1.02, 2.0], [1.0, 3.0, 1.01], [2.0, 1.0, 1.5], [1.0, 1.0, 1.0]])
    np.testing.assert_array_equal(tie_aware_labels(t), [0, 0, 1, 0])
    # without ties it is the argmin
    t = np.array([[3.0, 1.0, 2.0], [1.0, 2.0, 3.0], [2.0, 3.0, 1.0]])
    np.testing.assert_array_equal(tie_aware_labels(t), np.argmin(t, axis=1))
    # among equally frequent near-best classes the faster one wins
    t = np.array([[1.0, 1.01]])
    assert tie_aware_labels(t)[0] == 0
