"""The CPU network restatement pinned by finite differences (float64), and
its bf16 rounding by exact cases."""
import numpy as np

from oracle import mlp_oracle as M


def test_bf16_rounding_is_round_to_nearest_even():
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -7, 1.0 + 3 * 2 ** -8, -2.5, 3.0e38], dtype=np.float32)
    r = M.bf16_round(v)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: to even (1); 1 + 3*2^-8 is a
    # tie between 1 + 2^-7 and 1 + 2^-6: to even (1 + 2^-6)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == np.float32(1 + 2 ** -7) and r[3] == np.float32(1 + 2 ** -6)
    assert r[4] == -2.5
    assert np.all((M.bf16_round(r).view(np.uint32) & 0xFFFF) == 0)


def test_gradients_match_central_differences():
    rng = np.random.default_rng(0)
    dims = (4, 6, 5, 3, 7)
    _, n = M.layout(dims)
    p = (rng.standard_normal(n) * 0.5).astype(np.float32)
    x = rng.standard_normal((5, 4))
    t = rng.standard_normal((5, 7))
    loss, g = M.loss_and_gradients(p, dims, x, t, terms=None)
    h = 1e-3
    num = np.zeros(n)
    pp = p.astype(np.float64)
    for i in range(n):
        # float32 parameters: difference quotients on float64 copies
        a = pp.copy(); a[i] += h
        b = pp.copy(); b[i] -= h
        la = _loss64(a, dims, x, t)
        lb = _loss64(b, dims, x, t)
        num[i] = (la - lb) / (2 * h)
    assert np.max(np.abs(num - g)) < 1e-6 * max(1.0, np.max(np.abs(g)))


def _loss64(p64, dims, x, t):
    lay, _ = M.layout(dims)
    a = np.asarray(x, dtype=np.float64)
    for l, (wo, ws, bo, bn) in enumerate(lay):
        z = a @ p64[wo:wo + ws[0] * ws[1]].reshape(ws) + p64[bo:bo + bn]
        a = np.tanh(z) if l < 3 else z
    return float(np.sum((a - t) ** 2) / a.size)


def test_split_products_are_closer_to_float64_than_bf16():
    rng = np.random.default_rng(1)
    a = rng.standard_normal((16, 64)).astype(np.float32)
    b = rng.standard_normal((64, 8)).astype(np.float32)
    exact = a.astype(np.float64) @ b.astype(np.float64)
    e1 = np.abs(M.matmul(a, b, 1) - exact).max()
    e3 = np.abs(M.matmul(a, b, 3) - exact).max()
    assert e3 < e1 / 100
