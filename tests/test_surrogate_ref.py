"""Pins of the surrogate's CPU reference (oracle/surrogate_ref.py), no GPU.

The reference's back-propagation is checked against central finite differences
of its own loss and against torch autograd (float64), its BF16 rounding against
torch's float32 -> bfloat16 conversion and hand-worked ties, and its GELU against
torch's exact-erf GELU.
"""
import numpy as np
import torch

from oracle import surrogate_ref as sr


def test_bf16_rounding_ties_and_torch():
    one = np.float32(1.0)
    ulp = 2.0 ** -7
    x = np.array([1.0, 1 + ulp / 2, 1 + 1.5 * ulp, -(1 + ulp / 2), 3.0e-3, 65504.0], np.float32)
    r = sr.bf16(x)
    assert r[0] == one and r[1] == one                     # tie to even (down)
    assert r[2] == np.float32(1 + 2 * ulp)                 # tie to even (up)
    assert r[3] == -one
    rng = np.random.default_rng(0)
    v = (rng.standard_normal(100000) * np.exp(rng.uniform(-20, 20, 100000))).astype(np.float32)
    t = torch.from_numpy(v).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(sr.bf16(v), t)


def test_gelu_matches_torch_and_derivative():
    z = np.linspace(-8, 8, 2001)
    g = torch.nn.functional.gelu(torch.from_numpy(z)).numpy()
    assert np.max(np.abs(sr.gelu(z) - g)) < 1e-15
    h = 1e-6
    fd = (sr.gelu(z + h) - sr.gelu(z - h)) / (2 * h)
    assert np.max(np.abs(sr.gelu_d(z) - fd)) < 1e-8


def _tiny(seed=1, widths=(7, 6, 5, 7), rows=4):
    rng = np.random.default_rng(seed)
    Ws = [rng.standard_normal((widths[l + 1], widths[l])) / np.sqrt(widths[l]) for l in range(len(widths) - 1)]
    bs = [0.1 * rng.standard_normal(widths[l + 1]) for l in range(len(widths) - 1)]
    f = rng.uniform(0, 1, (rows, widths[0]))
    q = rng.standard_normal((rows, widths[-1]))
    return Ws, bs, f, q


def test_backprop_matches_finite_differences():
    Ws, bs, f, q = _tiny()
    loss, dWs, dbs, _ = sr.forward_backward(Ws, bs, f, q, 1.3, 0.7, bf16_path=False)
    eps = 1e-6
    for l in range(len(Ws)):
        for (i, j) in [(0, 0), (1, 2), (Ws[l].shape[0] - 1, Ws[l].shape[1] - 1)]:
            Wp = [W.copy() for W in Ws]
            Wm = [W.copy() for W in Ws]
            Wp[l][i, j] += eps
            Wm[l][i, j] -= eps
            fd = (sr.forward_backward(Wp, bs, f, q, 1.3, 0.7, False)[0] -
                  sr.forward_backward(Wm, bs, f, q, 1.3, 0.7, False)[0]) / (2 * eps)
            assert abs(fd - dWs[l][i, j]) < 1e-8 * max(1.0, abs(fd)), (l, i, j, fd, dWs[l][i, j])
        for i in (0, len(bs[l]) - 1):
            bp = [b.copy() for b in bs]
            bm = [b.copy() for b in bs]
            bp[l][i] += eps
            bm[l][i] -= eps
            fd = (sr.forward_backward(Ws, bp, f, q, 1.3, 0.7, False)[0] -
                  sr.forward_backward(Ws, bm, f, q, 1.3, 0.7, False)[0]) / (2 * eps)
            assert abs(fd - dbs[l][i]) < 1e-8 * max(1.0, abs(fd))


def test_backprop_matches_torch_autograd():
    Ws, bs, f, q = _tiny(seed=5, widths=(9, 8, 8, 8, 9), rows=6)
    loss, dWs, dbs, y = sr.forward_backward(Ws, bs, f, q, 2.0, 0.5, bf16_path=False)
    tW = [torch.tensor(W, requires_grad=True) for W in Ws]
    tb = [torch.tensor(b, requires_grad=True) for b in bs]
    h = torch.tensor(f) / 2.0
    for l in range(len(Ws)):
        z = h @ tW[l].T + tb[l]
        h = torch.nn.functional.gelu(z) if l + 1 < len(Ws) else z
    tl = torch.mean((h - torch.tensor(q) / 0.5) ** 2)
    tl.backward()
    assert abs(tl.item() - loss) < 1e-14
    for l in range(len(Ws)):
        assert np.max(np.abs(tW[l].grad.numpy() - dWs[l])) < 1e-13
        assert np.max(np.abs(tb[l].grad.numpy() - dbs[l])) < 1e-13


def test_projection_removes_moments_and_is_idempotent():
    rng = np.random.default_rng(2)
    n = 200
    u, v, w = rng.uniform(-2, 2, (3, n))
    gw = rng.uniform(0.1, 1.0, n)
    q = rng.standard_normal((3, n))
    p = sr.project_invariants(q, u, v, w, gw)
    psi = np.stack([np.ones(n), u, v, w, u * u + v * v + w * w])
    assert np.max(np.abs((p * gw) @ psi.T)) < 1e-12
    assert np.max(np.abs(sr.project_invariants(p, u, v, w, gw) - p)) < 1e-12
    # a pure invariant combination projects to zero
    c = rng.standard_normal(5) @ psi
    assert np.max(np.abs(sr.project_invariants(c[None], u, v, w, gw))) < 1e-11


def test_adam_bias_correction_first_step():
    # with m = v = 0 the first bias-corrected step moves every weight by lr*sign(g)
    g = np.array([0.3, -2.0, 1e-3])
    W, m, v = sr.adam(np.zeros(3), np.zeros(3), np.zeros(3), g, 1, lr=0.01)
    assert np.allclose(W, -0.01 * np.sign(g), rtol=1e-4)
