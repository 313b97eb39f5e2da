"""The CPU oracle of the candidate-ranking MLP (learned screening form),
pinned by central finite differences of its own double-precision forward
pass, and the bf16 rounding it shares with the GPU path."""
import numpy as np

from oracle_bind import mlp_forward, mlp_loss_grad


def _params(din, h, rng):
    n = h * din + h + h * h + h + h + 1
    return (rng.normal(0, 0.4, n)).astype(np.float32)


def test_gradient_matches_finite_differences():
    """Every parameter's analytic gradient (double mode) against a central
    difference of the loss with step 1e-3 (float32 parameters, so the
    perturbed values are exactly representable only approximately: compare
    with the actual perturbation)."""
    rng = np.random.default_rng(1)
    din, h, n = 11, 6, 7
    P = _params(din, h, rng)
    X = rng.normal(0, 1, (n, din)).astype(np.float32)
    t = rng.normal(0, 1, n).astype(np.float32)
    loss, g = mlp_loss_grad(P, din, h, X, t, bf16=False)
    assert loss > 0
    for q in range(len(P)):
        Pp, Pm = P.copy(), P.copy()
        Pp[q] += np.float32(1e-3)
        Pm[q] -= np.float32(1e-3)
        step = float(Pp[q]) - float(Pm[q])
        fd = (mlp_loss_grad(Pp, din, h, X, t, bf16=False)[0] - mlp_loss_grad(Pm, din, h, X, t, bf16=False)[0]) / step
        assert abs(fd - g[q]) <= 1e-6 + 1e-5 * abs(g[q]), (q, fd, g[q])


def test_loss_is_half_mean_square_error():
    rng = np.random.default_rng(2)
    din, h, n = 9, 5, 13
    P = _params(din, h, rng)
    X = rng.normal(0, 1, (n, din)).astype(np.float32)
    t = rng.normal(0, 1, n).astype(np.float32)
    y = mlp_forward(P, din, h, X, bf16=False)
    loss, _ = mlp_loss_grad(P, din, h, X, t, bf16=False)
    assert abs(loss - 0.5 * np.mean((y - t.astype(np.float64)) ** 2)) < 1e-12


def test_forward_closed_form_single_unit():
    """din = 1, hidden = 1: y = w3 gelu(w2 gelu(w1 x + b1) + b2) + b3 with
    gelu(z) = z Phi(z), evaluated independently with math.erf."""
    import math

    def gelu(z):
        return 0.5 * z * (1 + math.erf(z / math.sqrt(2)))
    P = np.array([0.7, -0.2, 1.3, 0.1, -0.9, 0.25], dtype=np.float32)  # w1 b1 w2 b2 w3 b3
    X = np.array([[0.5], [-1.5], [2.0]], dtype=np.float32)
    y = mlp_forward(P, 1, 1, X, bf16=False)
    w1, b1, w2, b2, w3, b3 = [float(v) for v in P]
    for x, yy in zip(X[:, 0], y):
        ref = w3 * gelu(w2 * gelu(w1 * float(x) + b1) + b2) + b3
        assert abs(ref - yy) < 1e-14


def test_bf16_rounding_of_matmul_operands():
    """bf16 keeps 8 significand bits: with din = hidden = 1 and weights that
    are exact in bf16, an input of 1 + 2^-9 rounds to 1 (ties to even) and
    1 + 3 * 2^-9 to 1 + 2^-7."""
    P = np.array([1.0, 0.0, 1.0, 0.0, 1.0, 0.0], dtype=np.float32)
    X = np.array([[1 + 2 ** -9], [1 + 3 * 2 ** -9]], dtype=np.float32)
    import math

    def gelu(z):
        return 0.5 * z * (1 + math.erf(z / math.sqrt(2)))
    y = mlp_forward(P, 1, 1, X, bf16=True)
    # second product rounds gelu(x') to bf16 too
    def rb(v):
        a = np.array([v], dtype=np.float32).view(np.uint32)[0]
        a = (int(a) + 0x7FFF + ((int(a) >> 16) & 1)) & 0xFFFF0000
        return float(np.array([a], dtype=np.uint32).view(np.float32)[0])
    for x, yy, xr in zip(X[:, 0], y, [1.0, 1 + 2 ** -7]):
        assert rb(float(x)) == xr
        assert abs(yy - gelu(rb(gelu(xr)))) < 1e-14
