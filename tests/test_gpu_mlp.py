"""The learned screening form's MLP on the GPU (bf16 MFMA, FP32
accumulation) against the CPU oracle of the same network (double
accumulation over the same bf16-rounded operands)."""
import numpy as np
import pytest

from oracle_bind import mlp_forward as ref_forward, mlp_loss_grad as ref_loss_grad

pytestmark = pytest.mark.gpu

DIN, HID = 241, 256  # 24-orbital features: 48 occupation bits + 4 x 48 one-hot + log|c|


def _data(n, seed):
    rng = np.random.default_rng(seed)
    X = np.zeros((n, DIN), dtype=np.float32)
    X[:, :48] = rng.integers(0, 2, (n, 48))
    for blk in range(4):
        X[np.arange(n), 48 + 48 * blk + rng.integers(0, 48, n)] = 1.0
    X[:, 240] = rng.uniform(-6, 0, n)
    t = rng.normal(-4, 1.5, n).astype(np.float32)
    return X, t


def test_forward_matches_oracle():
    from dice_amd import shci
    P = shci.mlp_init(DIN, HID, seed=3)
    X, _ = _data(300, 1)
    y = shci.mlp_forward(P, DIN, HID, X)
    yr = ref_forward(P, DIN, HID, X, bf16=True)
    assert np.max(np.abs(y - yr)) <= 1e-4 * (1 + np.max(np.abs(yr)))


def test_gradient_matches_oracle():
    """Loss to 1e-5 relative; gradient to 1e-3 in norm and 1e-2 of its largest
    component element-wise (FP32 accumulation order and rare bf16 ties of
    intermediate values are the only differences)."""
    from dice_amd import shci
    P = shci.mlp_init(DIN, HID, seed=4)
    X, t = _data(512, 2)
    loss, g = shci.mlp_loss_grad(P, DIN, HID, X, t)
    lr, gr = ref_loss_grad(P, DIN, HID, X, t, bf16=True)
    assert abs(loss - lr) <= 1e-5 * abs(lr)
    assert np.linalg.norm(g - gr) <= 1e-3 * np.linalg.norm(gr)
    assert np.max(np.abs(g - gr)) <= 1e-2 * np.max(np.abs(gr))
    # every parameter block receives gradient
    n1 = HID * DIN
    blocks = [(0, n1), (n1, n1 + HID), (n1 + HID, n1 + HID + HID * HID)]
    for a, b in blocks:
        assert np.linalg.norm(gr[a:b]) > 0 and np.linalg.norm(g[a:b] - gr[a:b]) <= 2e-3 * np.linalg.norm(gr[a:b])


def test_training_fits_a_teacher_network():
    """Adam on the GPU fits targets produced by a random teacher network of the
    same shape: the loss falls by more than 10x."""
    from dice_amd import shci
    X, _ = _data(16384, 5)
    teacher = shci.mlp_init(DIN, HID, seed=11)
    t = shci.mlp_forward(teacher, DIN, HID, X)
    P0 = shci.mlp_init(DIN, HID, seed=12)
    l0, _ = shci.mlp_loss_grad(P0, DIN, HID, X, t)
    P, first, last = shci.mlp_train(P0, DIN, HID, X, t, epochs=8, batch=1024, lr=1e-3, seed=1)
    l1, _ = shci.mlp_loss_grad(P, DIN, HID, X, t)
    print("loss before %.4g after %.4g (minibatches %.4g -> %.4g)" % (l0, l1, first, last))
    assert l1 < l0 / 10
