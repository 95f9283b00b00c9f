"""Pins the corrector network oracle (oracle/corr_oracle.py).

* conv3x3 equals a loop-level zero-padded convolution;
* conv3x3_dgrad is the exact adjoint of conv3x3 (<Ax, y> = <x, A^T y>);
* conv3x3_wgrad is the exact derivative of <conv3x3(x; w), dy> in w;
* exact-mode parameter gradients of the loss match central finite differences;
* bf16 rounding is round-to-nearest-even on the float32 bits.
"""
import numpy as np
import pytest

from oracle import corr_oracle as C


def test_conv_matches_loops():
    rng = np.random.default_rng(0)
    x = rng.standard_normal((2, 5, 6, 3))
    w = rng.standard_normal((4, 3, 3, 3))
    b = rng.standard_normal(4)
    y = C.conv3x3(x, w, b)
    ref = np.zeros_like(y)
    for n in range(2):
        for i in range(5):
            for j in range(6):
                for o in range(4):
                    s = b[o]
                    for c in range(3):
                        for ky in range(3):
                            for kx in range(3):
                                ii, jj = i + ky - 1, j + kx - 1
                                if 0 <= ii < 5 and 0 <= jj < 6:
                                    s += w[o, c, ky, kx] * x[n, ii, jj, c]
                    ref[n, i, j, o] = s
    assert np.allclose(y, ref, rtol=1e-12, atol=1e-12)


def test_dgrad_is_adjoint_and_wgrad_is_derivative():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((2, 7, 5, 4))
    w = rng.standard_normal((3, 4, 3, 3))
    dy = rng.standard_normal((2, 7, 5, 3))
    lhs = np.sum(C.conv3x3(x, w) * dy)
    assert lhs == pytest.approx(np.sum(x * C.conv3x3_dgrad(dy, w)), rel=1e-12)
    # <conv(x; w), dy> is linear in w: its gradient is wgrad(x, dy)
    assert lhs == pytest.approx(np.sum(w * C.conv3x3_wgrad(x, dy)), rel=1e-12)


@pytest.mark.parametrize("band", [0, 1])
def test_exact_gradients_match_finite_differences(band):
    rng = np.random.default_rng(2)
    H, W = 6, 5
    raw = rng.standard_normal((2, H, W, C.C_IN))
    tgt = rng.standard_normal((2, H, W, 3)) * 0.1
    flat = rng.standard_normal(C.n_params()) * 0.2
    shift = np.zeros(C.C_IN)
    scale = np.ones(C.C_IN)
    osc = np.array([0.5, 2.0, 1.0])
    loss, g = C.loss_and_grads(flat, raw, tgt, shift, scale, osc, exact=True, band=band)
    # features are computed in float32 then widened: perturb parameters only
    idx = rng.choice(C.n_params(), 40, replace=False)
    eps = 1e-6
    for k in idx:
        fp = flat.copy()
        fm = flat.copy()
        fp[k] += eps
        fm[k] -= eps
        lp, _ = C.loss_and_grads(fp, raw, tgt, shift, scale, osc, exact=True, band=band)
        lm, _ = C.loss_and_grads(fm, raw, tgt, shift, scale, osc, exact=True, band=band)
        fd = (lp - lm) / (2 * eps)
        assert g[k] == pytest.approx(fd, rel=1e-4, abs=1e-9), k


def test_band_loss_counts_only_the_corrected_region():
    """band = 2 on a 7 x 9 image: the loss is the squared misfit summed over
    rows 2..4 and columns 2..6 only, divided by every point and channel."""
    rng = np.random.default_rng(5)
    H, W = 7, 9
    raw = rng.standard_normal((1, H, W, C.C_IN))
    tgt = rng.standard_normal((1, H, W, 3))
    flat = rng.standard_normal(C.n_params()) * 0.2
    shift, scale, osc = np.zeros(C.C_IN), np.ones(C.C_IN), np.array([0.5, 2.0, 1.0])
    y, _ = C.forward(flat, raw, shift, scale, exact=True)
    d = y - tgt / osc
    want = np.sum(d[0, 2:5, 2:7] ** 2) / d.size
    loss, g = C.loss_and_grads(flat, raw, tgt, shift, scale, osc, exact=True, band=2)
    assert loss == pytest.approx(want, rel=1e-12)
    # a target change outside the region moves neither the loss nor a gradient
    t2 = tgt.copy()
    t2[0, 1, :] += 5.0
    t2[0, :, 7] -= 3.0
    loss2, g2 = C.loss_and_grads(flat, raw, t2, shift, scale, osc, exact=True, band=2)
    assert loss2 == loss and np.array_equal(g, g2)
    assert C.band_mask(H, W, 2).sum() == 3 * 5


def test_bf16_rounding():
    vals = np.array([1.0, 1.00390625, 1.005859375, -3.0e-5, 65504.0, 1e-40], dtype=np.float32)
    r = C.bf16(vals)
    assert r[0] == 1.0
    assert r[1] == 1.0            # exactly halfway between 1 and 1+2^-7: ties to even
    assert r[2] == 1.0078125      # above halfway: rounds up
    assert r[3] == pytest.approx(-3.0e-5, rel=2 ** -8)
    normal = np.abs(vals) >= np.finfo(np.float32).tiny
    assert np.all((np.abs(r - vals) <= np.abs(vals) * 2 ** -8)[normal])
