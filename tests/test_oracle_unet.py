"""The U-Net oracle's backward pass pinned by central finite differences in float64, and its
building blocks by their definitions (a convolution with a delta kernel shifts, pooling and
upsampling are adjoint up to the factor 4, bf16 rounding is round-to-nearest-even)."""
import numpy as np
import pytest

from oracle import unet_oracle as uo


def _loss(img, W, b, proj):
    y, tape = uo.forward(img, W, b)
    return float(np.sum(y * proj)), tape


@pytest.mark.parametrize('seed', [0, 1])
def test_backward_matches_finite_differences(seed):
    rng = np.random.default_rng(seed)
    H = Wd = 16
    img = rng.standard_normal((H, Wd))
    W, b = uo.init_params(seed)
    b = rng.standard_normal(b.shape) * 0.05
    proj = rng.standard_normal((H, Wd))
    L0, tape = _loss(img, W, b, proj)
    gW, gb, gimg = uo.backward(proj, tape)
    eps = 1e-6
    for _ in range(12):
        i, k, o, c = rng.integers(14), rng.integers(9), rng.integers(16), rng.integers(16)
        Wp = W.copy(); Wp[i, k, o, c] += eps
        Wm = W.copy(); Wm[i, k, o, c] -= eps
        fd = (_loss(img, Wp, b, proj)[0] - _loss(img, Wm, b, proj)[0]) / (2 * eps)
        assert abs(fd - gW[i, k, o, c]) <= 1e-6 * (abs(fd) + 1e-3), (i, k, o, c, fd, gW[i, k, o, c])
    for _ in range(6):
        i, o = rng.integers(14), rng.integers(16)
        bp = b.copy(); bp[i, o] += eps
        bm = b.copy(); bm[i, o] -= eps
        fd = (_loss(img, W, bp, proj)[0] - _loss(img, W, bm, proj)[0]) / (2 * eps)
        assert abs(fd - gb[i, o]) <= 1e-6 * (abs(fd) + 1e-3)
    for _ in range(6):
        h, w = rng.integers(H), rng.integers(Wd)
        ip = img.copy(); ip[h, w] += eps
        im = img.copy(); im[h, w] -= eps
        fd = (_loss(ip, W, b, proj)[0] - _loss(im, W, b, proj)[0]) / (2 * eps)
        assert abs(fd - gimg[h, w]) <= 1e-6 * (abs(fd) + 1e-3)


def test_building_blocks():
    rng = np.random.default_rng(3)
    x = rng.standard_normal((8, 8, 16))
    w = np.zeros((9, 16, 16))
    w[5] = np.eye(16)                      # tap (dy, dx) = (0, +1)
    y = uo.conv(x, w, np.zeros(16))
    assert np.array_equal(y[:, :-1], x[:, 1:]) and np.all(y[:, -1] == 0)
    g = rng.standard_normal((4, 4, 16))
    # <pool(x), g> = <x, pool_back(g)>, <up(g), x> = <g, up_back(x)>
    assert np.isclose(np.sum(uo.pool(x) * g), np.sum(x * uo.pool_back(g)))
    assert np.isclose(np.sum(uo.up(g) * x), np.sum(g * uo.up_back(x)))
    # bf16 keeps 7 fraction bits: the spacing at 1 is 2^-7
    v = np.array([1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -8, 1 + 1.5 * 2 ** -8, 1 + 2 ** -9, -3.14159, 0.0])
    r = uo.bf16_round(v)
    assert list(r[:5]) == [1.0, 1.0, 1 + 2 ** -6, 1 + 2 ** -7, 1.0]   # ties go to the even neighbour
    assert r[5] == -3.140625 and r[6] == 0.
