"""CPU ORACLE -- test infrastructure only (never imported by the product).

The learned Poisson preconditioner's network (no counterpart in the reference, whose Poisson
solve is SuperLU): a 4-level U-Net with 16 channels, 3x3 convolutions with zero padding, ReLU,
2x2 average pooling down, nearest-neighbour upsampling with additive skips, and a final
convolution whose channel 0 is the output.  Images are NHWC float arrays (H, W, 16); the
input image carries its single channel in channel 0.  Weights: [14][9][16 out][16 in] with
tap k = 3 (dy + 1) + (dx + 1), biases [14][16].

    level 0   s0 = relu(c1(relu(c0(x))))
    level 1   s1 = relu(c3(relu(c2(pool(s0)))))
    level 2   s2 = relu(c5(relu(c4(pool(s1)))))
    level 3   u3 = relu(c7(relu(c6(pool(s2)))))
    up 2      u2 = relu(c9(relu(c8(up(u3) + s2))))
    up 1      u1 = relu(c11(relu(c10(up(u2) + s1))))
    up 0      y  = c13(relu(c12(up(u1) + s0)))[..., 0]

`bf16=True` rounds the weights and every stored activation to bfloat16 (round to nearest
even), as the device kernels store them; the arithmetic itself is float64 here (the device
accumulates in FP32).  `backward` is the exact gradient of this forward (bf16 rounding
treated as identity), pinned by finite differences in float64 (tests/test_oracle_unet.py).
"""
import numpy as np

C = 16
NCONV = 14
RELU = [True] * 13 + [False]


def bf16_round(a):
    """Round float values to the nearest bfloat16 (ties to even), returned as float64."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    lsb = (u >> np.uint64(16)) & np.uint64(1)
    u = (u + np.uint64(0x7FFF) + lsb) & np.uint64(0xFFFF0000)
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def init_params(seed=0, scale=1.0):
    """He-style random weights (std sqrt(2 / (9 * 16)) * scale), zero biases."""
    rng = np.random.default_rng(seed)
    W = rng.standard_normal((NCONV, 9, C, C)) * np.sqrt(2. / (9 * C)) * scale
    b = np.zeros((NCONV, C))
    return W, b


def conv(x, w, b):
    """3x3 convolution with zero padding: y[h, w, o] = b[o] + sum_k sum_i w[k, o, i] x[h+dy, w+dx, i]."""
    H, W_, _ = x.shape
    xp = np.zeros((H + 2, W_ + 2, C))
    xp[1:-1, 1:-1] = x
    y = np.broadcast_to(b, (H, W_, C)).copy()
    for k in range(9):
        dy, dx = k // 3 - 1, k % 3 - 1
        y += xp[1 + dy:1 + dy + H, 1 + dx:1 + dx + W_] @ w[k].T
    return y


def conv_back(x, w, gy):
    """Gradients of conv: (gx, gw, gb)."""
    H, W_, _ = x.shape
    xp = np.zeros((H + 2, W_ + 2, C))
    xp[1:-1, 1:-1] = x
    gxp = np.zeros_like(xp)
    gw = np.zeros_like(w)
    for k in range(9):
        dy, dx = k // 3 - 1, k % 3 - 1
        sl = (slice(1 + dy, 1 + dy + H), slice(1 + dx, 1 + dx + W_))
        gw[k] = np.einsum('hwo,hwi->oi', gy, xp[sl])
        gxp[sl] += gy @ w[k]
    return gxp[1:-1, 1:-1], gw, gy.sum(axis=(0, 1))


def pool(x):
    H, W_, _ = x.shape
    return 0.25 * (x[0::2, 0::2] + x[1::2, 0::2] + x[0::2, 1::2] + x[1::2, 1::2])


def pool_back(g):
    return 0.25 * np.repeat(np.repeat(g, 2, axis=0), 2, axis=1)


def up(x):
    return np.repeat(np.repeat(x, 2, axis=0), 2, axis=1)


def up_back(g):
    return g[0::2, 0::2] + g[1::2, 0::2] + g[0::2, 1::2] + g[1::2, 1::2]


def _input(img):
    x = np.zeros(img.shape + (C,))
    x[..., 0] = img
    return x


def forward(img, W, b, bf16=False):
    """img (H, W) with H, W multiples of 8 -> (output (H, W), tape for backward)."""
    rd = bf16_round if bf16 else (lambda a: a)
    Wr = rd(W)
    tape = {}

    def cv(i, x):
        z = conv(x, Wr[i], b[i])
        tape['in%d' % i] = x
        tape['z%d' % i] = z
        return rd(np.maximum(z, 0.)) if RELU[i] else z

    x = rd(_input(img))
    s0 = cv(1, cv(0, x))
    s1 = cv(3, cv(2, rd(pool(s0))))
    s2 = cv(5, cv(4, rd(pool(s1))))
    u3 = cv(7, cv(6, rd(pool(s2))))
    u2 = cv(9, cv(8, rd(up(u3) + s2)))
    u1 = cv(11, cv(10, rd(up(u2) + s1)))
    y = cv(13, cv(12, rd(up(u1) + s0)))
    tape['W'] = Wr
    return y[..., 0], tape


def backward(gout, tape, bf16=False):
    """gout (H, W): dLoss/doutput -> (gW [14][9][16][16], gb [14][16], gimg (H, W)).
    bf16=True rounds every stored gradient image to bfloat16 where the device stores one
    (convolution input gradients, the upsampling adjoint, the pooling adjoint sums); the
    parameter gradients stay unrounded."""
    rd = bf16_round if bf16 else (lambda a: a)
    Wr = tape['W']
    gW = np.zeros_like(Wr)
    gb = np.zeros((NCONV, C))

    def cvb(i, g):
        if RELU[i]:
            g = g * (tape['z%d' % i] > 0)
        gx, gW[i], gb[i] = conv_back(tape['in%d' % i], Wr[i], g)
        return rd(gx)

    g = np.zeros(gout.shape + (C,))
    g[..., 0] = rd(gout)
    gt0 = cvb(12, cvb(13, g))                       # grad of t0 = up(u1) + s0
    gt1 = cvb(10, cvb(11, rd(up_back(gt0))))        # grad of t1 = up(u2) + s1
    gt2 = cvb(8, cvb(9, rd(up_back(gt1))))          # grad of t2 = up(u3) + s2
    gp3 = cvb(6, cvb(7, rd(up_back(gt2))))          # grad of pool(s2)
    gs2 = rd(gt2 + pool_back(gp3))
    gp2 = cvb(4, cvb(5, gs2))
    gs1 = rd(gt1 + pool_back(gp2))
    gp1 = cvb(2, cvb(3, gs1))
    gs0 = rd(gt0 + pool_back(gp1))
    gx = cvb(0, cvb(1, gs0))
    return gW, gb, gx[..., 0]
