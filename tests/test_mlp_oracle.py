"""The numpy restatement of the learned-physics network: gradients pinned by central finite
differences in float64, BF16 rounding against known bit patterns, and the log-moment decoding
of the width classes against the soft targets it is trained on."""
import numpy as np

from oracle import mlp_ref as R


def _params(rng, scale=0.3):
    p = rng.normal(0.0, scale, R.NPARAM)
    return p


def test_bf16_rounding_known_values():
    x = np.array([1.0, 1.00390625, 1.0029296875, 1.0048828125, -2.5, 3.0e38, 1e-40, np.inf], np.float32)
    y = R.bf16(x)
    # BF16 ULP at 1 is 2^-7: 1 + 2^-8 is a tie (nearest even: 1), 1 + 0.375 ULP rounds down,
    # 1 + 0.625 ULP rounds up
    assert y[0] == 1.0 and y[1] == 1.0 and y[2] == 1.0 and y[3] == np.float32(1.0078125)
    assert y[4] == -2.5 and np.isinf(y[7])
    assert np.isnan(R.bf16(np.array([np.nan], np.float32)))[0]


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    p = _params(rng)
    B = 7
    x = rng.uniform(-1, 1, (B, R.NIN))
    dict4 = np.array([-6.0, 1.0, -9.0, -1.0])
    labels = np.stack([rng.uniform(-6, 1, B), rng.uniform(-9, -1, B), rng.uniform(-20, -3, B)], axis=1)
    loss, g = R.loss_and_grad(p, x, labels, dict4, w_att=0.5, w_mom=0.7, emulate=False)
    idx = np.concatenate([rng.choice(np.arange(off, off + int(np.prod(sh))), 6, replace=False)
                          for _, off, sh in R.SHAPES])
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        lp, _ = R.loss_and_grad(pp, x, labels, dict4, w_att=0.5, w_mom=0.7, emulate=False)
        lm, _ = R.loss_and_grad(pm, x, labels, dict4, w_att=0.5, w_mom=0.7, emulate=False)
        num = (lp - lm) / (2 * h)
        assert abs(num - g[i]) <= 1e-6 * max(1.0, abs(num)), (i, num, g[i])


def test_decoding_recovers_the_soft_targets():
    """The width classes are a lossless code of the widths: the log-moments of the soft
    targets equal the labels (inside the dictionary range)."""
    rng = np.random.default_rng(1)
    dict4 = np.array([-7.0, 0.5, -10.0, -0.5])
    lab = np.stack([rng.uniform(-7, 0.5, 500), rng.uniform(-10, -0.5, 500), np.full(500, -5.0)], axis=1)
    q = R.soft_targets(lab, dict4)
    z = np.concatenate([np.log(np.maximum(q, 1e-300)), lab[:, 2:3]], axis=1)
    dec = R.decode(z, dict4)
    assert np.allclose(dec[:, 0], lab[:, 0], atol=1e-9)
    assert np.allclose(dec[:, 1], lab[:, 1], atol=1e-9)
