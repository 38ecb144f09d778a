"""The oracle's irreversible (9/7) decode path, pinned on the CPU.

The reference ships the pieces -- the unscaled 9/7 lifting
(crates/ndic-htj2k/src/dwt/mod.rs:103-325), the irreversible step
Quant::irrev_delta (crates/ndic-codestream/src/quant.rs:204-221) and the
block decoder's mid-point reconstruction (block/decode.rs:5-6) -- but its
reader refuses 9/7 codestreams (reader.rs:326-330) and its writer lossy
encodes (htj2k.rs:293-297).  The oracle assembles them into a decode as
T.800 Annex E/F define it, with the K factor the unscaled lifting leaves in
each band moved into the band's step.  What pins it:

* the lifting against T.800 Table F.4's analysis filters: an impulse at an
  even or odd sample gives K h0 taps in the low band and g0 / K taps in the
  high band, along rows and along columns;
* a constant plane transforms to c K^2L in the LL band and zeros elsewhere,
  a cubic ramp to zero interior detail (four vanishing moments), and the
  inverse undoes the forward transform;
* with no decomposition levels the decoded samples are T.800 E.1.1.2's
  mid-point reconstruction sign(q) (|q| + 1/2) Delta 2^depth, level shifted
  and rounded, with Delta from the QCD exponent / mantissa and the LL gain;
* a constant image comes back exactly through five levels (a wrong K factor
  or band gain scales it), and every reduced-resolution decode equals the LL
  band of the forward transform of the full decode over K^2(L - r)
  (the lossy form of codestream_roundtrip.rs:103-128);
* without the switch, and for the wavelet / QCD style mixes it does not
  cover, the codestream is refused as the reference refuses it.
"""

from __future__ import annotations

import numpy as np
import pytest

K = 1.230174104914001  # T.800 Table F.4
H0 = {0: 0.6029490182363579, 1: 0.2668641184428723, 2: -0.07822326652898785, 3: -0.01686411844287495,
      4: 0.02674875741080976}
G0 = {0: 1.115087052456994, 1: -0.5912717631142470, 2: -0.05754352622849957, 3: 0.09127176311424948}


def expected_bands(n, pos):
    nl, nh = (n + 1) // 2, n // 2
    low = np.array([K * H0.get(abs(2 * i - pos), 0.0) for i in range(nl)])
    high = np.array([G0.get(abs(2 * i + 1 - pos), 0.0) / K for i in range(nh)])
    return low, high


@pytest.mark.parametrize("pos", [20, 21, 30, 31])
def test_impulse_responses_are_the_t800_filters(oracle, pos):
    n = 64
    x = np.zeros((1, n), np.float32)
    x[0, pos] = 1
    y = oracle.dwt97(True, x, 1)[0]
    low, high = expected_bands(n, pos)
    np.testing.assert_allclose(y[:32], low, atol=2e-7)
    np.testing.assert_allclose(y[32:], high, atol=2e-7)
    # the same along columns
    yc = oracle.dwt97(True, x.T.copy(), 1)[:, 0]
    np.testing.assert_allclose(yc[:32], low, atol=2e-7)
    np.testing.assert_allclose(yc[32:], high, atol=2e-7)


@pytest.mark.parametrize("w,h,levels", [(40, 24, 3), (33, 17, 5), (64, 64, 2)])
def test_constant_plane_keeps_only_scaled_dc(oracle, w, h, levels):
    c = 7.0
    y = oracle.dwt97(True, np.full((h, w), c, np.float32), levels)
    lw, lh = w, h
    for _ in range(levels):
        lw, lh = (lw + 1) // 2, (lh + 1) // 2
    ll = y[:lh, :lw]
    np.testing.assert_allclose(ll, c * K ** (2 * levels), rtol=2e-6)
    rest = y.copy()
    rest[:lh, :lw] = 0
    assert np.abs(rest).max() < 1e-4


def test_cubic_ramp_has_no_interior_detail(oracle):
    n = 96
    t = np.arange(n, dtype=np.float64) / n
    x = (1 + 2 * t - 3 * t ** 2 + 1.5 * t ** 3).astype(np.float32)[None, :]
    y = oracle.dwt97(True, x, 1)[0]
    high = y[n // 2:]
    assert np.abs(high[4:-4]).max() < 2e-6 and np.abs(high[-2:]).max() > 1e-4


@pytest.mark.parametrize("w,h,levels", [(37, 29, 4), (1, 9, 3), (130, 7, 5), (2, 2, 1)])
def test_inverse_undoes_forward(oracle, w, h, levels):
    x = np.random.default_rng(w * h).normal(0, 100, (h, w)).astype(np.float32)
    back = oracle.dwt97(False, oracle.dwt97(True, x, levels), levels)
    np.testing.assert_allclose(back, x, atol=2e-4)


def quant_step(eps, mu, gain, depth):
    """T.800 E-3 with R_b = depth + log2(gain)."""
    return gain * 2.0 ** (depth - eps) * (1 + mu / 2048)


@pytest.mark.parametrize("depth,signed,eps,mu", [(12, False, 9, 0), (8, False, 5, 1500), (16, True, 12, 77)])
def test_no_levels_is_the_midpoint_reconstruction(oracle, depth, signed, eps, mu):
    rng = np.random.default_rng(depth)
    lo, hi = (-(1 << (depth - 1)), (1 << (depth - 1)) - 1) if signed else (0, (1 << depth) - 1)
    img = rng.integers(lo, hi + 1, size=(23, 41)).astype(np.int32)
    cs = oracle.encode_image_irreversible(img, depth, signed, 0, eps, mu, guard=3)
    got = oracle.decode_codestream(cs, irreversible=True)
    half = 0 if signed else 1 << (depth - 1)
    step = quant_step(eps, mu, 1, depth)
    x = (img - half).astype(np.float64)
    q = np.floor(np.abs(x) / step)
    rec = np.where(q > 0, np.sign(x) * (q + 0.5) * step, 0.0)
    want = np.clip(np.rint(rec), -(1 << (depth - 1)), (1 << (depth - 1)) - 1) + half
    assert np.array_equal(got, want.astype(np.int32))


def test_constant_image_survives_five_levels(oracle):
    img = np.full((70, 90), 1234, np.int32)
    cs = oracle.encode_image_irreversible(img, 12, False, 5, 16, 0, guard=3)
    assert np.array_equal(oracle.decode_codestream(cs, irreversible=True), img)


def smooth(h, w, depth, seed):
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:h, 0:w]
    top = (1 << depth) - 1
    v = top / 2 + top / 3 * np.sin(xx / 9) * np.cos(yy / 7) + rng.normal(0, top / 200, (h, w))
    return np.clip(np.round(v), 0, top).astype(np.int32)


@pytest.mark.parametrize("h,w,levels,eps", [(77, 130, 3, 14), (64, 64, 5, 11), (31, 97, 4, 13)])
def test_reduced_resolutions_are_the_ll_of_the_full_decode(oracle, h, w, levels, eps):
    img = smooth(h, w, 12, h + w)
    cs = oracle.encode_image_irreversible(img, 12, False, levels, eps, 0, guard=3)
    full = oracle.decode_codestream(cs, irreversible=True)
    step = quant_step(eps, 0, 1, 12)
    assert np.sqrt(np.mean((full - img) ** 2.0)) < 4 * step
    for r in range(levels):
        got = oracle.decode_codestream(cs, max_res=r, irreversible=True)
        lh, lw = got.shape
        f = oracle.dwt97(True, (full - 2048).astype(np.float32), levels - r)
        ll = f[:lh, :lw].astype(np.float64) / K ** (2 * (levels - r)) + 2048
        assert np.abs(got - ll).max() <= 1.5, r


def test_refusals(oracle):
    img = smooth(16, 16, 8, 0)
    cs = oracle.encode_image_irreversible(img, 8, False, 2, 8, 0, guard=3)
    with pytest.raises(oracle.OracleError, match="irreversible"):
        oracle.decode_codestream(cs)
    # 5/3 with an expounded QCD, and 9/7 with a reversible QCD: neither path
    at_cod = cs.index(b"\xff\x52")
    wavelet_at = at_cod + 4 + 9
    assert cs[wavelet_at] == 0
    mixed = bytearray(cs)
    mixed[wavelet_at] = 1
    with pytest.raises(oracle.OracleError, match="irreversible"):
        oracle.decode_codestream(bytes(mixed), irreversible=True)
    rev = oracle.encode_image(img, 8, False, 2)
    at_cod = rev.index(b"\xff\x52")
    mixed = bytearray(rev)
    mixed[at_cod + 4 + 9] = 0
    with pytest.raises(oracle.OracleError, match="irreversible"):
        oracle.decode_codestream(bytes(mixed), irreversible=True)
    # reversible streams decode the same with the switch on
    assert np.array_equal(oracle.decode_codestream(rev, irreversible=True), img)
