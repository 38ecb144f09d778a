"""The decode-only 9/7 path on the device against the oracle.

ndic_htj2k_decode_plane_batch_irreversible decodes lossy codestreams (9/7,
expounded QCD) bit-exactly as the oracle's irreversible decode does (whose
pins are in test_lossy_oracle.py): at full and reduced resolution, for odd
sizes, 0..5 levels, 8..16-bit signed and unsigned samples and a range of
step exponents / mantissas; with reversible codestreams in the same batch;
and with enough planes in one call that the two-pass lane decoder runs.  The
reversible entry point keeps refusing 9/7 codestreams as the reference does.
"""

from __future__ import annotations

import numpy as np
import pytest

import batch_tools

pytestmark = pytest.mark.gpu

NDIC_UNSUPPORTED = 3


def image(h, w, depth, signed, seed):
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:h, 0:w]
    top = (1 << depth) - 1
    v = top / 2 + top / 3 * np.sin(xx / 7.0 + seed) * np.cos(yy / 5.0) + rng.normal(0, top / 100, (h, w))
    v = np.clip(np.round(v), 0, top)
    return (v - (1 << (depth - 1)) if signed else v).astype(np.int32)


CASES = [
    # (h, w, levels, depth, signed, eps, mu)
    (64, 64, 5, 12, False, 13, 0),
    (97, 311, 4, 12, False, 11, 700),
    (130, 65, 3, 16, True, 15, 1234),
    (33, 7, 2, 8, False, 6, 2047),
    (1, 9, 2, 8, False, 8, 0),
    (200, 150, 0, 10, True, 9, 55),
    (257, 129, 5, 16, False, 12, 300),
]


@pytest.mark.parametrize("h,w,levels,depth,signed,eps,mu", CASES)
def test_decode_matches_oracle(ndic, oracle, h, w, levels, depth, signed, eps, mu):
    img = image(h, w, depth, signed, h + w)
    cs = oracle.encode_image_irreversible(img, depth, signed, levels, eps, mu, guard=3)
    for r in [-1] + list(range(levels)):
        want = oracle.decode_codestream(cs, max_res=r, irreversible=True)
        [(st, got)] = batch_tools.plane_batch(ndic, [cs], w, h, levels, max_res=r, irreversible=True)
        assert st == 0, (r, st)
        assert np.array_equal(got, want), (r, int(np.abs(got.astype(np.int64) - want).max()))


def test_mixed_batch_and_refusal(ndic, oracle):
    rev_img = image(40, 50, 12, False, 1)
    rev = oracle.encode_image(rev_img, 12, False, 3)
    lossy = oracle.encode_image_irreversible(image(40, 50, 12, False, 2), 12, False, 3, 12, 100, guard=3)
    res = batch_tools.plane_batch(ndic, [lossy, rev, lossy], 50, 40, 3, irreversible=True)
    want = oracle.decode_codestream(lossy, irreversible=True)
    assert [s for s, _ in res] == [0, 0, 0]
    assert np.array_equal(res[0][1], want) and np.array_equal(res[2][1], want)
    assert np.array_equal(res[1][1], rev_img)
    # the reversible entry point refuses the 9/7 codestream alone
    res = batch_tools.plane_batch(ndic, [lossy, rev], 50, 40, 3)
    assert res[0][0] == NDIC_UNSUPPORTED and res[1][0] == 0 and np.array_equal(res[1][1], rev_img)


def test_many_planes_take_the_lane_decoder(ndic, oracle):
    # 160 planes of 512 x 512 at 5 levels: 70 code-blocks each, 11,200 in one
    # window, above the wavefront decoder's default limit (8,192), so the
    # two-pass lane decoder runs
    img = image(512, 512, 16, False, 9)
    cs = oracle.encode_image_irreversible(img, 16, False, 5, 13, 0, guard=3)
    want = oracle.decode_codestream(cs, irreversible=True)
    res = batch_tools.plane_batch(ndic, [cs] * 160, 512, 512, 5, irreversible=True)
    for i, (st, got) in enumerate(res):
        assert st == 0 and np.array_equal(got, want), i
