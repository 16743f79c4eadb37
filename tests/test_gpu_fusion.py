"""Two-level S/T stages against the one-level schedule.

A two-level stage evaluates the same expressions in the same order as two
one-level stages (only the middle level is never stored), so the product must
be bit-identical with the setting on or off, for every algorithm, dtype, leaf
mode, depth and peeling pattern below.  Parity of the default (fused) path
with the oracle is covered by the golden cases in test_gpu_parity.py; here the
one-level path is the reference and the tolerance is zero.
"""
import numpy as np
import pytest
import torch

from cases import make_inputs

pytestmark = pytest.mark.gpu

DEV = "cuda:0"


def _dev(a, fm):
    return fm.col_major(torch.from_numpy(np.asfortranarray(a)).to(DEV))


def _both(fm, alg, A, B, make_c, steps, alpha=1.0, beta=0.0, mode="plain"):
    """(C fused, C one-level, S+T launches fused, S+T launches one-level);
    make_c() gives a fresh C operand (or None) for each run."""
    out = []
    for fuse in (True, False):
        fm.set_stage_fusion(fuse)
        try:
            C = make_c()
            fm.profile_enable(True)
            fm.profile_read()
            C = fm.fastmatmul(alg, A, B, C, num_steps=steps, alpha=alpha, beta=beta, mode=mode)
            torch.cuda.synchronize()
            prof = fm.profile_read()
        finally:
            fm.profile_enable(False)
            fm.set_stage_fusion(True)
        out.append((C, prof["S"][1] + prof["T"][1]))
    return out[0][0], out[1][0], out[0][1], out[1][1]


CASES = [
    # alg, (m, k, n), steps, fused S+T launches expected (None: any)
    ("strassen", (256, 256, 256), 2, 2),
    ("strassen", (512, 384, 256), 3, 4),         # levels (1, 2) paired, level 0 alone
    ("strassen", (512, 512, 512), 4, 4),         # (2, 3) and (0, 1)
    ("strassen", (129, 133, 137), 2, 2),         # top level peels, middle level does not
    ("strassen", (130, 130, 130), 2, 4),         # middle level peels: one-level stages
    ("grey424_26_206", (512, 256, 512), 2, 2),
    ("grey424_26_206", (1024, 256, 512), 3, 4),
    ("smirnov333_23_139", (243, 243, 243), 2, 2),
    ("grey423_20_144", (256, 128, 288), 2, 2),
    ("grey323_15_103", (288, 128, 288), 2, 2),
    ("bini322_10_52_approx", (144, 64, 64), 2, 2),
    ("grey234_20_144", (128, 288, 256), 2, 3),   # B side has 12 blocks: one-level T
    ("grey234_20_144", (128, 192, 256), 2, 4),   # 64 % 3 != 0: middle level peels
]


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("alg,shape,steps,launches", CASES)
def test_two_level_bit_identical(fm, alg, shape, steps, launches, dtype):
    m, k, n = shape
    A, B, C0 = make_inputs(m, k, n, 11, "u11", dtype, with_c=True)
    Ad, Bd = _dev(A, fm), _dev(B, fm)
    Cf, Cu, lf, lu = _both(fm, alg, Ad, Bd, lambda: _dev(C0, fm), steps, alpha=0.75, beta=-1.25)
    assert torch.equal(Cf, Cu), float((Cf - Cu).abs().max())
    if launches is not None:
        assert lf == launches
    assert lf <= lu


@pytest.mark.parametrize("alg,shape,steps", [("strassen", (256, 256, 256), 2),
                                             ("grey424_26_206", (512, 256, 512), 2)])
def test_two_level_on_strided_views(fm, alg, shape, steps):
    """Operands that are sub-matrices of larger buffers (lda > m etc.)."""
    m, k, n = shape
    A, B, C0 = make_inputs(m + 7, k + 5, n + 3, 12, "u11", np.float64, with_c=True)
    Ad = _dev(A, fm)[3:3 + m, 1:1 + k]
    Bd = _dev(B, fm)[2:2 + k, :n]
    big = np.asfortranarray(np.ones((m + 7, n + 3)))
    Cf, Cu, lf, lu = _both(fm, alg, Ad, Bd, lambda: _dev(big, fm)[1:1 + m, 2:2 + n], steps,
                           beta=0.5)
    assert Cf.stride() == (1, m + 7)
    assert torch.equal(Cf, Cu)
    assert lf < lu


@pytest.mark.parametrize("mode,dtype", [("split_bf16", np.float32), ("ozaki_int8", np.float64)])
def test_two_level_leaf_modes(fm, mode, dtype):
    A, B, _ = make_inputs(256, 256, 256, 13, "u11", dtype)
    Cf, Cu, lf, lu = _both(fm, "strassen", _dev(A, fm), _dev(B, fm), lambda: None, 2, mode=mode)
    assert torch.equal(Cf, Cu)
    assert (lf, lu) == (2, 4)


def test_two_level_exact_on_integers(fm):
    """Integer data: both schedules equal the exact product."""
    A, B, _ = make_inputs(512, 256, 512, 14, "int", np.float64)
    Cf, Cu, _, _ = _both(fm, "grey424_26_206", _dev(A, fm), _dev(B, fm), lambda: None, 2)
    exact = A @ B
    assert np.array_equal(np.asfortranarray(Cf.cpu().numpy()), exact)
    assert torch.equal(Cf, Cu)
