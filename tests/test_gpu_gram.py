"""GPU parity of the signature-kernel Gram matrix (MFMA, bf16 hi/lo split)
against a float64 GEMM of the same float32 signatures
(examples/signature-kernel-by-signature-dot.py:20-23 defines the kernel as
np.inner of the dense signatures).  The max relative error, normalised by
sqrt(G_ii G_jj), is stated and bounded by 1e-4."""
import numpy as np
import pytest

from roughpy_amd import streams as rs

pytestmark = pytest.mark.gpu


def _rel_err(G, ref):
    d = np.sqrt(np.outer(np.diag(ref), np.diag(ref)))
    return float(np.abs(G - ref).max() / d.max()), float((np.abs(G - ref) / d).max())


@pytest.mark.parametrize("n,k", [(1, 7), (5, 13), (64, 121), (100, 3906), (257, 1000)])
def test_gram_random(gpu, rng, n, k):
    S = rng.normal(size=(n, k)).astype(np.float32)
    G = rs.signature_gram(S)
    ref = S.astype(np.float64) @ S.astype(np.float64).T
    err, rel = _rel_err(G, ref)
    print(f"gram n={n} k={k}: max err / max scale {err:.2e}, max entrywise-normalised {rel:.2e}")
    assert rel < 1e-4
    np.testing.assert_array_equal(G, G.T)


def test_gram_of_signatures_w6_d5(gpu):
    """Signatures of 512 streams, width 6, depth 5 (9331 coefficients),
    length 512, then their Gram matrix."""
    rng = np.random.default_rng(12345 + 5)
    n, L, w, d = 512, 512, 6, 5
    inc = (rng.normal(size=(n * L, w)) / np.sqrt(L)).astype(np.float32)
    offs = np.arange(0, n * L + 1, L, dtype=np.int64)
    S = rs.signature_batch(gpu.from_numpy(inc).cuda(), gpu.from_numpy(offs).cuda(), w, d)
    G = rs.signature_gram(S).cpu().numpy()
    S64 = S.double().cpu().numpy()
    ref = S64 @ S64.T
    err, rel = _rel_err(G, ref)
    print(f"gram of signatures: max err / max scale {err:.2e}, entrywise-normalised {rel:.2e}")
    assert rel < 1e-4
