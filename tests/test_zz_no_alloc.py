"""pond_mlp_forward_noalloc (the entry point that allocates no device memory) against
pond_mlp_forward: the same outputs bit for bit on consecutive calls, the second smaller than the
first (the static scratch keeps the first call's contents), and a clear error beyond its
maximum of 16384 rows."""
import numpy as np
import pytest

from lbmpond_amd import Pond, PondError

pytestmark = pytest.mark.gpu


def test_forward_noalloc_matches_forward(gpu):
    rng = np.random.default_rng(21)
    with Pond("D3Q27", (16, 8, 4), "weno5c", "split", use_mlp=True) as p:
        p.mlp_init(seed=9)
        for B in (16384, 4000, 37):
            print(f"[noalloc] rows {B}", flush=True)
            X = rng.standard_normal((B, 135)).astype(np.float32)
            ref = p.mlp_forward(X)
            got = p.mlp_forward_noalloc(X)
            assert got.shape == (B, 4)
            assert np.isfinite(ref).all()
            assert np.array_equal(got, ref)
        with pytest.raises(PondError, match="exceeds the maximum"):
            p.mlp_forward_noalloc(np.zeros((16385, 135), np.float32))


def test_forward_noalloc_two_contexts_share_the_scratch(gpu):
    """two networks, one after the other on the same GPU: each call sees only its own rows"""
    rng = np.random.default_rng(22)
    X1 = rng.standard_normal((300, 135)).astype(np.float32)
    X2 = rng.standard_normal((20, 135)).astype(np.float32)
    with Pond("D3Q27", (16, 8, 4), "weno5c", "split", use_mlp=True) as a, \
            Pond("D3Q27", (16, 8, 4), "weno5c", "split", use_mlp=True) as b:
        a.mlp_init(seed=1)
        b.mlp_init(seed=2)
        ya = a.mlp_forward_noalloc(X1)
        yb = b.mlp_forward_noalloc(X2)
        assert np.array_equal(ya, a.mlp_forward(X1))
        assert np.array_equal(yb, b.mlp_forward(X2))
