"""The stratified G-vector sample of the full-size GPU parity tests (CPU)."""
import numpy as np

from tests.helpers import stratified_sample


def test_stratified_sample_meets_every_tile():
    """The sample puts an element pair (i >= j) into every lower tile of both
    tilings at the largest size (CPU check of the sampling itself)."""
    G = 24995
    idx = stratified_sample(G)
    assert len(idx) == 196 and np.all(np.diff(idx) > 0) and idx[-1] < G
    for bm, bn in ((128, 128), (256, 192)):
        hit = set()
        for i in idx:
            for j in idx[idx <= i]:
                hit.add((i // bm, j // bn))
        tiles = {(bi, bj) for bi in range((G + bm - 1) // bm) for bj in range((G + bn - 1) // bn)
                 if bj * bn < (bi + 1) * bm}
        # a tile the diagonal cuts may hold only pairs with i < j: at most one such tile per tile row
        missed = tiles - hit
        assert len(missed) <= (G + bm - 1) // bm, (bm, bn, len(missed))
        assert all(bj * bn + bn > bi * bm for bi, bj in missed)
