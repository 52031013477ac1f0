"""CPU: the segment layout of a batched scan."""
import numpy as np

from openecloud_amd.scan import PB, plan_segments, seed_key


def test_segments_are_push_block_multiples_with_room_to_grow():
    counts = [1, 10000, 255, 2000000, 0]
    seg, ab = plan_segments(counts, growth=1.5, slack=1024, absFrac=1. / 16, absMin=1024)
    sizes = np.diff(seg)
    assert seg[0] == 0 and np.all(sizes % PB == 0)
    assert np.all(sizes >= np.asarray(counts) * 1.5 + 1024)
    assert np.all(sizes < np.asarray(counts) * 1.5 + 1024 + PB)
    hits = np.diff(ab)
    assert ab[0] == 0 and np.all(hits >= 1024) and np.all(hits >= sizes / 16)


def test_seed_key_matches_randomGen():
    from openecloud_amd import randomGen
    for s in (0, 1, 2 ** 32 + 5, 2 ** 64 - 1, 123456789012345):
        randomGen.seed(s)
        assert randomGen.key() == seed_key(s)
