"""The domain-cache oracle (oracle/lru_ref.py, a restatement of
render/lru_cache.cc) against known answers."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

from lru_ref import LruRef  # noqa: E402

# The textbook LRU reference string (Silberschatz et al., Operating System
# Concepts): 10 faults with 3 frames, 8 with 4.
BELADY = [1, 2, 3, 4, 1, 2, 5, 1, 2, 3, 4, 5]


@pytest.mark.parametrize("frames,faults", [(3, 10), (4, 8), (5, 5)])
def test_lru_known_fault_counts(frames, faults):
    lru = LruRef(8, frames)
    for d in BELADY:
        lru.load(d)
    assert lru.misses == faults
    assert lru.hits == len(BELADY) - faults


def test_lru_blocks_and_order():
    lru = LruRef(10, 3)
    assert [lru.load(d) for d in (4, 5, 6)] == [(False, 0), (False, 1), (False, 2)]  # fill in order
    assert lru.load(5) == (True, 1)  # hit promotes
    assert lru.lru_order() == [4, 6, 5]
    assert lru.load(7) == (False, 0)  # evicts 4, reuses its block
    assert lru.lru_order() == [6, 5, 7]


def test_lru_capacity_rules():
    assert LruRef(6, -1).capacity == 6  # lru_cache.cc:46-47: negative = all domains
    assert LruRef(6, 9).capacity == 6
    assert LruRef(6, 2).capacity == 2
    with pytest.raises(ValueError):
        LruRef(6, 0)  # lru_cache.cc:44
