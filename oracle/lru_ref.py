"""Test infrastructure (oracle): a restatement of the reference's domain LRU
cache, render/lru_cache.cc:41-171, used by the tests to check the paged
domain pool's residency and hit/miss counts.  Not part of the product.

LruCache::init (lru_cache.cc:41-63): capacity = all domains when the cache
size is negative or at least the domain count, else the cache size (0 is an
error).  LruCache::load (lru_cache.cc:65-171): a loaded domain is a hit and
moves to the most recently used end; a missing one takes the next block
while the cache is not full, else evicts the least recently used domain.
"""
from collections import OrderedDict


class LruRef:
    def __init__(self, num_domains, cache_size):
        if cache_size == 0:
            raise ValueError("cache size must not be 0")  # lru_cache.cc:44
        self.capacity = num_domains if (cache_size < 0 or cache_size >= num_domains) else cache_size
        self.blocks = OrderedDict()  # domain -> block, least recently used first
        self.hits = 0
        self.misses = 0

    def load(self, dom):
        """Returns (hit, block) as LruCache::load does."""
        if dom in self.blocks:
            self.hits += 1
            self.blocks.move_to_end(dom)
            return True, self.blocks[dom]
        self.misses += 1
        if len(self.blocks) < self.capacity:
            block = len(self.blocks)
        else:
            _, block = self.blocks.popitem(last=False)
        self.blocks[dom] = block
        return False, block

    def lru_order(self):
        return list(self.blocks.keys())
