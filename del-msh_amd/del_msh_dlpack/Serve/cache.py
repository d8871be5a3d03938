"""Host-side bookkeeping of the serving loop (no GPU needed): the byte-budget
LRU cache of device-resident trees with pinning, and request admission."""

from __future__ import annotations

from collections import OrderedDict
from dataclasses import dataclass, field


@dataclass
class Entry:
    key: int
    nbytes: int
    value: object = None
    pins: int = 0


class ByteLRU:
    """Entries keyed by mesh id with a byte budget.  Pinned entries (used by a
    step in flight) are never evicted; unpinned ones leave least recently
    used first.  reserve() makes room or says that it cannot."""

    def __init__(self, budget: int):
        self.budget = int(budget)
        self.used = 0
        self.entries: OrderedDict[int, Entry] = OrderedDict()
        self.evicted = 0
        self.hits = 0
        self.misses = 0

    def __contains__(self, key: int) -> bool:
        return key in self.entries

    def touch(self, key: int) -> Entry:
        e = self.entries[key]
        self.entries.move_to_end(key)
        return e

    def evictable_bytes(self) -> int:
        return sum(e.nbytes for e in self.entries.values() if e.pins == 0)

    def can_fit(self, nbytes: int) -> bool:
        return nbytes <= self.budget - self.used + self.evictable_bytes()

    def reserve(self, nbytes: int) -> bool:
        """Evict unpinned entries (oldest first) until nbytes fit; False (and
        nothing evicted) if they cannot."""
        if not self.can_fit(nbytes):
            return False
        for key in list(self.entries.keys()):
            if self.used + nbytes <= self.budget:
                break
            e = self.entries[key]
            if e.pins == 0:
                del self.entries[key]
                self.used -= e.nbytes
                self.evicted += 1
        return self.used + nbytes <= self.budget

    def insert(self, key: int, nbytes: int, value, pins: int = 0) -> Entry:
        if key in self.entries:
            raise KeyError(f"{key} already cached")
        e = Entry(key, int(nbytes), value, pins)
        self.entries[key] = e
        self.used += e.nbytes
        return e

    def pin(self, key: int) -> None:
        self.entries[key].pins += 1

    def unpin(self, key: int) -> None:
        e = self.entries.get(key)
        if e is not None and e.pins > 0:
            e.pins -= 1


@dataclass
class Request:
    rid: int
    arrival: float          # seconds from the start of the stream
    mesh_id: int
    width: int
    height: int
    ndc2world: object       # 4x4 row-major (numpy)
    done: float = -1.0      # completion time, same clock as arrival
    result: object = field(default=None, repr=False)


def admit(queue: list, cache: ByteLRU, tree_bytes, max_requests: int, max_pixels: int):
    """Admission for one step, first come first served: a request joins if its
    tree is cached or room can be made for it (evicting unpinned trees), and
    the step stays within max_requests / max_pixels.  Admitted requests' trees
    are pinned (new trees are reserved; the caller inserts them).  Returns
    (admitted, new_mesh_ids); the rest stay queued, in order."""
    admitted, new_ids, keep = [], [], []
    pixels = 0
    planned = {}
    for r in queue:
        px = r.width * r.height
        fits_step = len(admitted) < max_requests and pixels + px <= max_pixels
        ok = False
        if fits_step:
            if r.mesh_id in cache:
                cache.touch(r.mesh_id)
                cache.pin(r.mesh_id)
                cache.hits += 1
                ok = True
            elif r.mesh_id in planned:
                planned[r.mesh_id] += 1
                cache.hits += 1
                ok = True
            else:
                nb = tree_bytes(r.mesh_id)
                if cache.reserve(nb):
                    # account the bytes now so later admissions see them
                    cache.used += nb
                    planned[r.mesh_id] = 1
                    new_ids.append(r.mesh_id)
                    cache.misses += 1
                    ok = True
        if ok:
            admitted.append(r)
            pixels += px
        else:
            keep.append(r)
    queue[:] = keep
    for mid in new_ids:
        cache.used -= tree_bytes(mid)  # insert() adds them back with the entry
    return admitted, new_ids, planned
