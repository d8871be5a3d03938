"""CPU model of a single-launch chained fallback for the radix sort (no GPU
code): P passes of T tiles, one ticket counter for every pass, a tile of pass
p waiting until all T tiles of pass p - 1 are done and, inside a pass, until
the previous tile has published its prefix (the look-back).  At most R
workgroups are resident; the dispatcher starts workgroups in index order as
slots free up.  The model shows when such a launch drains and when it stalls
for good (no resident workgroup can advance while work is left):

  ticket / persistent / static: tiles handed out in pass-major start order --
      drains for every residency limit R >= 1;
  static_rm: tile index not in start order (passes interleaved over block
      indices) -- stalls: a resident tile waits for a tile whose workgroup is
      not resident yet;
  short_count: the wait target counts T_cap tiles while tiles past the data
      (x >= T) exit without counting themselves -- stalls.
"""

import itertools


def simulate(P, T, R, mode, T_cap=None):
    T_cap = T_cap or T
    tiles_total = P * T_cap
    done = [0] * P
    published = [[False] * T_cap for _ in range(P)]
    next_ticket = 0
    next_block = 0
    resident = []  # dicts: tile (p, x) or None (persistent, needs a ticket)
    finished = 0
    grid = R if mode == "persistent" else tiles_total

    def start_block(b):
        nonlocal next_ticket
        if mode in ("ticket", "persistent", "short_count"):
            t = next_ticket
            next_ticket += 1
        elif mode == "static":
            t = b
        else:  # static_rm
            t = (b % P) * T_cap + b // P
        return {"tile": divmod(t, T_cap) if t < tiles_total else None}

    for step in itertools.count():
        while len(resident) < R and next_block < grid:
            resident.append(start_block(next_block))
            next_block += 1
        if not resident:
            return "drains", step
        progressed = False
        for wg in list(resident):
            tile = wg["tile"]
            if tile is None:  # no work left for this workgroup
                resident.remove(wg)
                progressed = True
                continue
            p, x = tile
            beyond = x >= T  # past the data (short_count only)
            if beyond:
                resident.remove(wg)  # exits without counting itself
                progressed = True
                continue
            target = T_cap if mode == "short_count" else T
            if p > 0 and done[p - 1] < target:
                continue  # waits for the previous pass
            if x > 0 and not published[p][x - 1]:
                continue  # look-back waits for the previous tile
            published[p][x] = True
            done[p] += 1
            finished += 1
            progressed = True
            if mode == "persistent":
                t = next_ticket
                next_ticket += 1
                wg["tile"] = divmod(t, T_cap) if t < tiles_total else None
            else:
                resident.remove(wg)
        if not progressed:
            return "stalls", finished




def test_in_order_tickets_drain_under_any_residency():
    for mode in ("ticket", "persistent", "static"):
        for P, T, R in ((8, 64, 192), (8, 64, 20), (8, 64, 1), (3, 50, 7)):
            assert simulate(P, T, R, mode)[0] == "drains", (mode, P, T, R)


def test_out_of_order_tiles_stall():
    for P, T, R in ((8, 64, 192), (3, 50, 7), (8, 64, 1)):
        assert simulate(P, T, R, "static_rm")[0] == "stalls"


def test_wait_target_beyond_the_counted_tiles_stalls():
    assert simulate(8, 50, 192, "short_count", T_cap=64)[0] == "stalls"
    assert simulate(8, 64, 192, "short_count", T_cap=64)[0] == "drains"
