"""The activation kernel's candidate walk against the reference's in-place
walk (CheckIfNearBoundaries, src/TimeStepper.f90:935-1150; AddTile,
src/UpdateTiles.f90:56-78).

The reference walks the sorted active list in four passes (N, S, E, W). A
flagged side whose neighbour is activatable inserts that neighbour in order,
and the later entries of the pass shift. The device walks only the candidate
sides: those whose action at the start of the step is to activate or to halt.
A candidate at initial position i0 is reached at i0 plus the number of new
tiles below it, and is visited if that position is below the pass's initial
length. This test restates both walks over small random tile grids and checks
that they agree: the final list, the tile states, the ghost list (kept in
order) and the order of the new tiles."""
import random

OFF, GHOST, ACTIVE = 0, 1, 2


def side(g, k, r):
    """Neighbour of tile k on side r (N, S, E, W): raw index arithmetic, or
    wrapped with periodic boundaries (nb_tile)."""
    ntx, nty, periodic = g["ntx"], g["nty"], g["periodic"]
    di = 1 if r == 2 else (-1 if r == 3 else 0)
    dj = 1 if r == 0 else (-1 if r == 1 else 0)
    if not periodic:
        return k + di + dj * ntx
    return (k % ntx + di) % ntx + ((k // ntx + dj) % nty) * ntx


def action(g, ts, n):
    """0 nothing, 1 activate, 2 halt (AddTile)."""
    ntx, nty, N = g["ntx"], g["nty"], g["ntx"] * g["nty"]
    if n <= -1 or (n < N and ts[n] == ACTIVE):
        return 0
    ti, tj = n % ntx, n // ntx
    edge = ti in (0, ntx - 1) or (nty > 1 and tj in (0, nty - 1))
    if n >= N or (not g["periodic"] and edge):
        return 2 if g["halt"] else 0
    return 1


def walk_in_place(g, active, ts, flags, ghosts):
    active, ts, ghosts, new = list(active), list(ts), list(ghosts), []
    for q in range(4):
        size, tt = len(active), 0
        while tt < size:
            k = active[tt]
            if flags.get(k, 0) & (1 << q):
                n = side(g, k, q)
                a = action(g, ts, n)
                if a == 2:
                    return None
                if a == 1:
                    pos = len(active)
                    active.append(None)
                    while pos > 0 and active[pos - 1] > n:
                        active[pos] = active[pos - 1]
                        pos -= 1
                    active[pos] = n
                    if ts[n] == GHOST:
                        ghosts.remove(n)
                    ts[n] = ACTIVE
                    new.append(n)
            tt += 1
    return active, ts, ghosts, new


def walk_candidates(g, active, ts, flags, ghosts):
    ts, n0, cand = list(ts), len(active), []
    for i0, k in enumerate(active):
        bits = 0
        for r in range(4):
            if flags.get(k, 0) & (1 << r) and action(g, ts, side(g, k, r)) != 0:
                bits |= 1 << r
        if bits:
            cand.append((k, i0, bits))
    new, size = [], n0
    for q in range(4):
        for k, i0, bits in cand:
            if not bits & (1 << q):
                continue
            if i0 + sum(1 for m in new if m < k) >= size:
                break
            n = side(g, k, q)
            a = action(g, ts, n)
            if a == 2:
                return None
            if a == 1:
                ts[n] = ACTIVE
                new.append(n)
        size = n0 + len(new)
    merged = sorted(active + new)
    return merged, ts, [x for x in ghosts if ts[x] != ACTIVE], new


def test_candidate_walk_equals_in_place_walk():
    rng = random.Random(20261203)
    grown = halted = 0
    for _ in range(20000):
        g = {"ntx": rng.randint(2, 7), "nty": rng.randint(1, 6),
             "periodic": rng.random() < 0.3, "halt": rng.random() < 0.2}
        N = g["ntx"] * g["nty"]
        ts = [OFF] * N
        active = sorted(rng.sample(range(N), rng.randint(1, N)))
        for k in active:
            ts[k] = ACTIVE
        ghosts = [k for k in rng.sample(range(N), N) if ts[k] == OFF and rng.random() < 0.7]
        for k in ghosts:
            ts[k] = GHOST
        flags = {k: (rng.randint(0, 15) if rng.random() < 0.6 else 0) for k in active}
        a = walk_in_place(g, active, ts, flags, ghosts)
        b = walk_candidates(g, active, ts, flags, ghosts)
        assert a == b
        halted += a is None
        grown += a is not None and len(a[3]) > 0
    assert grown > 5000 and halted > 500


def test_planted_mistake_fails():
    """Counting only the candidates' own position (no shift by new tiles
    below) differs from the in-place walk on some list."""
    rng = random.Random(7)
    differs = False
    for _ in range(5000):
        g = {"ntx": rng.randint(2, 6), "nty": rng.randint(2, 5), "periodic": False, "halt": False}
        N = g["ntx"] * g["nty"]
        ts = [OFF] * N
        active = sorted(rng.sample(range(N), rng.randint(1, N)))
        for k in active:
            ts[k] = ACTIVE
        flags = {k: rng.randint(0, 15) for k in active}
        a = walk_in_place(g, active, ts, flags, [])
        ts2, new = list(ts), []
        for q in range(4):
            for k in active:
                if flags[k] & (1 << q) and action(g, ts2, side(g, k, q)) == 1:
                    n = side(g, k, q)
                    ts2[n] = ACTIVE
                    new.append(n)
        if a is not None and new != a[3]:
            differs = True
            break
    assert differs
