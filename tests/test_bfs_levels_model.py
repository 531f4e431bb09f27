"""CPU model of the GPU's level-synchronous RSDF search, against the FIFO oracle.

The GPU search (mmnorm_amd/csrc/rsdf_bfs.hip) replaces the reference's FIFO
queue (sdf.py:195-222) with per-level claims: frontier position p of a level
whose first entry has queue index q0 offers the key 6 (q0 + p) + d to its
enterable neighbour in direction d; the neighbour keeps the smallest key; the
entries that hold their own keys emit those children in (p, d) order.  Seeds
hold -1, and voxels reached earlier hold smaller keys, so no visited test is
needed.  The multi-block path offers the next level's keys while it emits, and
the host moves between the single-workgroup path (frontiers up to a cap) and
the multi-block path (batches of levels, the first claimed by a kernel of its
own).  This model runs those steps with NumPy, in the same order and with the
same hand-over rule, and must give the oracle's rsdf and hops bit for bit for
every cap and batch size.  Each planted mistake (on this CPU model only) is a
plausible slip in that design and must make some field differ.

The multi-block levels compute their output offsets with a single-pass scan
across blocks (k_level); its protocol is simulated at the end of this file
under random schedules, dispatch orders and residency limits.
"""
import numpy as np
import pytest
from scipy import ndimage

SIX = ndimage.generate_binary_structure(3, 1)
INT32_MAX = np.iinfo(np.int32).max
# the reference's neighbour order (sdf.py:145-147)
DIRS = np.array([(-1, 0, 0), (0, 1, 0), (1, 0, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)])


def level_search(normals, labels, seeds, dx, cap, batch, plant=None):
    nx, ny, nz = labels.shape
    V = nx * ny * nz
    nrm = normals.reshape(-1, 3)
    lab = labels.ravel()
    step = DIRS @ np.array([ny * nz, nz, 1])
    xyz = np.stack(np.unravel_index(np.arange(V), labels.shape), -1)
    emask = np.zeros(V, np.int64)
    for d in range(6):
        c = xyz + DIRS[d]
        inside = np.all((c >= 0) & (c < labels.shape), axis=1)
        nb = np.where(inside, np.arange(V) + step[d], 0)
        ok = inside & ~np.isnan(nrm[nb, 0]) & (lab[nb] == lab)
        emask |= ok.astype(np.int64) << d

    claim = np.full(V, INT32_MAX, np.int64)
    rsdf = np.full(V, np.nan)
    hops = np.full(V, -1, np.int32)
    front = np.asarray(seeds, np.int64)
    if plant != "seed_unclaimed":
        claim[front] = -1
    rsdf[front] = 0.0
    hops[front] = 0

    def offer(fr, q):
        if plant == "no_queue_offset":
            q = 0
        for d in range(6):
            sel = np.nonzero(emask[fr] >> d & 1)[0]
            np.minimum.at(claim, fr[sel] + step[d], 6 * (q + sel) + d)

    def run_level(fr, q0, level, claim_first, claim_next):
        if claim_first:
            offer(fr, q0)
        keys = 6 * (q0 + np.arange(len(fr)))[:, None] + np.arange(6)
        if plant == "no_queue_offset":
            keys = keys - 6 * q0
        nb = fr[:, None] + step[None, :]
        enter = (emask[fr][:, None] >> np.arange(6) & 1).astype(bool)
        won = enter & (claim[np.where(enter, nb, 0)] == keys)
        if plant == "emit_reversed":
            won = won[:, ::-1]
        p, dd = np.nonzero(won)
        if plant == "emit_reversed":
            dd = 5 - dd
        ch = fr[p] + step[dd]
        dots = np.array([np.dot(DIRS[d] * dx, nrm[c]) for d, c in zip(dd, ch)])
        rsdf[ch] = rsdf[fr[p]] + dots
        hops[ch] = level + 1
        q1 = q0 + (0 if plant == "next_keys_no_offset" else len(fr))
        if claim_next and len(ch):
            offer(ch, q1)
        return ch, q0 + len(fr)

    level, q0 = 0, 0
    while len(front) and level <= V:
        if len(front) <= cap:
            # the single workgroup: claims at the start of each level
            while 0 < len(front) <= cap and level <= V:
                front, q0 = run_level(front, q0, level, True, False)
                level += 1
            continue
        # a multi-block batch: a claim kernel, then count + emit-and-claim
        for b in range(batch):
            front, q0 = run_level(front, q0, level, b == 0, True)
            level += 1
    return rsdf.reshape(labels.shape), hops.reshape(labels.shape)


def field(shape, density, seed):
    rng = np.random.default_rng(seed)
    noise = ndimage.gaussian_filter(rng.normal(size=shape), 1.2)
    mask = noise > np.quantile(noise, 1 - density)
    n = rng.normal(size=shape + (3,))
    n /= np.linalg.norm(n, axis=-1, keepdims=True)
    n[~mask] = np.nan
    lab, nl = ndimage.label(mask, structure=SIX)
    lab = lab + 1
    # one seed per component with two voxels or more, at a random voxel of it
    seeds = []
    for c in range(2, nl + 2):
        idx = np.flatnonzero(lab.ravel() == c)
        if len(idx) >= 2:
            seeds.append(int(rng.choice(idx)))
    return n, lab, np.asarray(seeds, np.int64)


FIELDS = [((9, 8, 7), 0.5, 1), ((14, 12, 10), 0.35, 2), ((12, 12, 12), 0.7, 3)]


@pytest.fixture(scope="module")
def cases(oracle):
    out = []
    for shape, density, seed in FIELDS:
        n, lab, seeds = field(shape, density, seed)
        out.append((n, lab, seeds, oracle.bfs_fifo(n, lab, seeds, 1e-3)))
    return out


def same(got, want):
    return (got[0].view(np.uint64).tobytes() == want[0].view(np.uint64).tobytes()
            and np.array_equal(got[1], want[1]))


@pytest.mark.parametrize("cap,batch", [(0, 1), (0, 32), (1, 2), (3, 1), (8, 3), (40, 32),
                                       (10 ** 9, 1)])
def test_level_search_equals_fifo(cases, cap, batch):
    assert max(len(c[2]) for c in cases) >= 2  # components share the frontier
    for n, lab, seeds, want in cases:
        assert same(level_search(n, lab, seeds, 1e-3, cap, batch), want)


@pytest.mark.parametrize("plant", ["no_queue_offset", "seed_unclaimed", "emit_reversed",
                                   "next_keys_no_offset"])
def test_planted_slips_are_caught(cases, plant):
    # the early next-level offers exist only on the multi-block path (cap 0)
    caught = any(not same(level_search(n, lab, seeds, 1e-3, 0, 4, plant=plant), want)
                 for n, lab, seeds, want in cases)
    assert caught, plant


# ---------------------------------------------------------------------------
# The single-pass scan of k_level (decoupled look-back), on the CPU.
# ---------------------------------------------------------------------------
AGG, PRE, VAL = 1 << 30, 2 << 30, (1 << 30) - 1


def lookback_scan(counts, resident, window, rng, plant=None, max_steps=200000):
    """Run k_level's scan for blocks whose child counts are `counts` (indexed
    by blockIdx) under a random schedule: at most `resident` blocks run at a
    time, a block takes its scan position from a counter when it starts
    (plant "blockidx": its blockIdx instead), publishes its aggregate, then
    reads the status words of the `window` positions before it, nearest
    first, and either sums up to the nearest inclusive prefix (all nearer
    words being ready), moves one window back when all are aggregates, or
    waits (reads again later).  Returns each scan position's exclusive
    prefix, or None if the schedule stops making progress."""
    G = len(counts)
    status = [0] * G
    counter = 0
    # the hardware need not start blocks in blockIdx order: a shuffled order
    waiting = [int(b) for b in rng.permutation(G)]
    running = {}                      # blockIdx -> state dict
    done = {}
    steps = 0
    while len(done) < G:
        steps += 1
        if steps > max_steps:
            return None
        while waiting and len(running) < resident:
            b = waiting.pop(0)
            pos = b if plant == "blockidx" else counter
            counter += 1
            running[b] = {"pos": pos, "phase": 0, "base": pos - 1, "excl": 0}
        b = rng.choice(sorted(running))
        s = running[b]
        pos, tot = s["pos"], counts[s["pos"]]
        if s["phase"] == 0:                       # pass 1 done: publish
            if pos == 0 or plant == "early_prefix":
                status[pos] = PRE | tot
                if pos == 0:
                    done[pos] = 0
                    del running[b]
                    continue
            else:
                status[pos] = AGG | tot
            s["phase"] = 1
            continue
        lanes = [status[j] if j >= 0 else PRE for j in range(s["base"], s["base"] - window, -1)]
        states = [w & ~VAL for w in lanes]
        pre = [i for i, st in enumerate(states) if st == PRE]
        if plant == "stop_at_agg":
            pre = [i for i, st in enumerate(states) if st != 0]
        upto = pre[0] if pre else window - 1
        if all(st != 0 for st in states[:upto + 1]):
            s["excl"] += sum(w & VAL for w in lanes[:upto + 1])
            if pre:
                status[pos] = PRE | (s["excl"] + tot)
                done[pos] = s["excl"]
                del running[b]
            else:
                s["base"] -= window
        # else: not ready, read again on a later turn
    return [done[p] for p in range(G)]


@pytest.mark.parametrize("G,resident,window", [(1, 1, 4), (13, 13, 4), (13, 3, 4), (64, 7, 64),
                                               (70, 9, 8), (40, 1, 4)])
def test_lookback_scan_every_schedule(G, resident, window):
    rng = np.random.default_rng(G * 100 + resident)
    for trial in range(30):
        counts = rng.integers(0, 50, size=G)
        counts[rng.uniform(size=G) < 0.3] = 0
        got = lookback_scan(list(counts), resident, window, rng)
        assert got is not None
        assert got == list(np.concatenate([[0], np.cumsum(counts)[:-1]]))


@pytest.mark.parametrize("plant", ["blockidx", "stop_at_agg", "early_prefix"])
def test_lookback_planted_slips_are_caught(plant):
    """Scan order by blockIdx can wait on a block that has not started
    (dispatch order is not scan order when blocks are taken out of order, and
    with few resident slots the wait never ends); stopping at an aggregate or
    publishing an inclusive prefix before the look-back gives wrong offsets."""
    rng = np.random.default_rng(5)
    caught = False
    for trial in range(40):
        counts = list(rng.integers(1, 50, size=12))
        got = lookback_scan(counts, 3, 4, rng, plant=plant, max_steps=20000)
        want = list(np.concatenate([[0], np.cumsum(counts)[:-1]]))
        if got != want:
            caught = True
            break
    assert caught, plant
