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
