"""Restatement of MDManager's units of one process: the particle boxes of a unit grid,
the halo relay between units and each unit's pair list, kicks and observers, in numpy.

TEST INFRASTRUCTURE ONLY: imported by tests/, never by the product package.

The reference (kaityo256/mdacp) splits the box into one MDUnit per OpenMP thread
(MDManager::MDManager, mdmanager.cc:25-45; ParaInfo's grid).  Each unit owns the particles
inside its box [s, e) (MDUnit::MDUnit, mdunit.cc:8-27; MDRect::IsInside, mdrect.h:35-38)
and receives ghosts through six sequential passes, one per direction
(MDManager::MakePairList, mdmanager.cc:262-280): every unit marks the particles it holds
so far (owned and ghosts) that lie strictly inside the border strip of that direction
(MDUnit::FindBorderParticles, mdunit.cc:125-135; MDRect::IsInsideEdge, mdrect.cc:24-47),
sends their positions, shifted by +-L when the neighbour lies across the periodic boundary
(MDUnit::MakeBufferForBorderParticles, mdunit.cc:138-176; ParaInfo::IsOverBoundary,
parainfo.cc:256-275), and every unit appends what its neighbour on the opposite side sent
(MDManager::SendBorderParticlesSub, mdmanager.cc:297-348; MDUnit::ReceiveBorderParticles,
mdunit.cc:179-191).  The pair list of a unit holds every pair of its particles within the
search length that is not a pair of two ghosts (MeshList::SearchMesh, meshlist.cc:220-271,
r2 > S2 rejected); forces act on the pairs within the cutoff (ForceCalculator,
fcalculator.cc:53-77) and the observers weigh a pair with a ghost by 1/2
(observer.cc:18-79).

Here the relay is carried out literally, each ghost keeping the id of the particle it
images and the shift it picked up, and the pair search is brute force (a k-d tree for
candidates, then the reference's r2 test in its operation order), which the reference's
cell search equals whenever every cell is at least a search length wide.
"""
import numpy as np

# the reference's directions (mdconfig.h:25): D_LEFT, D_RIGHT, D_BACK, D_FORWARD, D_DOWN, D_UP
N_DIR = 6


def unit_boxes(L, grid):
    """Grid positions and boxes of the units, id = x + gx * (y + gy * z): per axis
    ul = L / g, s = ul * pos, e = s + ul (mdunit.cc:17-24)."""
    L = [float(v) for v in L]
    out = []
    for z in range(grid[2]):
        for y in range(grid[1]):
            for x in range(grid[0]):
                pos = (x, y, z)
                s = [L[a] / float(grid[a]) * float(pos[a]) for a in range(3)]
                e = [s[a] + L[a] / float(grid[a]) for a in range(3)]
                out.append((pos, np.array(s), np.array(e)))
    return out


def distribute(L, grid, q, ids):
    """Owned particles of every unit: those inside its box, in the given order
    (MDUnit::AddParticle, mdunit.cc:36-49)."""
    units = []
    for pos, s, e in unit_boxes(L, grid):
        inside = np.all((q >= s) & (q < e), axis=1)
        units.append(dict(pos=pos, s=s, e=e, ids=ids[inside].copy(), q=q[inside].copy()))
    return units


def _neighbor(units, grid, u, direction):
    """ParaInfo::GetNeighborID (parainfo.cc:191-224)."""
    a, step = direction // 2, (1 if direction % 2 else -1)
    p = list(units[u]["pos"])
    p[a] = (p[a] + step) % grid[a]
    for k, v in enumerate(units):
        if tuple(v["pos"]) == tuple(p):
            return k
    raise ValueError("grid not filled")


def relay(units, grid, L, SL):
    """Ghosts of every unit after the six passes: for unit u, arrays gid (id of the imaged
    particle), gq (position) and gshift (shift in units of L per axis), in the order the
    reference appends them."""
    L = np.asarray(L, dtype=np.float64)
    hold = []  # per unit: (ids, q, shift) of owned + ghosts so far
    for v in units:
        n = len(v["ids"])
        hold.append([v["ids"].copy(), v["q"].copy(), np.zeros((n, 3), dtype=np.int64)])
    pn = [len(v["ids"]) for v in units]
    for direction in range(N_DIR):
        a = direction // 2
        lower = direction % 2 == 0
        sent = []
        for u, v in enumerate(units):
            ids, q, sh = hold[u]
            x = q[:, a]
            if lower:
                border = (x < v["s"][a] + SL) & (x > v["s"][a])
            else:
                border = (x > v["e"][a] - SL) & (x < v["e"][a])
            over = (v["pos"][a] - 1 < 0) if lower else (v["pos"][a] + 1 >= grid[a])
            qs = q[border].copy()
            shs = sh[border].copy()
            if over:
                if lower:
                    qs[:, a] = qs[:, a] + L[a]
                    shs[:, a] += 1
                else:
                    qs[:, a] = qs[:, a] + (0.0 - L[a])  # diff[X] -= L[X]: 0.0 - L, then q + diff
                    shs[:, a] -= 1
            sent.append((ids[border].copy(), qs, shs))
        for u in range(len(units)):
            src = _neighbor(units, grid, u, direction ^ 1)  # the unit whose `direction` neighbour is u
            ids, q, sh = hold[u]
            sids, sq, ssh = sent[src]
            hold[u] = [np.concatenate([ids, sids]), np.concatenate([q, sq]), np.concatenate([sh, ssh])]
    out = []
    for u in range(len(units)):
        ids, q, sh = hold[u]
        out.append(dict(gid=ids[pn[u]:], gq=q[pn[u]:], gshift=sh[pn[u]:]))
    return out


def shift_code(sh):
    sh = np.asarray(sh)
    return (sh[:, 0] + 1) + 3 * (sh[:, 1] + 1) + 9 * (sh[:, 2] + 1)


def _r2(qa, qb):
    dx = qa[:, 0] - qb[:, 0]
    dy = qa[:, 1] - qb[:, 1]
    dz = qa[:, 2] - qb[:, 2]
    return (dx * dx + dy * dy) + dz * dz


def unit_pairs(unit, ghosts, SL):
    """Pairs of one unit within SL (r2 <= SL^2), no ghost-ghost pair, as index pairs (i, j)
    into owned + ghosts with i owned, i < j."""
    from scipy.spatial import cKDTree

    q = np.concatenate([unit["q"], ghosts["gq"]])
    pn = len(unit["q"])
    if len(q) < 2:
        return np.zeros((0, 2), dtype=np.int64)
    cand = cKDTree(q).query_pairs(SL * (1.0 + 1e-9), output_type="ndarray")
    if len(cand) == 0:
        return np.zeros((0, 2), dtype=np.int64)
    i, j = np.minimum(cand[:, 0], cand[:, 1]), np.maximum(cand[:, 0], cand[:, 1])
    keep = (i < pn) & ~(_r2(q[i], q[j]) > SL * SL)
    return np.stack([i[keep], j[keep]], axis=1)


def canonical_unit_pairs(unit, ghosts, pairs):
    """The unit's pairs as (id, id, image code) triples: owned-owned as (smaller id, larger
    id, 13), owned-ghost as (owned id, imaged id, the ghost's shift code)."""
    pn = len(unit["q"])
    i, j = pairs[:, 0], pairs[:, 1]
    ids = unit["ids"]
    own = j < pn
    a = np.where(own, np.minimum(ids[i], ids[np.minimum(j, pn - 1)]), ids[i])
    b = np.where(own, np.maximum(ids[i], ids[np.minimum(j, pn - 1)]), 0)
    g = np.maximum(j - pn, 0)
    b = np.where(own, b, ghosts["gid"][g] if len(ghosts["gid"]) else 0)
    code = np.where(own, 13, shift_code(ghosts["gshift"])[g] if len(ghosts["gid"]) else 13)
    return np.stack([a, b, code], axis=1).astype(np.int64)


def lj_constants(rc):
    """C0, C2 of the smoothed potential (variables.cc:16-28)."""
    s2 = 1.0 / (rc * rc)
    s6 = s2 * s2 * s2
    s8 = s6 * s2
    s12 = s6 * s6
    s14 = s12 * s2
    C2 = 6.0 * s14 - 3.0 * s8
    C0 = -s12 + s6 - C2 / s2
    return C0, C2


def unit_kicks_and_observers(unit, ghosts, pairs, rc, dt):
    """Kick F*dt on every owned particle from the unit's pairs within rc
    (fcalculator.cc:53-77), its sum of pair-kick magnitudes, and the unit's potential
    energy and virial (observer.cc:18-79; a pair with a ghost weighs 1/2)."""
    C0, C2 = lj_constants(rc)
    q = np.concatenate([unit["q"], ghosts["gq"]])
    pn = len(unit["q"])
    i, j = pairs[:, 0], pairs[:, 1]
    d = q[j] - q[i]
    r2 = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
    inside = ~(r2 > rc * rc)
    i, j, d, r2 = i[inside], j[inside], d[inside], r2[inside]
    r6 = r2 * r2 * r2
    df = ((24.0 * r6 - 48.0) / (r6 * r6 * r2) + C2 * 8.0) * dt
    f = df[:, None] * d
    dp = np.zeros((pn, 3))
    mag = np.zeros(pn)
    np.add.at(dp, i, f)
    np.add.at(mag, i, np.abs(df) * np.sqrt(r2))
    own = j < pn
    np.add.at(dp, j[own], -f[own])
    np.add.at(mag, j[own], np.abs(df[own]) * np.sqrt(r2[own]))
    w = np.where(own, 1.0, 0.5)
    e = 4.0 * (1.0 / (r2 * r2 * r2 * r2 * r2 * r2) - 1.0 / (r2 * r2 * r2) + C2 * r2 + C0)
    phi = ((24.0 * r6 - 48.0) / (r6 * r6 * r2) + C2 * 8.0) * r2
    return dp, mag, float(np.sum(w * e)), float(np.sum(w * phi) / 3.0)
