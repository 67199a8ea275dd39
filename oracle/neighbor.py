"""Oracle: Verlet half list and skin trigger (test infrastructure only).

Restates core/neighbor_search/gpu_kdtree.py:90-151 (brute force over row
blocks of 2000, keep global i<j with d^2 < r_l^2, lexicographic (i, j)
order), io_utils/reader.py:477-484 (i<j re-filter, distances recomputed as
the norm of the minimum image, reader.py:364-376) and the skin/2 trigger of
io_utils/reader.py:136-163.

List radius: the reference forms r_l = cutoff + skin as a float32 tensor sum
(reader.py:50, 77-78) and squares it in float32 (gpu_kdtree.py:99); the
comparison is then made in the dtype of the squared distances.

Squared distance order: (x*x + y*y) + z*z, one rounding per op.

pinned by: brute force over 27 images, shell counts of perfect simple-cubic,
bcc and fcc lattices (tests/test_oracle_geometry.py) and the reference's own
pair sets for seeded FP32 and FP64 inputs (tests/golden/).
"""
import numpy as np

from . import box as obox


def list_radius_sq(cutoff, skin):
    """(fp32(cutoff) + fp32(skin))^2 in float32, as reader.py:50,77-78 and
    gpu_kdtree.py:99 form it."""
    rl = np.float32(cutoff) + np.float32(skin)
    return np.float32(rl * rl)


def sqnorm(r):
    return (r[..., 0] * r[..., 0] + r[..., 1] * r[..., 1]) + r[..., 2] * r[..., 2]


def brute_force_half_list(pos, H, rl2, Hinv=None, ortho=None, block=2000):
    """gpu_kdtree.py:90-151 + reader.py:477-484.

    pos: [N,3] float32 or float64; rl2: squared list radius (float32 value).
    Returns (edge_index int64 [2,E] lexicographic, dist [E] in pos.dtype).
    """
    pos = np.asarray(pos)
    n = pos.shape[0]
    dt = pos.dtype
    rl2 = dt.type(np.float32(rl2))
    ei, ej = [], []
    jg = np.arange(n)
    for i0 in range(0, n, block):
        i1 = min(i0 + block, n)
        rij = pos[i0:i1, None, :] - pos[None, :, :]
        rij = obox.minimum_image(rij, H, Hinv, ortho)
        d2 = sqnorm(rij)
        ig = np.arange(i0, i1)[:, None]
        bi, bj = np.nonzero((ig < jg[None, :]) & (d2 < rl2))
        ei.append(bi + i0)
        ej.append(bj)
    if not ei:
        return np.zeros((2, 0), np.int64), np.zeros(0, dt)
    ei = np.concatenate(ei).astype(np.int64)
    ej = np.concatenate(ej).astype(np.int64)
    return np.stack([ei, ej]), edge_distances(pos, np.stack([ei, ej]), H, Hinv, ortho)


def edge_distances(pos, edge_index, H, Hinv=None, ortho=None):
    """reader.py:364-376: |minimum_image(pos[i]-pos[j])|."""
    pos = np.asarray(pos)
    rij = obox.minimum_image(pos[edge_index[0]] - pos[edge_index[1]], H, Hinv, ortho)
    return np.sqrt(sqnorm(rij))


def cell_half_list(pos, H, rl2, Hinv=None, ortho=None):
    """Same pair set as brute_force_half_list, found through a cell grid.

    Cells are parallelepipeds in fractional coordinates whose perpendicular
    widths are at least r_l; every candidate is accepted or rejected with the
    exact brute-force criterion, so the result equals the brute force (this
    equality is tested), while the cost is O(N) instead of O(N^2). Used for
    configurations too large for the brute force.
    """
    pos = np.asarray(pos)
    n = pos.shape[0]
    dt = pos.dtype
    H = np.asarray(H, np.float64)
    if Hinv is None:
        Hinv = obox.inverse(H)
    if ortho is None:
        ortho = obox.is_orthogonal(H)
    rl = float(np.sqrt(np.float64(np.float32(rl2))))
    w = obox.perpendicular_widths(H)
    nc = np.maximum(1, np.floor(w / (rl * (1.0 + 1e-6))).astype(np.int64))
    s = pos.astype(np.float64) @ np.asarray(Hinv, np.float64)
    s = s - np.floor(s)
    c = np.minimum((s * nc).astype(np.int64), nc - 1)
    cid = (c[:, 2] * nc[1] + c[:, 1]) * nc[0] + c[:, 0]
    order = np.argsort(cid, kind="stable")
    ncell = int(nc.prod())
    counts = np.bincount(cid, minlength=ncell)
    start = np.concatenate([[0], np.cumsum(counts)])
    offs = []
    for k in range(3):
        o = [0] if nc[k] == 1 else ([0, 1] if nc[k] == 2 else [-1, 0, 1])
        offs.append(o)
    rl2t = dt.type(np.float32(rl2))
    ei_all, ej_all = [], []
    for dz in offs[2]:
        for dy in offs[1]:
            for dx in offs[0]:
                cc = c + np.array([dx, dy, dz])
                cc = np.mod(cc, nc)
                ncid = (cc[:, 2] * nc[1] + cc[:, 1]) * nc[0] + cc[:, 0]
                cnt = counts[ncid]
                tot = int(cnt.sum())
                if tot == 0:
                    continue
                ii = np.repeat(np.arange(n), cnt)
                base = np.repeat(start[ncid], cnt)
                within = np.arange(tot) - np.repeat(np.cumsum(cnt) - cnt, cnt)
                jj = order[base + within]
                keep = ii < jj
                ii, jj = ii[keep], jj[keep]
                rij = obox.minimum_image(pos[ii] - pos[jj], H, Hinv, ortho)
                m = sqnorm(rij) < rl2t
                ei_all.append(ii[m])
                ej_all.append(jj[m])
    ei = np.concatenate(ei_all).astype(np.int64)
    ej = np.concatenate(ej_all).astype(np.int64)
    key = ei * n + ej
    key = np.unique(key)
    ei, ej = key // n, key % n
    ed = np.stack([ei, ej])
    return ed, edge_distances(pos, ed, H, Hinv, ortho)


def full_list_from_half(edge_index, n):
    """Per-atom full neighbor sets {i: sorted j} from a half list."""
    i, j = edge_index
    a = np.concatenate([i, j])
    b = np.concatenate([j, i])
    o = np.lexsort((b, a))
    a, b = a[o], b[o]
    counts = np.bincount(a, minlength=n)
    return counts, b


def needs_rebuild(pos_new, pos_last, H, skin, Hinv=None, ortho=None):
    """reader.py:136-149: rebuild when max |mi(x - x_last)| > skin/2.

    The comparison mirrors `max_displacement > self.skin_thickness / 2`
    with both operands in float32 (skin is a float32 tensor)."""
    if pos_last is None:
        return True
    d = obox.minimum_image(np.asarray(pos_new) - np.asarray(pos_last), H, Hinv, ortho)
    m = np.max(np.sqrt(sqnorm(d))) if len(d) else 0.0
    return bool(m > np.float32(skin) / np.float32(2.0))
