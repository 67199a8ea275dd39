"""find_neighbors_gpu_pbc with the interface of the reference's
core/neighbor_search/gpu_kdtree.py:56-87.

Returns the half list {(i, j): i < j, d^2 < cutoff^2} in lexicographic order
(edge_index int64 [2, E]) and the distances |minimum_image(x_i - x_j)| in the
dtype of `positions`. The pair test is the reference's arithmetic in that
dtype (core/box.py:122-130, gpu_kdtree.py:130-138), evaluated by the cell-list
kernels (any box shape, no neighbor cap) instead of an O(N^2) scan; the
squared radius follows the reference's types: a tensor cutoff is squared in
its own dtype, a Python float in double precision and compared in the
positions' dtype.
"""
import math

import numpy as np
import torch

from . import _lib
from ._lib import call, ptr, stream_ptr
from .box import Box
from .engine import perpendicular_widths


def _radius_sq(cutoff, pos_dtype):
    if torch.is_tensor(cutoff):
        c = cutoff.detach().cpu()
        rl2 = float((c * c).item())            # squared in the tensor's dtype
    else:
        rl2 = float(cutoff) * float(cutoff)
    if pos_dtype == torch.float32:
        rl2 = float(np.float32(rl2))
    return rl2


def find_neighbors_gpu_pbc(positions, cutoff, box, batch_size=2000, max_nbr=None):
    """gpu_kdtree.py:56-87. positions: [N,3] float32/float64 on the GPU;
    box: Box or a cubic box length."""
    if not torch.is_tensor(positions) or positions.device.type != "cuda":
        raise RuntimeError("find_neighbors_gpu_pbc runs on the GPU: pass a CUDA (HIP) tensor")
    if positions.dtype not in (torch.float32, torch.float64):
        raise TypeError("positions must be float32 or float64")
    pos = positions.detach().contiguous()
    n = pos.shape[0]
    if not isinstance(box, Box):
        box = Box(float(box), device=pos.device)
    H = box.H.detach().cpu().numpy()
    Hinv = box.H_inv.detach().cpu().numpy()
    rl2 = _radius_sq(cutoff, pos.dtype)
    rl = math.sqrt(rl2)
    P = _lib.Replica()
    for k in range(9):
        P.h[k] = float(H.flat[k])
        P.hinv[k] = float(Hinv.flat[k])
    P.ortho = int(box.is_orthogonal)
    P.volume = box.volume
    P.rl2, P.rl = rl2, rl
    P.atom_base, P.natoms = 0, n
    nc = np.maximum(1, np.floor(perpendicular_widths(H) / (rl * (1 + 1e-6)))).astype(np.int64)
    nc = np.minimum(nc, 1024)
    P.cell_base, P.cell_cap = 0, int(nc.prod())
    dev = pos.device
    rep = _lib.struct_tensor([P], dev)
    ncells = int(nc.prod())
    i32 = dict(dtype=torch.int32, device=dev)
    if n == 0:
        return torch.zeros((2, 0), dtype=torch.long, device=dev), torch.zeros(0, device=dev)
    if max_nbr is None:
        rho = n / box.volume
        max_nbr = int(math.ceil((1.5 * 4.0 / 3.0 * math.pi * rl ** 3 * rho + 32) / 8) * 8)
    dt = 0 if pos.dtype == torch.float32 else 1
    while True:
        work = [torch.zeros(n, **i32) for _ in range(3)]
        cnt = torch.zeros(ncells + 1, **i32)
        start = torch.zeros(ncells + 1, **i32)
        nbr = torch.zeros((max_nbr, n), **i32)
        nnbr = torch.zeros(n, **i32)
        ovf = torch.zeros(1, **i32)
        call("sim_nl_build_xyz", ptr(rep), ptr(pos), dt, n, ptr(work[0]), ptr(work[1]), ptr(cnt),
             ptr(start), ptr(work[2]), ncells, ptr(nbr), ptr(nnbr), max_nbr, ptr(ovf), stream_ptr())
        if int(ovf.item()) == 0:
            break
        max_nbr = 2 * max_nbr                     # dense input: retry with room for it
        if max_nbr > n:
            max_nbr = n
    from .engine import half_list_from_full
    return half_list_from_full(nbr, nnbr, max_nbr, rep, pos)
