"""Drop-in for ``sdf._construct_rsdf_components`` (sdf.py:122-224).

Same signature and result: a float64 array shaped like the image holding
the RSDF of every valid connected component and NaN elsewhere.  The
breadth-first search itself (sdf.py:195-222, the interpreter-bound loop)
runs on the GPU (``mmnorm_amd.ops.rsdf_bfs``), bit-exact with the
reference's FIFO order.  The start voxel of each component is picked on the
host with the reference's own NumPy expressions (sdf.py:182-191: mean of the
component's coordinates, Euclidean distance, first minimum), so ties resolve
identically.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib, ops


def component_seeds(all_normals, coords_full, labeled, valid_labels):
    """Linear start voxel of each component in ``valid_labels`` order (sdf.py:180-191).

    Labels repeated in ``valid_labels`` are searched once (a second search
    from an already visited seed discovers nothing in the reference).
    """
    ok = ~np.isnan(all_normals[:, :, :, 0])
    seeds, seen = [], set()
    for comp in valid_labels:
        key = comp.item() if hasattr(comp, "item") else comp
        if key in seen:
            continue
        seen.add(key)
        mask = np.logical_and(labeled == comp, ok)
        idxs = np.argwhere(mask)
        if idxs.shape[0] == 0:
            raise ValueError(f"component {key} has no voxel with a valid normal")
        coords_sel = coords_full[mask, :]
        mean_coord = np.mean(coords_sel, axis=0)[None, :]
        dists = np.linalg.norm(coords_sel - mean_coord, axis=1)
        i0, j0, k0 = idxs[int(np.argmin(dists))]
        seeds.append(np.ravel_multi_index((i0, j0, k0), labeled.shape))
    return np.asarray(seeds, dtype=np.int64)


def rsdf_components(all_normals, labeled, seeds, delta_x, device=None, return_hops=False):
    """Run the GPU BFS from explicit seeds; NumPy in and out."""
    _lib.require_gpu()
    dev = device or torch.device("cuda", torch.cuda.current_device())
    labeled = np.asarray(labeled)
    if labeled.size and (labeled.min() < np.iinfo(np.int32).min or
                         labeled.max() > np.iinfo(np.int32).max):
        raise ValueError("labels do not fit in int32")
    nrm = torch.from_numpy(np.ascontiguousarray(all_normals, dtype=np.float64)).to(dev)
    lab = torch.from_numpy(np.ascontiguousarray(labeled, dtype=np.int32)).to(dev)
    sd = torch.from_numpy(np.ascontiguousarray(seeds, dtype=np.int32)).to(dev)
    rsdf, hops = ops.rsdf_bfs(nrm, lab, sd, float(delta_x))
    rsdf = rsdf.cpu().numpy()
    return (rsdf, hops.cpu().numpy()) if return_hops else rsdf


def _construct_rsdf_components(all_normals, coords_full, labeled, valid_labels, x_locs):
    """RSDF per connected component (sdf.py:122-224), BFS on the GPU."""
    delta_x = x_locs[1] - x_locs[0]
    seeds = component_seeds(all_normals, coords_full, labeled, valid_labels)
    return rsdf_components(all_normals, labeled, seeds, delta_x)
