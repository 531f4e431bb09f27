"""Drop-ins for the RSDF stage of ``sdf.py``: ``_separate_binary_image``
(sdf.py:33-86), ``_find_connected_components`` (sdf.py:88-120) and
``_construct_rsdf_components`` (sdf.py:122-224).  Same signatures, NumPy in
and out.

* The Li threshold of each exponent (sdf.py:46-70) runs on the GPU
  (``ops.li_threshold``); the binary-image volume test (sdf.py:77-84) uses
  ``volume_fn``: by default ``hull_face_box_volume``, open3d's documented
  minimal-box algorithm restated over scipy's qhull.
* The 6-connected labelling (cc3d, sdf.py:89) and the per-label areas
  (sdf.py:100) run on the GPU (``ops.ccl6``, ``ops.label_sizes``); the
  reference's size rule (sdf.py:92-119) is applied to the areas as written.
* The breadth-first search (sdf.py:195-222, the interpreter-bound loop) runs
  on the GPU (``ops.rsdf_bfs``), bit-exact with the reference's FIFO order.
  The start voxel of each component is picked with the reference's own
  NumPy expressions (sdf.py:182-191: mean of the component's coordinates,
  Euclidean distance, first minimum), so ties resolve identically: on the
  host for NumPy inputs (``component_seeds``), or from device fields with the
  whole-grid masks and gathers on the GPU (``component_seeds_gpu``).
"""
from __future__ import annotations

import os
import sys

import numpy as np
import torch

from . import _lib, ops

LI_EXPONENTS = (2, 2.5, 3.5, 4)     # sdf.py:73
LI_MAX_ITER = 1000                  # sdf.py:56-58
VOLUME_THRESHOLD = 0.03             # sdf.py:74 (m^3)


def _device():
    _lib.require_gpu()
    return torch.device("cuda", torch.cuda.current_device())


def pca_box_volume(points):
    """Volume of the box aligned with the principal axes of ``points``: a
    cheaper ``volume_fn`` for ``_separate_binary_image`` (equal to
    hull_face_box_volume for box-like point sets)."""
    p = np.asarray(points, dtype=np.float64)
    if p.shape[0] < 2:
        return 0.0
    c = p - p.mean(axis=0)
    _, _, vt = np.linalg.svd(c, full_matrices=False)
    q = c @ vt.T
    ext = q.max(axis=0) - q.min(axis=0)
    return float(np.prod(ext)) if ext.size == 3 else 0.0


def hull_face_box_volume(points, block=256):
    """Volume of open3d's "minimal" oriented bounding box (the volume test of
    sdf.py:79-81 calls ``get_minimal_oriented_bounding_box().volume()``),
    from the algorithm open3d documents for it: the convex hull of the
    points, and for each hull triangle the axis-aligned box of the hull's
    vertices in that triangle's frame -- an edge, the in-plane direction
    perpendicular to it, the normal -- the smallest of those volumes.  The
    hull is qhull's (scipy.spatial.ConvexHull), as open3d's is; every edge of
    a triangle is tried, so the volume is at most open3d's, whose frame takes
    one edge per triangle in its hull's vertex order.  Not the true minimal
    box: for a regular tetrahedron this gives sqrt(2)/2 edge^3, the true
    minimum is edge^3 / (2 sqrt 2).  A flat or degenerate set (no 3-D hull)
    has volume 0."""
    from scipy.spatial import ConvexHull
    from scipy.spatial import QhullError
    p = np.asarray(points, dtype=np.float64).reshape(-1, 3)
    if p.shape[0] < 4:
        return 0.0
    try:
        hull = ConvexHull(p)
    except QhullError:
        return 0.0
    hv = p[hull.vertices]
    tri = p[hull.simplices]                       # (T, 3, 3)
    w = np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0])
    w /= np.linalg.norm(w, axis=1, keepdims=True)
    best = np.inf
    for e in range(3):
        u = tri[:, (e + 1) % 3] - tri[:, e]
        u /= np.linalg.norm(u, axis=1, keepdims=True)
        v = np.cross(w, u)
        frames = np.stack([u, v, w], axis=1)      # (T, 3 axes, 3)
        for lo in range(0, len(frames), block):
            q = np.einsum("tak,nk->tan", frames[lo:lo + block], hv)
            ext = q.max(axis=2) - q.min(axis=2)
            best = min(best, float(np.min(np.prod(ext, axis=1))))
    return best


def _separate_binary_image(sum_img, coords_full, volume_fn=hull_face_box_volume):
    """Binary image of the SAR image (sdf.py:33-86): True where |sum_img|**exp is
    below the Li threshold, for the first exponent of (2, 2.5, 3.5, 4) whose Li
    iteration ends within 1000 callbacks and whose kept voxels span at most
    0.03 m^3 (``volume_fn`` of their coordinates); if no exponent passes the
    volume test, the last exponent's result if its threshold succeeded, else
    ``sys.exit(0)`` as the reference does."""
    dev = _device()
    z = torch.from_numpy(np.ascontiguousarray(sum_img, dtype=np.complex64)).to(dev)
    filter_idx, success = None, False
    for ex in LI_EXPONENTS:
        _, iters, _, below = ops.li_threshold(z, ex, max_iter=LI_MAX_ITER, below=True)
        # the reference's callback raises on its 1001st call: the initial
        # guess plus 1000 iterations
        success = iters < LI_MAX_ITER
        if success:
            filter_idx = below.cpu().numpy()
        points = coords_full[np.logical_not(filter_idx)] if filter_idx is not None else \
            coords_full.reshape(-1, 3)
        volume = volume_fn(points)
        if success and volume <= VOLUME_THRESHOLD:
            return filter_idx
    if success:
        return filter_idx
    sys.exit(0)


def size_rule(areas, dx):
    """Valid labels by the reference's size filter (sdf.py:92-119).

    ``areas[l - 1]`` is the voxel count of label l for l = 1..max+1 (the
    reference's ``np.arange(1, labeled.max() + 2)``).
    """
    labels = np.arange(1, len(areas) + 1)
    lo, lz = int(0.03 / dx), int(0.04 / dx)
    valid = labels[areas > lo ** 2 * lz]
    if len(valid) == 0:
        lo, lz = int(0.015 / dx), int(0.02 / dx)
        valid = labels[areas > lo ** 2 * lz]
    return valid


def find_components_gpu(normals):
    """GPU labelling of the non-NaN voxels of a device normal field.

    normals float64 device tensor (NX,NY,NZ,3).  Returns (labeled int32
    device tensor = component label + 1, background 1, as sdf.py:89;
    areas int64 NumPy array of labels 1..max+1).
    """
    mask = ~torch.isnan(normals[..., 0])
    lab, n = ops.ccl6(mask.contiguous())
    sizes = ops.label_sizes(lab, n).cpu().numpy()   # [0] background, [1..n]
    # labels 1..n+2 of the reference's +1 numbering: background (label 1)
    # has mask-sum 0, components 2..n+1 their sizes, label n+2 is empty
    areas = np.concatenate(([0], sizes[1:], [0])).astype(np.float64)
    return lab + 1, areas


def _find_connected_components(all_normals, x_locs):
    """6-connected components of the non-NaN normals and the valid labels
    (sdf.py:88-120); labels numbered in C-order first appearance, + 1."""
    dev = _device()
    nrm = torch.from_numpy(np.ascontiguousarray(all_normals, dtype=np.float64)).to(dev)
    labeled, areas = find_components_gpu(nrm)
    valid = size_rule(areas, x_locs[1] - x_locs[0])
    return labeled.cpu().numpy().astype(np.uint32), valid


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


def component_seeds_gpu(normals, labeled, valid_labels, x_locs, y_locs, z_locs):
    """``component_seeds`` on device fields, for coords_full built as the
    reference builds it (sdf.py:278-279: voxel (i, j, k) at
    (x_locs[i], y_locs[j], z_locs[k])).

    The whole-grid part runs on the GPU: the voxels of the listed labels with a
    normal, in C order (one mask, one compaction), then a stable sort by label,
    so each component's voxels form a run in the C order of
    ``np.argwhere``, and their coordinates, gathered from the axes (equal to
    coords_full[mask]).  Only those runs come to the host, where the start
    voxel follows sdf.py:182-191 with the same NumPy expressions on the same
    values: the mean is NumPy's sequential sum in C order, which a parallel
    reduction would not reproduce bit for bit.

    normals float64 (NX,NY,NZ,3) and labeled int32 (NX,NY,NZ) tensors on one
    device (the GPU's in the pipeline; the same torch code runs on CPU
    tensors).  Returns int64 NumPy linear indices in ``valid_labels`` order,
    repeated labels searched once.
    """
    if normals.dim() != 4 or normals.shape[3] != 3 or normals.shape[:3] != labeled.shape:
        raise ValueError("normals must be (NX,NY,NZ,3) and labeled (NX,NY,NZ)")
    nx, ny, nz = labeled.shape
    xl, yl, zl = (np.asarray(a, dtype=np.float64) for a in (x_locs, y_locs, z_locs))
    if (len(xl), len(yl), len(zl)) != (nx, ny, nz):
        raise ValueError("axis lengths do not match the grid")
    order, seen = [], set()
    for comp in np.asarray(valid_labels).ravel():
        key = int(comp)
        if key not in seen:
            seen.add(key)
            order.append(key)
    if not order:
        return np.zeros(0, np.int64)
    flat = labeled.reshape(-1)
    kmin, kmax = min(order), max(order)
    if kmax - kmin < 1 << 24:
        # a flag per label of [kmin, kmax]: one gather per voxel
        lut = torch.zeros(kmax - kmin + 2, dtype=torch.bool, device=labeled.device)
        lut[torch.tensor(order, device=labeled.device) - kmin] = True
        rel = flat.long() - kmin
        # labels outside [kmin, kmax] read the last flag, which stays False
        rel = torch.where((rel < 0) | (rel > kmax - kmin), kmax - kmin + 1, rel)
        member = lut[rel]
    else:
        member = torch.isin(flat, torch.tensor(order, dtype=labeled.dtype,
                                               device=labeled.device))
    sel = member & ~torch.isnan(normals.reshape(-1, 3)[:, 0])
    lin = torch.nonzero(sel).reshape(-1)
    labs = flat[lin]
    if len(order) > 1:
        labs, perm = torch.sort(labs, stable=True)
        lin = lin[perm]
    # each selected voxel's coordinates, gathered from the axes (exact copies)
    i = torch.div(lin, ny * nz, rounding_mode="floor")
    r = lin - i * (ny * nz)
    j = torch.div(r, nz, rounding_mode="floor")
    k = r - j * nz
    dev = labeled.device
    ax = [torch.from_numpy(a).to(dev) for a in (xl, yl, zl)]
    sel_coords = torch.stack([ax[0][i], ax[1][j], ax[2][k]], dim=1).cpu().numpy()
    lin = lin.cpu().numpy()
    labs = labs.cpu().numpy()
    seeds = []
    for key in order:
        lo, hi = np.searchsorted(labs, key, "left"), np.searchsorted(labs, key, "right")
        if hi == lo:
            raise ValueError(f"component {key} has no voxel with a valid normal")
        coords_sel = sel_coords[lo:hi]
        mean_coord = np.mean(coords_sel, axis=0)[None, :]
        dists = np.linalg.norm(coords_sel - mean_coord, axis=1)
        seeds.append(int(lin[lo + int(np.argmin(dists))]))
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


def compute_rsdf_file(normal_path, sdf_path, volume_fn=hull_face_box_volume, params=None):
    """The RSDF stage of compute_rsdf (sdf.py:255-300) on explicit paths:
    read normal_data.npz, the binary image (its voxels' normals set to NaN),
    unit normals, the components and their RSDF, and write sdf.npz with the
    keys of _save_sdf_file (sdf.py:226-253; the 77 GHz 'almost-square'
    chirp's wavelengths from ``params``, image_processor.PARAMS by default).
    Returns the written dictionary."""
    from . import formats, image_processor
    sum_img, (xl, yl, zl), antenna_locs, channels, normals = formats.load_normal_data(
        normal_path)
    m1, m2, m3 = np.meshgrid(yl, xl, zl)
    coords_full = np.concatenate((m2[..., None], m1[..., None], m3[..., None]), axis=-1)
    filter_idx = _separate_binary_image(sum_img, coords_full, volume_fn)
    all_normals = np.array(normals, copy=True)
    all_normals[filter_idx] = np.nan
    all_normals /= np.linalg.norm(all_normals, axis=-1)[..., np.newaxis].repeat(3, axis=-1)
    labeled, valid_labels = _find_connected_components(all_normals, xl)
    rsdf = _construct_rsdf_components(all_normals, coords_full, labeled, valid_labels, xl)
    p = image_processor.get_radar_parameters("77_ghz", False, "almost-square", params)
    data = formats.sdf_record(sum_img, xl, yl, zl, antenna_locs, channels, all_normals, rsdf,
                              labeled, valid_labels, coords_full, p["min_f"], p["max_f"],
                              p["num_samples"])
    formats.save_sdf_file(sdf_path, data)
    return data


def compute_rsdf(obj_id, name, is_los, ext, overwrite_existing, root=".", **kw):
    """compute_rsdf (sdf.py:255-300) over the dataset's folder scheme under
    ``root`` (experiment 2 and the angle folder that holds it, as there):
    normals/normal_data<ext>.npz in,
    sdf/sdf<ext>.npz out; skipped when the output exists and
    ``overwrite_existing`` is false."""
    from . import formats
    angles = formats.find_obj_angles(root, obj_id, name, "2")
    proc = formats.dataset_path(root, obj_id, name, "77_ghz", is_sim=False, is_los=is_los,
                                exp_num="2", angles=angles, is_processed=True)
    out = f"{proc}/sdf/sdf{ext}.npz"
    if not overwrite_existing and os.path.exists(out):
        print("The output file already exists and overwrite_existing is not set: skipping.")
        return None
    return compute_rsdf_file(f"{proc}/normals/normal_data{ext}.npz", out, **kw)


def main(argv=None):
    """The command line of sdf.py:290-310 (plus --root, the dataset folder)."""
    import argparse
    ap = argparse.ArgumentParser(description="RSDF of an object's surface normals on the GPU")
    ap.add_argument("--name", default="hammer")
    ap.add_argument("--id", default="048")
    ap.add_argument("--ext", default="")
    ap.add_argument("--is_los", default="y")
    ap.add_argument("--overwrite_existing", default="n")
    ap.add_argument("--root", default=".", help="folder holding the dataset")
    a = ap.parse_args(argv)
    yes = lambda v: v in ("y", "Y")
    compute_rsdf(a.id, a.name, yes(a.is_los), a.ext, yes(a.overwrite_existing), root=a.root)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
