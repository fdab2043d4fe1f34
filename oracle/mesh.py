"""RCM ordering and mesh geometry (oracle; test infrastructure only).

* ``rcm`` restates tools/renumber_mesh.py:115-239 (_build_adjacency, _degree,
  _reverse_cuthill_mckee): coalesced symmetric adjacency, neighbour lists in
  ascending column order stably sorted by degree, start = first unvisited
  index of minimum degree, FIFO BFS, reversed; perm[new] = old.
* ``geometry`` restates mesh/mesh_geometry.py:35-395 with the reference's
  per-face / per-cell loops (small meshes only, any polygon).
* ``geometry_vec`` is the same computation vectorised over faces and cells
  for meshes whose faces all have n_v vertices and whose cells all have the
  same number of faces, in the reference's operation order: sums from 0 in
  vertex / list order, cells face by face and edge by edge, the tet centroid
  ((ref + fc) + v_j + v_{j+1}) / 4 weighted by |t| (mesh_geometry.py:129-136,
  213-247), the distinct vertices summed in ascending vertex order.  The device
  geometry is compared against it; both loop and vectorised forms are pinned to
  the reference's own output (golden) at 1e-13.
"""

from __future__ import annotations

from collections import deque

import numpy as np


def rcm(n, owner, neighbour) -> np.ndarray:
    if n == 0:
        return np.zeros(0, np.int64)
    if len(owner) == 0:
        return np.arange(n, dtype=np.int64)
    pairs = set()
    for a, b in zip(np.asarray(owner).tolist(), np.asarray(neighbour).tolist()):
        pairs.add((a, b))
        pairs.add((b, a))
    rows = sorted(pairs)
    deg = np.zeros(n, np.int64)
    nbr = [[] for _ in range(n)]
    for r, c in rows:
        deg[r] += 1
        if r != c:
            nbr[r].append(c)
    for i in range(n):
        nbr[i].sort(key=lambda v: deg[v])
    visited = np.zeros(n, bool)
    order = []
    while len(order) < n:
        cand = np.nonzero(~visited)[0]
        start = int(cand[np.argmin(deg[cand])])
        q = deque([start])
        visited[start] = True
        while q:
            v = q.popleft()
            order.append(v)
            for w in nbr[v]:
                if not visited[w]:
                    visited[w] = True
                    q.append(w)
    order.reverse()
    return np.array(order, dtype=np.int64)


def geometry(points, faces, owner, neighbour, n_cells):
    """Returns dict(face_centres, face_areas, cell_volumes, cell_centres, delta)."""
    pts = np.asarray(points, dtype=np.float64)
    nf = len(faces)
    nfi = len(neighbour)
    fc = np.zeros((nf, 3))
    area = np.zeros((nf, 3))
    for i, f in enumerate(faces):
        v = pts[np.asarray(f)]
        fc[i] = v.mean(axis=0)
        acc = np.zeros(3)
        for j in range(len(f)):
            acc += np.cross(v[j], v[(j + 1) % len(f)])
        area[i] = acc * 0.5
    cell_faces = [[] for _ in range(n_cells)]
    for f in range(nf):
        cell_faces[int(owner[f])].append(f)
    for f in range(nfi):
        cell_faces[int(neighbour[f])].append(f)
    vol = np.zeros(n_cells)
    cen = np.zeros((n_cells, 3))
    for c in range(n_cells):
        allp = []
        for f in cell_faces[c]:
            allp.extend(np.asarray(faces[f]).tolist())
        ref = pts[sorted(set(allp))].mean(axis=0)
        tv = 0.0
        wc = np.zeros(3)
        for f in cell_faces[c]:
            v = pts[np.asarray(faces[f])]
            if f < nfi and int(neighbour[f]) == c:
                v = v[::-1]
            for j in range(len(v)):
                a, b = v[j], v[(j + 1) % len(v)]
                t = np.dot(fc[f] - ref, np.cross(a - ref, b - ref)) / 6.0
                if abs(t) < 1e-30:
                    continue
                t = abs(t)
                tv += t
                wc += t * (ref + fc[f] + a + b) / 4.0
        if abs(tv) > 1e-30:
            vol[c] = abs(tv)
            cen[c] = wc / tv
        else:
            cen[c] = fc[cell_faces[c]].mean(axis=0)
    delta = np.zeros(nf)
    for f in range(nfi):
        d = cen[int(neighbour[f])] - cen[int(owner[f])]
        m = np.linalg.norm(area[f])
        nh = area[f] / (m if m > 1e-30 else 1.0)
        dn = abs(np.dot(d, nh))
        delta[f] = 1.0 / (dn if dn > 1e-30 else 1.0)
    return {"face_centres": fc, "face_areas": area, "cell_volumes": vol, "cell_centres": cen,
            "delta": delta}


def _cross(a, b):
    return np.stack([a[..., 1] * b[..., 2] - a[..., 2] * b[..., 1],
                     a[..., 2] * b[..., 0] - a[..., 0] * b[..., 2],
                     a[..., 0] * b[..., 1] - a[..., 1] * b[..., 0]], axis=-1)


def _dot(a, b):
    return (a[..., 0] * b[..., 0] + a[..., 1] * b[..., 1]) + a[..., 2] * b[..., 2]


def geometry_vec(points, faces, owner, neighbour, n_cells):
    """Vectorised ``geometry``; faces is an (n_faces, n_v) array.  Returns the
    same dict plus ``mag_sf`` = |S_f| in the order of torch's norm (FMA)."""
    pts = np.asarray(points, dtype=np.float64)
    faces = np.asarray(faces, dtype=np.int64)
    owner = np.asarray(owner, dtype=np.int64)
    neighbour = np.asarray(neighbour, dtype=np.int64)
    nf, nv = faces.shape
    nfi = neighbour.shape[0]
    fv = pts[faces]                                   # (nf, nv, 3)
    fc = np.zeros((nf, 3))
    for j in range(nv):
        fc = fc + fv[:, j]
    fc = fc / float(nv)
    area = np.zeros((nf, 3))
    if nv >= 3:
        for j in range(nv):
            area = area + _cross(fv[:, j], fv[:, (j + 1) % nv])
    area = area * 0.5
    from .ldu import face_mag
    mag = face_mag(area)  # the reference's torch norm: sqrt(fma(z,z,fma(y,y,x*x)))

    fidx = np.concatenate([np.arange(nf), np.arange(nfi)])
    cell = np.concatenate([owner, neighbour])
    flip = np.concatenate([np.zeros(nf, bool), np.ones(nfi, bool)])
    order = np.argsort(cell, kind="stable")
    counts = np.bincount(cell, minlength=n_cells)
    if counts.min() != counts.max():
        raise ValueError("geometry_vec expects cells with equal face counts")
    nfc = int(counts[0])
    fidx = fidx[order].reshape(n_cells, nfc)
    flip = flip[order].reshape(n_cells, nfc)

    cverts = np.sort(faces[fidx].reshape(n_cells, nfc * nv), axis=1)
    first = np.ones_like(cverts, dtype=bool)
    first[:, 1:] = cverts[:, 1:] != cverts[:, :-1]
    ref = np.zeros((n_cells, 3))
    for k in range(nfc * nv):
        ref = np.where(first[:, k, None], ref + pts[cverts[:, k]], ref)
    ref = ref / first.sum(axis=1)[:, None].astype(np.float64)

    tv = np.zeros(n_cells)
    wc = np.zeros((n_cells, 3))
    loops = faces[fidx]                                # (n_cells, nfc, nv)
    loops = np.where(flip[..., None], loops[..., ::-1], loops)
    if nv >= 3:
        for k in range(nfc):
            fck = fc[fidx[:, k]]
            d = fck - ref
            rf = ref + fck
            for j in range(nv):
                va = pts[loops[:, k, j]]
                vb = pts[loops[:, k, (j + 1) % nv]]
                t = _dot(d, _cross(va - ref, vb - ref)) / 6.0
                keep = ~(np.abs(t) < 1e-30)
                t = np.abs(t)
                cen = ((rf + va) + vb) / 4.0
                tv = np.where(keep, tv + t, tv)
                wc = np.where(keep[:, None], wc + cen * t[:, None], wc)
    good = np.abs(tv) > 1e-30
    vol = np.where(good, np.abs(tv), 0.0)
    mean_fc = np.zeros((n_cells, 3))
    for k in range(nfc):
        mean_fc = mean_fc + fc[fidx[:, k]]
    mean_fc = mean_fc / float(nfc)
    cen = np.where(good[:, None], wc / np.where(good, tv, 1.0)[:, None], mean_fc)

    delta = np.zeros(nf)
    if nfi:
        d = cen[neighbour] - cen[owner[:nfi]]
        m = mag[:nfi]
        nh = area[:nfi] / np.where(m > 1e-30, m, 1.0)[:, None]
        dn = np.abs(_dot(d, nh))
        delta[:nfi] = 1.0 / np.where(dn > 1e-30, dn, 1.0)
    return {"face_centres": fc, "face_areas": area, "mag_sf": mag, "cell_volumes": vol,
            "cell_centres": cen, "delta": delta}


def dirichlet_delta(face_centres, face_areas, cell_centres, owner, face_ids):
    """src/applications/laplacian_foam.py:341-353: 1 / clamp(|(x_f - C_P).n_hat|, 1e-30)
    with n_hat = safe / |safe|, safe = S_f where |S_f| > 1e-30 and (1, 1, 1)
    otherwise; both norms in torch's FMA order, the dot summed x, y, z in turn."""
    from .ldu import face_mag
    f = np.asarray(face_ids, dtype=np.int64)
    d = face_centres[f] - cell_centres[np.asarray(owner)[f]]
    s = np.asarray(face_areas, dtype=np.float64)[f]
    safe = np.where((face_mag(s) > 1e-30)[:, None], s, 1.0)
    nh = safe / face_mag(safe)[:, None]
    dn = np.abs((d[:, 0] * nh[:, 0] + d[:, 1] * nh[:, 1]) + d[:, 2] * nh[:, 2])
    return 1.0 / np.maximum(dn, 1e-30)
