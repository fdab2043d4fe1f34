"""OpenFOAM polyMesh files (reference ``io/mesh_io.py``).

Same names, arguments and results as the reference: ``read_points`` /
``write_points`` (:174-252), ``read_faces`` / ``write_faces`` (:277-349),
``read_owner`` / ``read_neighbour`` / ``write_owner`` / ``write_neighbour``
(:369-487), ``read_boundary`` / ``write_boundary`` (:490-583),
``read_mesh`` (:586-614), ``BoundaryPatch`` and ``MeshData`` (:56-145).

The count is the first line of the body that starts with digits; the list
follows its first ``(``.  ASCII lists are parsed by the native
``pfm_parse_ascii_*`` (one pass over the bytes) with the reference's
semantics; binary blocks follow the reference's layout (big-endian doubles
and int32 labels; faces as n, n + 1 offsets, then the labels, in one
parenthesised block, binary_io.py:190-216), and files whose header names an
``arch`` with ``LSB`` are read little-endian and with ``label=64`` as int64
labels, as OpenFOAM writes them; a ``faceCompactList`` (OpenFOAM's binary
faces: an offsets list, then a labels list) is read as well.

``read_faces_csr`` returns the faces as (face_ptr, face_verts) int32 arrays,
the layout the device geometry takes, without building one array per face;
``MeshData`` keeps that layout and builds ``faces`` on first use.
"""

from __future__ import annotations

import ctypes as C
import re
from pathlib import Path
from typing import Union

import numpy as np
import torch

from .. import _lib
from .foam_file import FoamFileHeader, read_foam_file, write_foam_file

__all__ = ["BoundaryPatch", "MeshData", "read_points", "write_points", "read_faces",
           "read_faces_csr", "write_faces", "read_owner", "read_neighbour", "write_owner",
           "write_neighbour", "read_boundary", "write_boundary", "read_mesh"]

_COUNT = re.compile(r"^\s*(\d+)", re.MULTILINE)


class BoundaryPatch:
    def __init__(self, name: str, patch_type: str = "patch", n_faces: int = 0,
                 start_face: int = 0, in_groups: list[str] | None = None) -> None:
        self.name = name
        self.patch_type = patch_type
        self.n_faces = n_faces
        self.start_face = start_face
        self.in_groups = in_groups or []

    def __repr__(self) -> str:
        return (f"BoundaryPatch(name={self.name!r}, type={self.patch_type!r}, "
                f"nFaces={self.n_faces}, startFace={self.start_face})")


class MeshData:
    def __init__(self, points: torch.Tensor, faces, owner: torch.Tensor,
                 neighbour: torch.Tensor, boundary: list[BoundaryPatch]) -> None:
        self.points = points
        if isinstance(faces, tuple):  # (face_ptr, face_verts)
            self.face_ptr, self.face_verts = faces
            self._faces = None
        else:
            self._faces = list(faces)
            self.face_ptr, self.face_verts = _csr_of(self._faces)
        self.owner = owner
        self.neighbour = neighbour
        self.boundary = boundary

    @property
    def faces(self) -> list[np.ndarray]:
        if self._faces is None:
            self._faces = np.split(self.face_verts, self.face_ptr[1:-1])
        return self._faces

    @property
    def n_points(self) -> int:
        return self.points.shape[0]

    @property
    def n_faces(self) -> int:
        return int(self.face_ptr.shape[0]) - 1

    @property
    def n_internal_faces(self) -> int:
        return self.neighbour.shape[0]

    @property
    def n_cells(self) -> int:
        return int(self.owner.max().item()) + 1 if self.owner.numel() > 0 else 0

    @property
    def n_boundary_faces(self) -> int:
        return self.n_faces - self.n_internal_faces

    def __repr__(self) -> str:
        return (f"MeshData(n_points={self.n_points}, n_faces={self.n_faces}, "
                f"n_cells={self.n_cells}, n_patches={len(self.boundary)})")


def _csr_of(faces) -> tuple[np.ndarray, np.ndarray]:
    lens = np.fromiter((len(f) for f in faces), dtype=np.int64, count=len(faces))
    ptr = np.zeros(len(faces) + 1, np.int64)
    np.cumsum(lens, out=ptr[1:])
    verts = (np.concatenate([np.asarray(f, dtype=np.int32) for f in faces]) if len(faces)
             else np.zeros(0, np.int32))
    return ptr.astype(np.int32), verts.astype(np.int32)


# ------------------------------------------------------------------ helpers --
def _count_and_start(body: str, what: str) -> tuple[int, int]:
    m = _COUNT.search(body)
    if m is None:
        raise ValueError(f"Cannot find {what} count")
    p = body.find("(", m.end())
    if p < 0:
        raise ValueError(f"Cannot find '(' for {what} data")
    return int(m.group(1)), p + 1


def _arch(header: FoamFileHeader) -> tuple[str, str]:
    """(byte order, label dtype) of a binary block."""
    a = header.arch or ""
    order = "<" if "LSB" in a else ">"
    return order, ("i8" if "label=64" in a else "i4")


def _raw(body: str, start: int, nbytes: int, what: str) -> bytes:
    data = body[start:start + nbytes].encode("latin-1")
    if len(data) < nbytes:
        raise EOFError(f"Expected {nbytes} bytes of {what}, got {len(data)}")
    return data


def _ascii_bytes(body: str, start: int) -> bytes:
    return body[start:].encode("utf-8")


def _check(status: int, what: str) -> None:
    _lib.check(status, what)


# ------------------------------------------------------------------- points --
def read_points(path: Union[str, Path]) -> tuple[FoamFileHeader, torch.Tensor]:
    header, body = read_foam_file(path)
    n, start = _count_and_start(body, "point")
    if header.is_binary:
        order, _ = _arch(header)
        arr = np.frombuffer(_raw(body, start, 24 * n, "points"), dtype=order + "f8").astype(
            np.float64).reshape(n, 3)
    else:
        arr = np.empty((n, 3), np.float64)
        txt = _ascii_bytes(body, start)
        got = C.c_int64(0)
        _check(_lib.lib().pfm_parse_ascii_points(txt, len(txt), n, arr.ctypes.data, C.byref(got)),
               "points")
    return header, torch.from_numpy(arr)


def write_points(path: Union[str, Path], header: FoamFileHeader, points: torch.Tensor, *,
                 overwrite: bool = False) -> None:
    pts = torch.as_tensor(points).detach().cpu().numpy().astype(np.float64)
    n = pts.shape[0]
    if header.is_binary:
        order, _ = _arch(header)
        blob = pts.astype(order + "f8").tobytes().decode("latin-1")
        body = f"{n}\n({blob})"
    else:
        lines = [f"{n}", "("]
        lines += [f"({x:.10g} {y:.10g} {z:.10g})" for x, y, z in pts.tolist()]
        lines.append(")")
        body = "\n".join(lines)
    write_foam_file(path, header, body, overwrite=overwrite)


# -------------------------------------------------------------------- faces --
def read_faces_csr(path: Union[str, Path]) -> tuple[FoamFileHeader, np.ndarray, np.ndarray]:
    header, body = read_foam_file(path)
    n, start = _count_and_start(body, "face")
    if header.is_binary:
        order, lab = _arch(header)
        size = int(lab[1])
        if header.class_name == "faceCompactList":
            # OpenFOAM: n offsets "( ... )", then m labels "( ... )"
            offs = np.frombuffer(_raw(body, start, size * n, "face offsets"),
                                 dtype=order + lab).astype(np.int64)
            rest = body[start + size * n:]
            m, s2 = _count_and_start(rest, "face label")
            verts = np.frombuffer(_raw(rest, s2, size * m, "face labels"),
                                  dtype=order + lab).astype(np.int32)
            ptr = offs.astype(np.int32)
        else:
            hdr = np.frombuffer(_raw(body, start, size, "faces"), dtype=order + lab)
            nl = int(hdr[0])
            offs = np.frombuffer(_raw(body, start + size, size * (nl + 1), "face offsets"),
                                 dtype=order + lab).astype(np.int64)
            total = int(offs[-1])
            verts = np.frombuffer(_raw(body, start + size * (nl + 2), size * total,
                                       "face labels"), dtype=order + lab).astype(np.int32)
            ptr = offs.astype(np.int32)
        return header, ptr, verts
    txt = _ascii_bytes(body, start)
    ptr = np.zeros(n + 1, np.int32)
    cap = max(1, len(txt) // 2)  # every label takes at least a digit and a separator
    verts = np.empty(cap, np.int32)
    got, nv = C.c_int64(0), C.c_int64(0)
    _check(_lib.lib().pfm_parse_ascii_faces(txt, len(txt), n, cap, ptr.ctypes.data,
                                            verts.ctypes.data, C.byref(got), C.byref(nv)),
           "faces")
    return header, ptr, verts[: nv.value].copy()


def read_faces(path: Union[str, Path]) -> tuple[FoamFileHeader, list[np.ndarray]]:
    header, ptr, verts = read_faces_csr(path)
    return header, np.split(verts, ptr[1:-1]) if len(ptr) > 1 else []


def write_faces(path: Union[str, Path], header: FoamFileHeader, faces: list[np.ndarray], *,
                overwrite: bool = False) -> None:
    n = len(faces)
    if header.is_binary:
        order, lab = _arch(header)
        ptr, verts = _csr_of(faces)
        blob = (np.array([n], order + lab).tobytes() + ptr.astype(order + lab).tobytes()
                + verts.astype(order + lab).tobytes())
        body = f"{n}\n({blob.decode('latin-1')})"
    else:
        lines = [f"{n}", "("]
        lines += [f"{len(f)}({' '.join(str(int(v)) for v in f)})" for f in faces]
        lines.append(")")
        body = "\n".join(lines)
    write_foam_file(path, header, body, overwrite=overwrite)


# ------------------------------------------------------------------- labels --
def _read_labels(path) -> tuple[FoamFileHeader, torch.Tensor]:
    header, body = read_foam_file(path)
    n, start = _count_and_start(body, "label")
    if header.is_binary:
        order, lab = _arch(header)
        arr = np.frombuffer(_raw(body, start, int(lab[1]) * n, "labels"),
                            dtype=order + lab).astype(np.int64)
    else:
        arr = np.empty(n, np.int64)
        txt = _ascii_bytes(body, start)
        got = C.c_int64(0)
        _check(_lib.lib().pfm_parse_ascii_labels(txt, len(txt), n, arr.ctypes.data,
                                                 C.byref(got)), "labels")
    return header, torch.from_numpy(arr)


def read_owner(path: Union[str, Path]) -> tuple[FoamFileHeader, torch.Tensor]:
    return _read_labels(path)


def read_neighbour(path: Union[str, Path]) -> tuple[FoamFileHeader, torch.Tensor]:
    return _read_labels(path)


def _write_labels(path, header, labels, overwrite) -> None:
    arr = torch.as_tensor(labels).detach().cpu().numpy().astype(np.int64)
    n = arr.shape[0]
    if header.is_binary:
        order, lab = _arch(header)
        body = f"{n}\n({arr.astype(order + lab).tobytes().decode('latin-1')})"
    else:
        body = "\n".join([f"{n}", "("] + [str(v) for v in arr.tolist()] + [")"])
    write_foam_file(path, header, body, overwrite=overwrite)


def write_owner(path: Union[str, Path], header: FoamFileHeader, owner: torch.Tensor, *,
                overwrite: bool = False) -> None:
    _write_labels(path, header, owner, overwrite)


def write_neighbour(path: Union[str, Path], header: FoamFileHeader, neighbour: torch.Tensor, *,
                    overwrite: bool = False) -> None:
    _write_labels(path, header, neighbour, overwrite)


# ----------------------------------------------------------------- boundary --
_BLOCK = re.compile(r"(\w+)\s*\{([^}]*)\}", re.DOTALL)
_KV = re.compile(r"(\w+)\s+(.+?)\s*;")


def read_boundary(path: Union[str, Path]) -> tuple[FoamFileHeader, list[BoundaryPatch]]:
    header, body = read_foam_file(path)
    patches = []
    for bm in _BLOCK.finditer(body):
        p = BoundaryPatch(bm.group(1))
        for kv in _KV.finditer(bm.group(2)):
            key, val = kv.group(1), kv.group(2).strip().strip('"')
            if key == "type":
                p.patch_type = val
            elif key == "nFaces":
                p.n_faces = int(val)
            elif key == "startFace":
                p.start_face = int(val)
            elif key == "inGroups":
                val = val.strip("()")
                p.in_groups = [g.strip().strip('"') for g in val.split()] if val else []
        patches.append(p)
    return header, patches


def write_boundary(path: Union[str, Path], header: FoamFileHeader,
                   patches: list[BoundaryPatch], *, overwrite: bool = False) -> None:
    lines = [f"{len(patches)}", "("]
    for p in patches:
        lines += [f"    {p.name}", "    {", f"        type            {p.patch_type};",
                  f"        nFaces          {p.n_faces};",
                  f"        startFace       {p.start_face};"]
        if p.in_groups:
            groups = " ".join(f'"{g}"' for g in p.in_groups)
            lines.append(f"        inGroups        ({groups});")
        lines.append("    }")
    lines.append(")")
    write_foam_file(path, header, "\n".join(lines), overwrite=overwrite)


# --------------------------------------------------------------------- mesh --
def read_mesh(mesh_dir: Union[str, Path]) -> MeshData:
    d = Path(mesh_dir)
    _, points = read_points(d / "points")
    _, ptr, verts = read_faces_csr(d / "faces")
    _, owner = read_owner(d / "owner")
    _, neighbour = read_neighbour(d / "neighbour")
    _, boundary = read_boundary(d / "boundary")
    return MeshData(points=points, faces=(ptr, verts), owner=owner, neighbour=neighbour,
                    boundary=boundary)
