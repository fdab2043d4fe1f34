"""polyMesh file reading (oracle; test infrastructure only).

Restates the reference's readers with its own regular expressions, for small
files:

* ``split``      -- io/foam_file.py:245-306 (header block by brace matching,
                    ``binary`` in the header -> latin-1 body, leading
                    whitespace stripped)
* ``points``     -- io/mesh_io.py:154-213 (_parse_ascii_points / read_points)
* ``faces``      -- io/mesh_io.py:255-310 and binary_io.py:190-216 (the
                    compact list: n, n + 1 offsets, labels; big-endian int32)
* ``labels``     -- io/mesh_io.py:352-403
* ``boundary``   -- io/mesh_io.py:490-546

Pinned by the reference tests' own fixtures (tests/unit/io/test_mesh_io.py:33-148,
150-268: a one-cell tetrahedron and a two-cell mesh with their asserted
values, tests/golden/polymesh/) and by write/read round trips.
"""

from __future__ import annotations

import re
import struct
from pathlib import Path

import numpy as np

_COUNT = re.compile(r"^\s*(\d+)", re.MULTILINE)


def split(path):
    raw = Path(path).read_bytes()
    k = raw.find(b"FoamFile")
    b = raw.find(b"{", k)
    depth, end = 0, None
    for i in range(b, len(raw)):
        if raw[i:i + 1] == b"{":
            depth += 1
        elif raw[i:i + 1] == b"}":
            depth -= 1
            if depth == 0:
                end = i
                break
    head = raw[:end + 1]
    binary = b"binary" in head
    body = raw[end + 1:].decode("latin-1" if binary else "utf-8")
    return binary, body.lstrip("\n\r ")


def _matching(text, start):
    depth = 0
    for i in range(start, len(text)):
        if text[i] == "(":
            depth += 1
        elif text[i] == ")":
            depth -= 1
            if depth == 0:
                return i
    raise ValueError("Unmatched parenthesis")


def _list_text(body):
    m = _COUNT.search(body)
    n = int(m.group(1))
    p = body.find("(", m.end())
    return n, p, body[p:_matching(body, p) + 1] if p >= 0 else ""


def points(path):
    binary, body = split(path)
    if binary:
        m = _COUNT.search(body)
        n = int(m.group(1))
        p = body.find("(", m.end()) + 1
        raw = body[p:p + 24 * n].encode("latin-1")
        return np.array(struct.unpack(f">{3 * n}d", raw)).reshape(n, 3)
    n, _, txt = _list_text(body)
    txt = txt.strip().strip("()")
    pts = [[float(c) for c in mm.group(1).split()] for mm in re.finditer(r"\(\s*([^)]+)\)", txt)]
    if len(pts) != n:
        raise ValueError(f"Expected {n} points, got {len(pts)}")
    return np.array(pts, dtype=np.float64).reshape(n, 3)


def faces(path):
    binary, body = split(path)
    if binary:
        m = _COUNT.search(body)
        p = body.find("(", m.end()) + 1
        raw = body[p:].encode("latin-1")
        nl = struct.unpack(">i", raw[:4])[0]
        offs = struct.unpack(f">{nl + 1}i", raw[4:4 + 4 * (nl + 1)])
        data = struct.unpack(f">{offs[-1]}i", raw[4 + 4 * (nl + 1):4 + 4 * (nl + 1) + 4 * offs[-1]])
        return [np.array(data[offs[i]:offs[i + 1]], np.int32) for i in range(nl)]
    n, _, txt = _list_text(body)
    txt = txt.strip().strip("()")
    out = [np.array([int(v) for v in mm.group(1).split()], np.int32)
           for mm in re.finditer(r"\d+\(([^)]*)\)", txt)]
    if len(out) != n:
        raise ValueError(f"Expected {n} faces, got {len(out)}")
    return out


def labels(path):
    binary, body = split(path)
    if binary:
        m = _COUNT.search(body)
        n = int(m.group(1))
        p = body.find("(", m.end()) + 1
        return np.array(struct.unpack(f">{n}i", body[p:p + 4 * n].encode("latin-1")), np.int64)
    n, _, txt = _list_text(body)
    vals = [int(v) for v in txt.strip().strip("()").split() if v.strip()]
    if len(vals) != n:
        raise ValueError(f"Expected {n} labels, got {len(vals)}")
    return np.array(vals, np.int64)


def boundary(path):
    _, body = split(path)
    out = []
    for bm in re.finditer(r"(\w+)\s*\{([^}]*)\}", body, re.DOTALL):
        rec = {"name": bm.group(1), "type": "patch", "nFaces": 0, "startFace": 0}
        for kv in re.finditer(r"(\w+)\s+(.+?)\s*;", bm.group(2)):
            key, val = kv.group(1), kv.group(2).strip().strip('"')
            if key == "type":
                rec["type"] = val
            elif key in ("nFaces", "startFace"):
                rec[key] = int(val)
        out.append(rec)
    return out
