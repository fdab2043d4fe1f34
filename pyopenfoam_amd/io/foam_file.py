"""OpenFOAM file header and body (reference ``io/foam_file.py``).

``FoamFileHeader`` (:50-100), ``parse_header`` (:140-175: comments stripped,
``key value;`` pairs of the first ``FoamFile { }`` block, quotes removed),
``split_header_body`` (:180-225), ``read_foam_file`` (:245-306: the header
ends at the brace matching the one after ``FoamFile``; a header containing
``binary`` makes the body latin-1 text, byte for byte; leading newlines,
carriage returns and spaces are stripped from the body) and
``write_foam_file`` (:309-331).

Difference: ``write_foam_file`` writes the body's characters as single bytes
(latin-1) when the header is binary.  The reference writes every body as
UTF-8 text, which turns each byte >= 0x80 of a binary block into two bytes,
so the binary files it writes do not read back; ASCII files are identical.
"""

from __future__ import annotations

import re
from dataclasses import dataclass
from enum import Enum
from pathlib import Path
from typing import Union

__all__ = ["FileFormat", "FoamFileHeader", "parse_header", "split_header_body", "read_foam_file",
           "write_foam_file"]


class FileFormat(Enum):
    ASCII = "ascii"
    BINARY = "binary"


@dataclass
class FoamFileHeader:
    version: str = "2.0"
    format: FileFormat = FileFormat.ASCII
    class_name: str = ""
    location: str = ""
    object: str = ""
    note: str = ""
    arch: str = ""

    @property
    def is_binary(self) -> bool:
        return self.format == FileFormat.BINARY

    @property
    def is_ascii(self) -> bool:
        return self.format == FileFormat.ASCII

    def to_header_string(self) -> str:
        out = ["FoamFile", "{", f"    version     {self.version};",
               f"    format      {self.format.value};"]
        if self.class_name:
            out.append(f"    class       {self.class_name};")
        if self.location:
            out.append(f'    location    "{self.location}";')
        if self.object:
            out.append(f"    object      {self.object};")
        if self.note:
            out.append(f'    note        "{self.note}";')
        if self.arch:
            out.append(f'    arch        "{self.arch}";')
        out.append("}")
        return "\n".join(out)

    def __repr__(self) -> str:
        return (f"FoamFileHeader(version={self.version!r}, format={self.format.value}, "
                f"class={self.class_name!r}, object={self.object!r})")


_BLOCK = re.compile(r"FoamFile\s*\{([^}]*)\}", re.DOTALL)
_KV = re.compile(r"(\w+)\s+(.+?)\s*;", re.DOTALL)
_FIELDS = {"version": "version", "class": "class_name", "location": "location",
           "object": "object", "note": "note", "arch": "arch"}


def _strip_comments(text: str) -> str:
    text = re.sub(r"//[^\n]*", "", text)
    return re.sub(r"/\*.*?\*/", "", text, flags=re.DOTALL)


def parse_header(content: str) -> FoamFileHeader:
    m = _BLOCK.search(_strip_comments(content))
    if m is None:
        raise ValueError("No FoamFile header found in content")
    h = FoamFileHeader()
    for kv in _KV.finditer(m.group(1)):
        key = kv.group(1).strip()
        val = kv.group(2).strip()
        if len(val) >= 2 and val[0] == '"' and val[-1] == '"':
            val = val[1:-1]
        if key == "format":
            h.format = FileFormat(val.lower())
        elif key in _FIELDS:
            setattr(h, _FIELDS[key], val)
    return h


def _header_end(data, start: int) -> int:
    """Index of the brace closing the block opened at ``start``."""
    depth = 0
    for i in range(start, len(data)):
        c = data[i:i + 1]
        if c in (b"{", "{"):
            depth += 1
        elif c in (b"}", "}"):
            depth -= 1
            if depth == 0:
                return i
    raise ValueError("Unmatched braces in FoamFile header")


def split_header_body(content: str) -> tuple[FoamFileHeader, str]:
    m = re.search(r"\bFoamFile\b", content)
    if m is None:
        raise ValueError("No FoamFile header found in content")
    brace = content.find("{", m.end())
    if brace < 0:
        raise ValueError("No opening brace found after FoamFile")
    depth, in_str, end = 0, False, None
    for i in range(brace, len(content)):
        ch = content[i]
        if ch == '"' and (i == 0 or content[i - 1] != "\\"):
            in_str = not in_str
        elif not in_str:
            if ch == "{":
                depth += 1
            elif ch == "}":
                depth -= 1
                if depth == 0:
                    end = i
                    break
    if end is None:
        raise ValueError("Unmatched braces in FoamFile header")
    return parse_header(content), content[end + 1:].lstrip("\n\r ")


def read_foam_file(source: Union[str, Path]) -> tuple[FoamFileHeader, str]:
    path = Path(source)
    if not path.exists():
        raise FileNotFoundError(f"File not found: {path}")
    raw = path.read_bytes()
    k = raw.find(b"FoamFile")
    if k < 0:
        raise ValueError("No FoamFile header found in content")
    brace = raw.find(b"{", k)
    if brace < 0:
        raise ValueError("No opening brace found after FoamFile")
    end = _header_end(raw, brace)
    head = raw[:end + 1]
    if b"binary" in head:
        body = raw[end + 1:].decode("latin-1")
    else:
        body = raw[end + 1:].decode("utf-8", errors="replace")
    return parse_header(head.decode("ascii", errors="replace")), body.lstrip("\n\r ")


def write_foam_file(dest: Union[str, Path], header: FoamFileHeader, body: str, *,
                    overwrite: bool = False) -> None:
    path = Path(dest)
    if path.exists() and not overwrite:
        raise FileExistsError(f"File already exists: {path}")
    content = f"{header.to_header_string()}\n\n{body}"
    path.parent.mkdir(parents=True, exist_ok=True)
    path.write_bytes(content.encode("latin-1" if header.is_binary else "utf-8"))
