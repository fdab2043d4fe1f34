"""OpenFOAM dictionaries: fvSolution, fvSchemes, controlDict (reference ``io/dictionary.py``).

``parse_dict(text)`` / ``parse_dict_file(path)`` -> ``FoamDict``, with the
reference's grammar and results (:126-606):

* tokens: ``"strings"`` (escapes \\n \\t \\\\ \\"), ``{ } ( ) [ ] ; , $ #``,
  numbers (optional sign, one ``.``, one exponent with optional sign), and
  words: a run of letters, digits and ``_-.:+*^/\\<>()',`` -- so
  ``div(phi,U)`` is one word; ``//`` and ``/* */`` comments are skipped;
* entries: ``key value;``, ``key { ... }`` (optional ``;``), multi-token
  values joined by single spaces (``value uniform 0;`` -> ``"uniform 0"``);
  a standalone ``$name;`` entry and ``#include``/``#calc``/``#if...``
  directives are consumed and otherwise ignored, as the reference does;
* values: numbers as ``int`` when ``int()`` accepts them, else ``float``;
  ``( ... )`` -> ``FoamList``; ``[ ... ]`` -> the string ``"[a b c]"``;
  ``$name`` -> the string ``"$name"``; ``{ ... }`` -> ``FoamDict``.

``FoamDict`` adds ``d["a/b"]`` path lookup, ``get_path`` and
``resolve_macro`` (``$var`` through the parent chain, ``$sub/var`` below a
sub-dictionary); ``expand_macros`` replaces ``$...`` strings by their values
(:643-691).  ``solver_dicts(fv_solution)`` lists the ``solvers`` entries the
linear-solver factory (``solver_from_dict``) takes.
"""

from __future__ import annotations

from pathlib import Path
from typing import Any, Optional, Union

__all__ = ["FoamDict", "FoamList", "parse_dict", "parse_dict_file", "expand_macros",
           "solver_dicts"]

_WORD_EXTRA = set("_-.:+*^/\\<>()',")
_PUNCT = {"{": "LBRACE", "}": "RBRACE", "(": "LPAREN", ")": "RPAREN", "[": "LBRACKET",
          "]": "RBRACKET", ";": "SEMI", ",": "COMMA", "$": "DOLLAR", "#": "HASH"}


class FoamDict(dict):
    def __init__(self, *args: Any, parent: Optional["FoamDict"] = None, **kwargs: Any) -> None:
        super().__init__(*args, **kwargs)
        self._parent = parent

    @property
    def parent(self) -> Optional["FoamDict"]:
        return self._parent

    def set_parent(self, parent: "FoamDict") -> None:
        self._parent = parent

    def __missing__(self, key: str) -> Any:
        if isinstance(key, str) and "/" in key:
            head, rest = key.split("/", 1)
            sub = self[head]
            if isinstance(sub, FoamDict):
                return sub[rest]
        raise KeyError(key)

    def get_path(self, path: str, default: Any = None) -> Any:
        try:
            return self[path]
        except KeyError:
            return default

    def resolve_macro(self, name: str) -> Any:
        if "/" in name:
            head, rest = name.split("/", 1)
            sub = self.get(head)
            if isinstance(sub, FoamDict):
                return sub.resolve_macro(rest)
            raise KeyError(f"Macro ${name}: subdict '{head}' not found")
        if name in self:
            return self[name]
        if self._parent is not None:
            return self._parent.resolve_macro(name)
        raise KeyError(f"Macro ${name} not found")


class FoamList(list):
    def __init__(self, *args: Any, foam_type: str = "", **kwargs: Any) -> None:
        super().__init__(*args)
        self.foam_type = foam_type


# ---------------------------------------------------------------- tokens ----
def _tokens(text: str) -> list[tuple[str, str]]:
    out: list[tuple[str, str]] = []
    i, n = 0, len(text)
    while True:
        while i < n:  # whitespace and comments
            ch = text[i]
            if ch.isspace():
                i += 1
            elif ch == "/" and i + 1 < n and text[i + 1] == "/":
                j = text.find("\n", i + 2)
                i = n if j < 0 else j + 1
            elif ch == "/" and i + 1 < n and text[i + 1] == "*":
                j = text.find("*/", i + 2)
                i = n if j < 0 else j + 2
            else:
                break
        if i >= n:
            out.append(("EOF", ""))
            return out
        ch = text[i]
        if ch == '"':
            i += 1
            buf = []
            while i < n:
                c = text[i]
                i += 1
                if c == "\\":
                    if i < n:
                        e = text[i]
                        i += 1
                        buf.append({"n": "\n", "t": "\t"}.get(e, e))
                elif c == '"':
                    break
                else:
                    buf.append(c)
            out.append(("STRING", "".join(buf)))
        elif ch in _PUNCT:
            out.append((_PUNCT[ch], ch))
            i += 1
        elif ch.isdigit() or (ch in "+-" and i + 1 < n and (text[i + 1].isdigit()
                                                            or text[i + 1] == ".")):
            j = i + 1 if ch in "+-" else i
            dot = exp = False
            while j < n:
                c = text[j]
                if c.isdigit():
                    j += 1
                elif c == "." and not dot:
                    dot = True
                    j += 1
                elif c in "eE" and not exp:
                    exp = True
                    j += 1
                    if j < n and text[j] in "+-":
                        j += 1
                else:
                    break
            out.append(("NUMBER", text[i:j]))
            i = j
        else:
            j = i
            while j < n and (text[j].isalnum() or text[j] in _WORD_EXTRA):
                j += 1
            if j == i:
                j = i + 1  # an unrecognised character is a word of its own
            out.append(("WORD", text[i:j]))
            i = j


# ---------------------------------------------------------------- parser ----
class _Parser:
    def __init__(self, toks: list[tuple[str, str]]) -> None:
        self.t = toks
        self.i = 0

    def peek(self) -> tuple[str, str]:
        return self.t[self.i] if self.i < len(self.t) else ("EOF", "")

    def take(self) -> tuple[str, str]:
        tok = self.peek()
        self.i += 1
        return tok

    def at(self, kind: str) -> bool:
        return self.peek()[0] == kind

    def expect(self, kind: str) -> tuple[str, str]:
        tok = self.take()
        if tok[0] != kind:
            raise ValueError(f"Expected token type {kind}, got {tok[0]} ({tok[1]!r}) "
                             f"at position {self.i}")
        return tok

    def dict_body(self, parent: Optional[FoamDict]) -> FoamDict:
        d = FoamDict(parent=parent)
        while not self.at("EOF") and not self.at("RBRACE"):
            self.entry(d)
        return d

    def entry(self, d: FoamDict) -> None:
        kind = self.peek()[0]
        if kind in ("SEMI", "COMMA"):
            self.take()
            return
        if kind == "DOLLAR":  # a standalone macro entry is consumed only
            self.take()
            self.take()
            if self.at("SEMI"):
                self.take()
            return
        if kind == "HASH":
            self.take()
            directive = self.take()[1]
            if directive == "include" and self.peek()[0] in ("STRING", "WORD"):
                self.take()
            elif directive == "includeEtc" and self.at("STRING"):
                self.take()
            elif directive == "calc" and not self.at("EOF"):
                self.take()
            elif directive in ("ifeq", "if"):
                while not self.at("EOF") and not self.at("LBRACE"):
                    self.take()
            if self.at("SEMI"):
                self.take()
            return
        kind, key = self.take()
        if kind not in ("WORD", "STRING"):
            raise ValueError(f"Expected key (word or string), got {kind} ({key!r})")
        if self.at("LBRACE"):
            self.take()
            sub = self.dict_body(d)
            self.expect("RBRACE")
            if self.at("SEMI"):
                self.take()
            d[key] = sub
            return
        value = self.value(d)
        if isinstance(value, str) and self.peek()[0] not in ("SEMI", "LBRACE", "EOF", "RBRACE"):
            parts = [value]
            while self.peek()[0] not in ("SEMI", "LBRACE", "EOF", "RBRACE"):
                parts.append(self.take()[1])
            value = " ".join(parts)
        if self.at("SEMI"):
            self.take()
        d[key] = value

    def value(self, d: FoamDict) -> Any:
        kind, text = self.peek()
        if kind == "NUMBER":
            self.take()
            try:
                return int(text)
            except ValueError:
                return float(text)
        if kind in ("STRING", "WORD"):
            self.take()
            return text
        if kind == "LPAREN":
            self.take()
            items = FoamList()
            while not self.at("EOF") and not self.at("RPAREN"):
                items.append(self.value(d))
            self.expect("RPAREN")
            return items
        if kind == "LBRACKET":
            self.take()
            parts = []
            while not self.at("EOF") and not self.at("RBRACKET"):
                parts.append(self.take()[1])
            self.expect("RBRACKET")
            return f"[{' '.join(parts)}]"
        if kind == "DOLLAR":
            self.take()
            return f"${self.take()[1]}"
        if kind == "LBRACE":
            self.take()
            sub = self.dict_body(d.parent)
            self.expect("RBRACE")
            return sub
        raise ValueError(f"Unexpected token for value: {kind} {text!r}")


def parse_dict(text: str, parent: Optional[FoamDict] = None) -> FoamDict:
    """Parse dictionary text; a FoamFile header block, if present, is skipped
    (split_header_body)."""
    from .foam_file import split_header_body
    try:
        _, body = split_header_body(text)
    except ValueError:
        body = text
    return _Parser(_tokens(body)).dict_body(parent)


def parse_dict_file(path: Union[str, Path]) -> FoamDict:
    return parse_dict(Path(path).read_text(encoding="utf-8", errors="replace"))


def expand_macros(d: FoamDict, *, top_level: Optional[FoamDict] = None) -> FoamDict:
    """A new FoamDict with ``$var`` / ``$sub/var`` strings (also in lists and
    sub-dictionaries) replaced by their values through ``resolve_macro`` and
    ``$!var`` by ``top_level[var]``; unresolvable ones are kept as they are."""
    top = d if top_level is None else top_level
    out = FoamDict(parent=d.parent)

    def resolve(ref: str, bang: bool) -> Any:
        name = ref[1:]
        try:
            if bang and name.startswith("!"):
                return top[name[1:]]
            return d.resolve_macro(name)
        except KeyError:
            return ref

    for key, val in d.items():
        if isinstance(val, str) and val.startswith("$"):
            out[key] = resolve(val, True)
        elif isinstance(val, FoamDict):
            out[key] = expand_macros(val, top_level=top)
        elif isinstance(val, FoamList):
            out[key] = FoamList([resolve(x, False) if isinstance(x, str) and x.startswith("$")
                                 else x for x in val], foam_type=val.foam_type)
        else:
            out[key] = val
    return out


def solver_dicts(fv_solution: FoamDict) -> dict[str, FoamDict]:
    """The ``solvers`` sub-dictionary of an fvSolution: field pattern ->
    solver dictionary (macros expanded)."""
    solvers = fv_solution.get("solvers")
    if not isinstance(solvers, FoamDict):
        return {}
    solvers = expand_macros(solvers, top_level=fv_solution)
    return {k: v for k, v in solvers.items() if isinstance(v, FoamDict)}
