"""Parameter files in the reference's ``input.inf`` style.

Restates ``readini`` (src/shared/m_readini.f90:29-166) and
``system__expenv`` (src/shared/m_system.f90:75-102):

* the file is scanned from the top for every key; lines are read into a
  256-character buffer and left-adjusted (leading blanks only);
* lines starting with ``#`` or ``!`` are comments;
* a line matches when it starts with the key (case-sensitive) and the next
  non-blank character is ``=``; the first matching line wins;
* the value is the first list-directed item after ``=``: a quoted string
  ('...' or "...", doubled quotes stand for one), or else the characters up
  to the first blank, comma or slash;
* ``${NAME}`` is replaced by the environment variable (only for values read
  from the file, not for defaults); a ``{`` alone also starts a (nameless)
  substitution, as ``scan(str, "${")`` does in the reference;
* numbers and logicals are then read from that item list-directed
  (``1.5d-3`` and ``.true.``/``T``/``F`` forms included);
* a missing key yields the default, or raises in strict mode.
"""
from __future__ import annotations

import os
import re

import numpy as np

LINE = 256


class IniError(ValueError):
    pass


def _first_item(s: str) -> str:
    """First list-directed character item of ``s``."""
    s = s.lstrip(" ")
    if not s:
        raise IniError("empty value")
    q = s[0]
    if q in "'\"":
        out = []
        i = 1
        while i < len(s):
            if s[i] == q:
                if i + 1 < len(s) and s[i + 1] == q:
                    out.append(q)
                    i += 2
                    continue
                break
            out.append(s[i])
            i += 1
        return "".join(out)
    m = re.match(r"[^ ,/]*", s)
    item = m.group(0)
    if not item:
        raise IniError("null value")
    return item


def expenv(s: str) -> str:
    """``system__expenv``: expand ${NAME} using the environment."""
    out = ""
    p = 0
    while True:
        k1 = min([i for i in (s.find("$", p), s.find("{", p)) if i >= 0], default=-1)
        if k1 < 0:
            break
        k2 = s.find("}", p)
        if k2 < 0:
            raise IniError(f"unterminated ${{...}} in {s!r}")
        out = out.rstrip(" ") + s[p:k1]
        name = s[k1 + 2:k2]
        out = out.rstrip(" ") + (os.environ.get(name, "") if name else "").rstrip(" ")
        p = k2 + 1
    return (out.rstrip(" ") + s[p:].rstrip(" "))


_REAL = re.compile(r"^([+-]?(?:\d+\.?\d*|\.\d+))(?:[eEdDqQ]([+-]?\d+)|([+-]\d+))?$")


def _to_real(tok: str) -> float:
    m = _REAL.match(tok.strip())
    if not m:
        raise IniError(f"cannot read a real from {tok!r}")
    exp = m.group(2) or m.group(3)
    return float(m.group(1) + (("e" + exp) if exp else ""))


def _to_int(tok: str) -> int:
    t = tok.strip()
    if not re.fullmatch(r"[+-]?\d+", t):
        raise IniError(f"cannot read an integer from {tok!r}")
    v = int(t)
    if not -2 ** 31 <= v < 2 ** 31:
        raise IniError(f"integer {t} overflows")
    return v


def _to_logical(tok: str) -> bool:
    t = tok.strip()
    if t.startswith("."):
        t = t[1:]
    if t[:1] in ("T", "t"):
        return True
    if t[:1] in ("F", "f"):
        return False
    raise IniError(f"cannot read a logical from {tok!r}")


def _repeat(tok: str) -> str:
    m = re.fullmatch(r"\d+\*(.+)", tok)
    return m.group(1) if m else tok


class InputFile:
    """An ``input.inf``-style parameter file held as lines.

    ``get(key, default)`` returns a value of the default's type (bool, int,
    float as double precision, str); ``kind='s'`` reads a single-precision
    real (returned as numpy.float32), matching ``readini_s``.
    """

    def __init__(self, path=None, text=None, strict=False):
        if text is None:
            with open(path, "r", errors="replace") as f:
                text = f.read()
        self.path = path
        self.lines = [ln[:LINE] for ln in text.splitlines()]
        self.strict = strict
        self.used = {}

    def raw(self, key):
        """The value item for ``key`` after environment expansion, or None."""
        kw = key.strip()
        n = len(kw)
        for ln in self.lines:
            c = ln.lstrip(" ")
            if c[:1] in ("#", "!"):
                continue
            if c[:n] == kw:
                rest = c[n:].lstrip(" ")
                if rest[:1] == "=":
                    return expenv(_first_item(rest[1:]))
        return None

    def get(self, key, default, kind=None):
        v = self.raw(key)
        if v is None:
            if self.strict:
                raise KeyError(f"key {key} is not found")
            out = default if kind != "s" else np.float32(default)
            self.used[key] = out
            return out
        if kind is None:
            kind = {bool: "l", int: "i", float: "d", str: "c"}.get(type(default), "c")
            if isinstance(default, np.floating):
                kind = "s" if default.dtype == np.float32 else "d"
        if kind == "c":
            out = v
        else:
            tok = _repeat(_first_item(v))
            out = {"l": _to_logical, "i": _to_int, "d": _to_real,
                   "s": lambda t: np.float32(_to_real(t))}[kind](tok)
        self.used[key] = out
        return out


def readini(f: InputFile, key, default, kind=None):
    """Function form of ``InputFile.get`` with the reference's name."""
    return f.get(key, default, kind)
