"""input.inf reading (src/shared/m_readini.f90, m_system.f90 system__expenv)
against values returned by the reference's own readini: golden fixtures in
tests/golden/ref_readini.json and, where oracle/_ref is built, live calls.
"""
import ctypes as C
import json
from pathlib import Path

import numpy as np
import pytest

from openswpc_amd.inifile import InputFile, IniError, expenv
from oracle.oracle import ref_lib

G = json.loads((Path(__file__).parent / "golden" / "ref_readini.json").read_text())
KIND = {0: "i", 1: "s", 2: "d", 3: "c", 4: "l"}


def defaults(kind):
    d = G["defaults"]
    return {0: d["i"], 1: np.float32(d["r"]), 2: float(d["r"]), 3: d["c"], 4: d["l"]}[kind]


@pytest.fixture
def ini(monkeypatch):
    for k, v in G["env"].items():
        monkeypatch.setenv(k, v)
    return InputFile(text=G["text"])


@pytest.mark.parametrize("q", G["queries"], ids=lambda q: f"{q['key'].strip()}-{q['kind']}")
def test_readini_golden(ini, q):
    got = ini.get(q["key"], defaults(q["kind"]), KIND[q["kind"]])
    if q["kind"] == 1:
        assert np.float32(got) == np.float32(q["value"])
    else:
        assert got == q["value"]
        assert type(got) is type(q["value"])


def test_strict_mode_and_errors():
    f = InputFile(text="a = 1\nb =\nc = x\n", strict=True)
    assert f.get("a", 0) == 1
    with pytest.raises(KeyError):
        f.get("zz", 0)
    with pytest.raises(IniError):
        f.get("b", 0)
    with pytest.raises(IniError):
        f.get("c", 0)
    with pytest.raises(IniError):
        f.get("c", 0.0)


def test_expenv(monkeypatch):
    monkeypatch.setenv("AA", "xx")
    monkeypatch.delenv("NOPE_NOT_SET", raising=False)
    assert expenv("${AA}/b") == "xx/b"
    assert expenv("p${NOPE_NOT_SET}q") == "pq"
    assert expenv("plain") == "plain"


def test_live_reference_readini(tmp_path, monkeypatch):
    L = ref_lib()
    if L is None:
        pytest.skip("oracle/_ref not built (reference sources absent)")
    rng = np.random.default_rng(4)
    lines = []
    keys = []
    for n in range(40):
        key = "k%d" % n
        kind = int(rng.integers(0, 5))
        if kind == 0:
            val = str(int(rng.integers(-10 ** 6, 10 ** 6)))
        elif kind in (1, 2):
            val = "%.*g" % (int(rng.integers(1, 12)), rng.standard_normal() * 10.0 ** int(rng.integers(-5, 6)))
            val = val.replace("e", str(rng.choice(["e", "d", "E", "D"])))
        elif kind == 3:
            val = str(rng.choice(["abc", "'q u o t e d'", "x,y", "a/b", '"dq"', "${HOME}z"]))
        else:
            val = str(rng.choice([".true.", "T", "F", ".false.", "t", "f"]))
        pad = " " * int(rng.integers(0, 3))
        lines.append(f"{pad}{key}{pad} ={pad}{val}{pad} ! c")
        keys.append((key, kind))
    text = "\n".join(lines) + "\n"
    fn = tmp_path / "r.inf"
    fn.write_text(text)
    f = InputFile(text=text)
    fb = str(fn).encode()
    for key, kind in keys:
        i, r, d, l = C.c_int(-7), C.c_float(2.5), C.c_double(2.5), C.c_int(1)
        s = C.create_string_buffer(b"dflt", 256)
        kb = key.encode()
        L.ref_readini(fb, len(fb), kb, len(kb), kind, C.byref(i), C.byref(r), C.byref(d), s, C.byref(l))
        want = [i.value, r.value, d.value, s.value.decode().rstrip(" "), bool(l.value)][kind]
        got = f.get(key, defaults(kind), KIND[kind])
        if kind == 1:
            assert np.float32(got) == np.float32(want), (key, lines[int(key[1:])])
        else:
            assert got == want, (key, lines[int(key[1:])])
