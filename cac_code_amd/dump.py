"""Atom and element dumps (DumpAtom, dump_atom.cpp:35-157; DumpElement,
dump_element.cpp:72-263), written by the native library
(cac_dump_atom_write / cac_dump_element_write, include/cac_io.h), and
readers of the same snapshots.

Atom dump, one snapshot:
    ITEM: TIMESTEP / step / ITEM: NUMBER OF ATOMS / count /
    ITEM: BOX BOUNDS pp pp pp / three "lo hi" lines (%g) /
    ITEM: ATOMS id type x y z fx fy fz vx vy vz / one line per atom, then
    one per element node with type 1.
Element dump (Tecplot finite-element point format), one snapshot:
    title="results for cac" / variables="x", ..., "fz" /
    zone t="load step N", n=  8M e =  M datapacking=point, zonetype=febrick /
    8M node lines "x y z vx vy vz fx fy fz" (%g) / M connectivity lines.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import List, Optional

import numpy as np

from . import _lib

ATOM_COLUMNS = "id type x y z fx fy fz vx vy vz"


def _p(a, ctype=C.c_double):
    return None if a is None else a.ctypes.data_as(C.POINTER(ctype))


def _check(rc: int):
    if rc != _lib.CAC_OK:
        msg = _lib.lib().cac_io_last_error().decode()
        if rc == _lib.CAC_ERR_IO:
            raise OSError(msg)
        raise _lib.CacError(rc, msg)


def write_dump_atom(path: str, step: int, lo, hi, x, f, v, nodex=None, nodef=None, nodev=None, tag=None,
                    type=None, nodetag=None, periodic=(1, 1, 1), append: bool = False) -> int:
    """One DumpAtom snapshot of atoms [n][3] and element nodes [ne][8][3];
    returns the number of lines (atoms + 8 ne).  A '*' in path is replaced by
    the step."""
    f64 = lambda a: np.ascontiguousarray(a, dtype=np.float64)  # noqa: E731
    x, f, v = f64(x).reshape(-1, 3), f64(f).reshape(-1, 3), f64(v).reshape(-1, 3)
    n = x.shape[0]
    if nodex is None:
        nodex = nodef = nodev = np.zeros((0, 8, 3))
    nodex, nodef, nodev = f64(nodex).reshape(-1, 8, 3), f64(nodef).reshape(-1, 8, 3), f64(nodev).reshape(-1, 8, 3)
    ne = nodex.shape[0]
    if f.shape[0] != n or v.shape[0] != n or nodef.shape[0] != ne or nodev.shape[0] != ne:
        raise ValueError("dump arrays differ in length")
    tag = None if tag is None else np.ascontiguousarray(tag, dtype=np.int64)
    type = None if type is None else np.ascontiguousarray(type, dtype=np.int32)
    nodetag = None if nodetag is None else np.ascontiguousarray(nodetag, dtype=np.int64).reshape(-1)
    lo, hi = f64(lo), f64(hi)
    per = np.ascontiguousarray(periodic, dtype=np.int32)
    _check(_lib.lib().cac_dump_atom_write(str(path).encode(), int(append), int(step), _p(lo), _p(hi),
                                          _p(per, C.c_int), n, _p(tag, C.c_int64), _p(type, C.c_int), _p(x),
                                          _p(f), _p(v), ne, _p(nodetag, C.c_int64), _p(nodex), _p(nodef),
                                          _p(nodev)))
    return n + 8 * ne


def write_dump_element(path: str, step: int, nodex, nodev, nodef, append: bool = False) -> int:
    """One DumpElement snapshot of element nodes [ne][8][3]; returns ne."""
    f64 = lambda a: np.ascontiguousarray(a, dtype=np.float64).reshape(-1, 8, 3)  # noqa: E731
    nodex, nodev, nodef = f64(nodex), f64(nodev), f64(nodef)
    ne = nodex.shape[0]
    if nodev.shape[0] != ne or nodef.shape[0] != ne:
        raise ValueError("dump arrays differ in length")
    _check(_lib.lib().cac_dump_element_write(str(path).encode(), int(append), int(step), ne, _p(nodex), _p(nodev),
                                             _p(nodef)))
    return ne


@dataclass
class AtomSnapshot:
    step: int
    bounds: str
    box: np.ndarray           # [3][2]
    id: np.ndarray
    type: np.ndarray
    x: np.ndarray
    f: np.ndarray
    v: np.ndarray


@dataclass
class ElementSnapshot:
    step: int
    nodes: np.ndarray         # [8M][9]: x y z vx vy vz fx fy fz
    conn: np.ndarray          # [M][8], 1-based node lines


def read_dump_atom(path: str) -> List[AtomSnapshot]:
    """Every snapshot of an atom dump."""
    lines = open(path).read().split("\n")
    out, i = [], 0
    while i < len(lines) and lines[i].startswith("ITEM: TIMESTEP"):
        step = int(lines[i + 1])
        count = int(lines[i + 3])
        bounds = lines[i + 4][len("ITEM: BOX BOUNDS "):]
        box = np.array([[float(w) for w in lines[i + 5 + d].split()] for d in range(3)])
        if lines[i + 8] != "ITEM: ATOMS " + ATOM_COLUMNS:
            raise ValueError(f"not an atom dump: {lines[i + 8]!r}")
        rows = np.array(" ".join(lines[i + 9:i + 9 + count]).split(), dtype=np.float64).reshape(count, 11)
        out.append(AtomSnapshot(step, bounds, box, rows[:, 0].astype(np.int64), rows[:, 1].astype(np.int64),
                                rows[:, 2:5], rows[:, 5:8], rows[:, 8:11]))
        i += 9 + count
    return out


def read_dump_element(path: str) -> List[ElementSnapshot]:
    """Every snapshot of an element dump."""
    lines = open(path).read().split("\n")
    out, i = [], 0
    while i + 2 < len(lines) and lines[i].startswith("title="):
        zone = lines[i + 2]
        step = int(zone.split('"load step ')[1].split('"')[0])
        w = zone.split(",")[1].split()
        n, ne = int(w[1]), int(w[4])
        a = i + 3
        nodes = np.array(" ".join(lines[a:a + n]).split(), dtype=np.float64).reshape(n, 9)
        conn = np.array(" ".join(lines[a + n:a + n + ne]).split(), dtype=np.int64).reshape(ne, 8)
        out.append(ElementSnapshot(step, nodes, conn))
        i = a + n + ne
    return out
