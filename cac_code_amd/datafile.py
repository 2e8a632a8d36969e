"""CAC data file (read_data) reader and writer.

The format is the reference's (read_data.cpp:277-437; element.cpp:195-330,
478-540, 1086-1166; atom_vec_atomic.cpp:49-71):

- line 1 is a comment;
- header lines, matched by substring after '#' comments are cut:
  "N atoms", "M elements", "lo hi xlo xhi" (and y, z), "T atom types",
  "E element types"; the first other non-blank line is a section keyword;
- a section keyword line is followed by one more line, which is skipped
  (read_data.cpp:397), then the section's lines:
    Masses    type mass                                   (atom types)
    NAAE      etype nx ny nz, each >= 2                   (element types)
    Numint    etype nint                                  (element types)
    Intp      etype m ix iy iz w, m 1-based               (sum of nint)
    Elements  tag etype xc yc zc [ix iy iz]               (M)
    Nodes     etag k x y z nodetag | etag k x y z vx vy vz nodetag  (8 M)
    Atoms     tag type x y z [ix iy iz]                   (N)
- an unknown keyword is an error ("Unknown identifier in data file: ...").

Nodes name their element by tag, so they come after Elements (the element
map, element.cpp:496).  Node tags name mesh vertices: the writer gives the
nodes of elements meeting at one vertex the same tag, and the reader groups
integration points by them (`CACSystem.node_tags`, `lattice.vertex_clusters`).
Numbers are written with 17 significant digits, so a written and re-read
system is the same bit for bit.  Atom velocities are not in the format (the
Atoms lines carry none, and ReadData::command rejects a Velocities
section); node velocities are, in 9-word Nodes lines.

`read_data` and `write_data` run in the native library (include/cac_io.h:
the file is read whole and its large sections parsed and formatted in
parallel); `read_data_py` and `write_data_py` do the same in Python and
produce the same arrays and the same bytes.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Tuple

import numpy as np

from . import _lib
from .lattice import CACSystem, vertex_clusters

# the sections ReadData::command handles (read_data.cpp:155-215); any other
# keyword is "Unknown identifier in data file"
SECTIONS = ("Atoms", "Masses", "NAAE", "Numint", "Intp", "Elements", "Nodes")


class DataFileError(ValueError):
    pass


@dataclass
class DataFile:
    comment: str
    box_lo: np.ndarray
    box_hi: np.ndarray
    ntypes: int = 1
    entypes: int = 1
    masses: Dict[int, float] = field(default_factory=dict)
    nae: Dict[int, Tuple[int, int, int]] = field(default_factory=dict)
    numint: Dict[int, int] = field(default_factory=dict)
    intp: Dict[int, np.ndarray] = field(default_factory=dict)      # etype -> [nint][4]: ix iy iz w
    atom_tag: np.ndarray = None
    atom_type: np.ndarray = None
    x: np.ndarray = None                                         # [N][3]
    atom_image: Optional[np.ndarray] = None                      # [N][3]
    elem_tag: np.ndarray = None
    elem_type: np.ndarray = None
    xc: np.ndarray = None                                        # [M][3]
    elem_image: Optional[np.ndarray] = None
    nodex: np.ndarray = None                                     # [M][8][3]
    nodev: Optional[np.ndarray] = None                           # [M][8][3]
    nodetag: np.ndarray = None                                   # [M][8]


def _numbers(lines: List[str], nwords: int, what: str) -> np.ndarray:
    txt = " ".join(lines)
    a = np.array(txt.split(), dtype=np.float64) if txt.strip() else np.zeros(0)
    if a.size != nwords * len(lines):
        raise DataFileError(f"Incorrect {what} format in data file")
    return a.reshape(len(lines), nwords)


def _word_count(line: str) -> int:
    return len(line.split())


def read_data_py(path: str) -> DataFile:
    """Parse a CAC data file in Python (ReadData::command, read_data.cpp:70-240)."""
    with open(path) as fh:
        lines = fh.read().split("\n")
    if lines and lines[-1] == "":
        lines.pop()                                           # the newline that ends the last line
    if not lines or (len(lines) == 1 and not lines[0]):
        raise DataFileError("Unexpected end of data file")
    d = DataFile(comment=lines[0], box_lo=np.zeros(3), box_hi=np.zeros(3))
    natoms = nelements = 0
    i = 1
    keyword = None
    # header (ReadData::header, read_data.cpp:277-368)
    while i < len(lines):
        line = lines[i].split("#", 1)[0]
        i += 1
        if not line.strip():
            continue
        if "atoms" in line:
            natoms = int(line.split()[0])
        elif "elements" in line:
            nelements = int(line.split()[0])
        elif "xlo xhi" in line:
            d.box_lo[0], d.box_hi[0] = map(float, line.split()[:2])
        elif "ylo yhi" in line:
            d.box_lo[1], d.box_hi[1] = map(float, line.split()[:2])
        elif "zlo zhi" in line:
            d.box_lo[2], d.box_hi[2] = map(float, line.split()[:2])
        elif "atom types" in line:
            d.ntypes = int(line.split()[0])
        elif "element types" in line:
            d.entypes = int(line.split()[0])
        else:
            keyword = line.strip()
            break
    if natoms < 0 or nelements < 0:
        raise DataFileError("System in data file is too big")

    def take(k: int) -> List[str]:
        nonlocal i
        if i + k > len(lines):
            raise DataFileError("Unexpected end of data file")
        out = lines[i:i + k]
        i += k
        return out

    elem_row: Dict[int, int] = {}
    while keyword:
        if keyword not in SECTIONS:
            raise DataFileError(f"Unknown identifier in data file: {keyword}")
        take(1)                                               # the line after the keyword (read_data.cpp:397)
        if keyword == "Masses":
            for ln in take(d.ntypes):
                w = ln.split()
                d.masses[int(w[0])] = float(w[1])
        elif keyword == "NAAE":                               # Element::set_nae, element.cpp:1086-1115
            for ln in take(d.entypes):
                w = ln.split()
                if len(w) < 4:
                    raise DataFileError("Invalid nae line in data file")
                t, nx, ny, nz = (int(v) for v in w[:4])
                if t < 1 or t > d.entypes:
                    raise DataFileError("Invalid type for nae set")
                if nx < 2 or ny < 2 or nz < 2:
                    raise DataFileError("Invalid naae values")
                d.nae[t] = (nx, ny, nz)
        elif keyword == "Numint":                             # Element::set_numint, element.cpp:1126-1140
            for ln in take(d.entypes):
                w = ln.split()
                if len(w) < 2:
                    raise DataFileError("Invalid numint line in data file")
                t = int(w[0])
                if t < 1 or t > d.entypes:
                    raise DataFileError("Invalid type for numint set")
                d.numint[t] = int(w[1])
        elif keyword == "Intp":                               # Element::set_intp, element.cpp:1144-1166
            total = sum(d.numint.values())
            rows = take(total)
            for t, n in d.numint.items():
                d.intp[t] = np.zeros((n, 4))
            for ln in rows:
                w = ln.split()
                t, m = int(w[0]), int(w[1]) - 1
                if t not in d.intp or not 0 <= m < d.intp[t].shape[0]:
                    raise DataFileError("Invalid Intp line in data file")
                d.intp[t][m] = [int(w[2]), int(w[3]), int(w[4]), float(w[5])]
        elif keyword == "Elements":                           # Element::data_elements, element.cpp:195-314
            rows = take(nelements)
            nw = _word_count(rows[0]) if rows else 5
            if nw not in (5, 8):
                raise DataFileError("Incorrect element format in data file")
            a = _numbers(rows, nw, "element")
            d.elem_tag = a[:, 0].astype(np.int64)
            d.elem_type = a[:, 1].astype(np.int64)
            if np.any((d.elem_type < 1) | (d.elem_type > d.entypes)):
                raise DataFileError("Invalid atom type in Elements section of data file")
            d.xc = np.ascontiguousarray(a[:, 2:5])
            d.elem_image = a[:, 5:8].astype(np.int64) if nw == 8 else None
            elem_row = {int(t): r for r, t in enumerate(d.elem_tag)}
            d.nodex = np.zeros((nelements, 8, 3))
            d.nodev = np.zeros((nelements, 8, 3))
            d.nodetag = np.zeros((nelements, 8), dtype=np.int64)
        elif keyword == "Nodes":                              # Element::data_nodes, element.cpp:478-540
            rows = take(8 * nelements)
            if not elem_row and nelements:
                raise DataFileError("Nodes section before Elements in data file")
            nw = _word_count(rows[0]) if rows else 6
            if nw not in (6, 9):
                raise DataFileError("Incorrect node format in data file")
            a = _numbers(rows, nw, "node")
            r = np.array([elem_row.get(int(t), -1) for t in a[:, 0]])
            k = a[:, 1].astype(np.int64) - 1
            ok = r >= 0                                       # nodes of unknown elements are skipped (:496)
            if np.any((k[ok] < 0) | (k[ok] > 7)):
                raise DataFileError("Invalid node index in Nodes section of data file")
            d.nodex[r[ok], k[ok]] = a[ok, 2:5]
            if nw == 9:
                d.nodev[r[ok], k[ok]] = a[ok, 5:8]
            d.nodetag[r[ok], k[ok]] = a[ok, nw - 1].astype(np.int64)
        elif keyword == "Atoms":                              # AtomVecAtomic::data_atom, atom_vec_atomic.cpp:49-71
            rows = take(natoms)
            nw = _word_count(rows[0]) if rows else 5
            if nw not in (5, 8):
                raise DataFileError("Incorrect atom format in data file")
            a = _numbers(rows, nw, "atom")
            d.atom_tag = a[:, 0].astype(np.int64)
            d.atom_type = a[:, 1].astype(np.int64)
            if np.any((d.atom_type < 1) | (d.atom_type > d.ntypes)):
                raise DataFileError("Invalid atom type in Atoms section of data file")
            d.x = np.ascontiguousarray(a[:, 2:5])
            d.atom_image = a[:, 5:8].astype(np.int64) if nw == 8 else None
        else:
            raise DataFileError(f"Unsupported section in data file: {keyword}")
        # next keyword (ReadData::parse_keyword): the next non-blank line
        keyword = None
        while i < len(lines):
            ln = lines[i]
            i += 1
            s = ln.strip()
            if s and not s.startswith("#"):
                keyword = s.split("#", 1)[0].strip()
                break
    if natoms > 0 and d.x is None:
        raise DataFileError("No atoms in data file")
    if nelements > 0 and d.xc is None:
        raise DataFileError("No elements in data file")
    if d.x is None:
        d.atom_tag = np.zeros(0, np.int64)
        d.atom_type = np.zeros(0, np.int64)
        d.x = np.zeros((0, 3))
    if d.xc is None:
        d.elem_tag = np.zeros(0, np.int64)
        d.elem_type = np.zeros(0, np.int64)
        d.xc = np.zeros((0, 3))
        d.nodex = np.zeros((0, 8, 3))
        d.nodev = np.zeros((0, 8, 3))
        d.nodetag = np.zeros((0, 8), np.int64)
    return d


def _g(v: float) -> str:
    return format(float(v), ".17g")


def write_data_py(path: str, d: DataFile, node_velocities: bool = True) -> None:
    """Write d in the reference's format in Python (one blank line after each
    section keyword, as parse_keyword expects)."""
    N, M = d.x.shape[0], d.xc.shape[0]
    out = [d.comment.replace("\n", " "), ""]
    out.append(f"{N} atoms")
    out.append(f"{M} elements")
    out.append(f"{d.ntypes} atom types")
    out.append(f"{d.entypes} element types")
    for ax, c in enumerate("xyz"):
        out.append(f"{_g(d.box_lo[ax])} {_g(d.box_hi[ax])} {c}lo {c}hi")
    out.append("")

    def section(name, rows):
        out.extend([name, ""])
        out.extend(rows)
        out.append("")

    if d.masses:
        section("Masses", [f"{t} {_g(m)}" for t, m in sorted(d.masses.items())])
    if M:
        section("NAAE", [f"{t} {a} {b} {c}" for t, (a, b, c) in sorted(d.nae.items())])
        section("Numint", [f"{t} {n}" for t, n in sorted(d.numint.items())])
        rows = []
        for t in sorted(d.intp):
            for m, (ix, iy, iz, w) in enumerate(d.intp[t]):
                rows.append(f"{t} {m + 1} {int(ix)} {int(iy)} {int(iz)} {_g(w)}")
        section("Intp", rows)
        img = d.elem_image
        section("Elements", [
            f"{int(d.elem_tag[e])} {int(d.elem_type[e])} {_g(d.xc[e, 0])} {_g(d.xc[e, 1])} {_g(d.xc[e, 2])}"
            + (f" {int(img[e, 0])} {int(img[e, 1])} {int(img[e, 2])}" if img is not None else "")
            for e in range(M)])
        rows = []
        for e in range(M):
            for k in range(8):
                p = d.nodex[e, k]
                s = f"{int(d.elem_tag[e])} {k + 1} {_g(p[0])} {_g(p[1])} {_g(p[2])}"
                if node_velocities:
                    v = d.nodev[e, k] if d.nodev is not None else np.zeros(3)
                    s += f" {_g(v[0])} {_g(v[1])} {_g(v[2])}"
                rows.append(s + f" {int(d.nodetag[e, k])}")
        section("Nodes", rows)
    if N:
        img = d.atom_image
        section("Atoms", [
            f"{int(d.atom_tag[i])} {int(d.atom_type[i])} {_g(d.x[i, 0])} {_g(d.x[i, 1])} {_g(d.x[i, 2])}"
            + (f" {int(img[i, 0])} {int(img[i, 1])} {int(img[i, 2])}" if img is not None else "")
            for i in range(N)])
    with open(path, "w") as fh:
        fh.write("\n".join(out))


def _arr(ptr, shape, dtype):
    n = int(np.prod(shape))
    if n == 0:
        return np.zeros(shape, dtype=dtype)
    return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype, copy=True).reshape(shape)


def read_data(path: str) -> DataFile:
    """Parse a CAC data file in the native library (cac_data_read; the
    reference's ReadData::command, read_data.cpp:70-240).  Raises
    DataFileError with the reference's message for a malformed file, OSError
    for one that cannot be read."""
    L = _lib.lib()
    h = C.c_void_p()
    rc = L.cac_data_read(str(path).encode(), C.byref(h))
    if rc != _lib.CAC_OK:
        msg = L.cac_io_last_error().decode()
        raise (OSError if rc == _lib.CAC_ERR_IO else DataFileError)(msg)
    try:
        v = L.cac_data_get(h).contents
        N, M = int(v.natoms), int(v.nelements)
        d = DataFile(comment=v.comment.decode(), box_lo=np.array(v.lo[:]), box_hi=np.array(v.hi[:]),
                     ntypes=int(v.ntypes), entypes=int(v.entypes))
        if v.masses:
            m = _arr(v.masses, (d.ntypes,), np.float64)
            d.masses = {t + 1: float(m[t]) for t in range(d.ntypes) if not np.isnan(m[t])}
        if v.nae:
            a = _arr(v.nae, (d.entypes, 3), np.int64)
            d.nae = {t + 1: tuple(int(c) for c in a[t]) for t in range(d.entypes) if a[t, 0] > 0}
        if v.numint:
            ni = _arr(v.numint, (d.entypes,), np.int64)
            d.numint = {t + 1: int(ni[t]) for t in range(d.entypes) if ni[t] >= 0}
            if v.intp_w or v.nintp:
                ijk = _arr(v.intp_ijk, (int(v.nintp), 3), np.float64)
                w = _arr(v.intp_w, (int(v.nintp),), np.float64)
                row = 0
                for t in range(1, d.entypes + 1):
                    n = d.numint.get(t, -1)
                    if n < 0:
                        continue
                    d.intp[t] = np.column_stack([ijk[row:row + n], w[row:row + n]]) if n else np.zeros((0, 4))
                    row += n
        d.atom_tag = _arr(v.atom_tag, (N,), np.int64)
        d.atom_type = _arr(v.atom_type, (N,), np.int64)
        d.x = _arr(v.x, (N, 3), np.float64)
        d.atom_image = _arr(v.atom_image, (N, 3), np.int64) if v.atom_image else None
        d.elem_tag = _arr(v.elem_tag, (M,), np.int64)
        d.elem_type = _arr(v.elem_type, (M,), np.int64)
        d.xc = _arr(v.xc, (M, 3), np.float64)
        d.elem_image = _arr(v.elem_image, (M, 3), np.int64) if v.elem_image else None
        d.nodex = _arr(v.nodex, (M, 8, 3), np.float64)
        d.nodev = _arr(v.nodev, (M, 8, 3), np.float64)
        d.nodetag = _arr(v.nodetag, (M, 8), np.int64)
    finally:
        L.cac_data_free(h)
    return d


def write_data(path: str, d: DataFile, node_velocities: bool = True) -> None:
    """Write d in the reference's format in the native library
    (cac_data_write): the bytes write_data_py writes."""
    N, M = d.x.shape[0], d.xc.shape[0]
    keep = []

    def ptr(a, dtype, ctype):
        a = np.ascontiguousarray(a, dtype=dtype)
        keep.append(a)
        return a.ctypes.data_as(C.POINTER(ctype))

    v = _lib.DataView()
    v.comment = d.comment.encode()
    v.lo[:] = [float(c) for c in d.box_lo]
    v.hi[:] = [float(c) for c in d.box_hi]
    v.ntypes, v.entypes = int(d.ntypes), int(d.entypes)
    if d.masses:
        m = np.full(d.ntypes, np.nan)
        for t, val in d.masses.items():
            m[t - 1] = val
        v.masses = ptr(m, np.float64, C.c_double)
    nae = np.zeros((d.entypes, 3), np.int32)
    for t, val in d.nae.items():
        nae[t - 1] = val
    ni = np.full(d.entypes, -1, np.int32)
    for t, val in d.numint.items():
        ni[t - 1] = val
    rows = []
    for t in range(1, d.entypes + 1):
        if t in d.intp:
            if ni[t - 1] != d.intp[t].shape[0]:
                raise DataFileError("Intp rows differ from Numint")
            rows.append(np.asarray(d.intp[t], np.float64).reshape(-1, 4))
    ip = np.concatenate(rows) if rows else np.zeros((0, 4))
    v.nae = ptr(nae, np.int32, C.c_int)
    v.numint = ptr(ni, np.int32, C.c_int)
    v.nintp = ip.shape[0]
    v.intp_ijk = ptr(ip[:, :3], np.int32, C.c_int)
    v.intp_w = ptr(ip[:, 3], np.float64, C.c_double)
    v.natoms, v.nelements = N, M
    v.atom_tag = ptr(d.atom_tag, np.int64, C.c_int64)
    v.atom_type = ptr(d.atom_type, np.int32, C.c_int)
    v.x = ptr(d.x, np.float64, C.c_double)
    if d.atom_image is not None:
        v.atom_image = ptr(d.atom_image, np.int32, C.c_int)
    v.elem_tag = ptr(d.elem_tag, np.int64, C.c_int64)
    v.elem_type = ptr(d.elem_type, np.int32, C.c_int)
    v.xc = ptr(d.xc, np.float64, C.c_double)
    if d.elem_image is not None:
        v.elem_image = ptr(d.elem_image, np.int32, C.c_int)
    v.nodex = ptr(d.nodex.reshape(M, 8, 3) if M else np.zeros((0, 8, 3)), np.float64, C.c_double)
    nodev = d.nodev if d.nodev is not None else np.zeros((M, 8, 3))
    v.nodev = ptr(nodev.reshape(M, 8, 3) if M else np.zeros((0, 8, 3)), np.float64, C.c_double)
    v.nodetag = ptr(d.nodetag, np.int64, C.c_int64)
    L = _lib.lib()
    rc = L.cac_data_write(str(path).encode(), C.byref(v), int(bool(node_velocities)))
    if rc != _lib.CAC_OK:
        msg = L.cac_io_last_error().decode()
        raise (OSError if rc == _lib.CAC_ERR_IO else DataFileError)(msg)


def from_system(s: CACSystem, comment: str = "CAC data file") -> DataFile:
    """DataFile of a generated system: one atom type and one element type,
    node tags = mesh vertex ids (1-based) where the mesh is known, else one
    tag per node."""
    M = s.n_elements
    d = DataFile(comment=comment, box_lo=np.array(s.box_lo, float), box_hi=np.array(s.box_hi, float))
    d.masses = {1: float(s.mass)}
    d.atom_tag = np.arange(1, s.n_atoms + 1)
    d.atom_type = np.ones(s.n_atoms, np.int64)
    d.x = np.ascontiguousarray(s.x, dtype=np.float64)
    d.elem_tag = np.arange(1, M + 1)
    d.elem_type = np.ones(M, np.int64)
    d.nodex = np.ascontiguousarray(s.nodex, dtype=np.float64)
    d.nodev = np.ascontiguousarray(s.nodev, dtype=np.float64)
    d.xc = d.nodex.sum(axis=1) / 8.0 if M else np.zeros((0, 3))
    if M:
        d.nae = {1: tuple(int(v) for v in s.nae)}
        d.numint = {1: int(s.intp_w.size)}
        d.intp = {1: np.column_stack([s.intp_ijk, s.intp_w]).astype(np.float64)}
        tags = np.arange(1, 8 * M + 1, dtype=np.int64)
        cl = vertex_clusters(s)
        if cl is not None:
            start, slots = cl
            for c in range(len(start) - 1):
                tags[slots[start[c]:start[c + 1]]] = c + 1
        d.nodetag = tags.reshape(M, 8)
    else:
        d.nodetag = np.zeros((0, 8), np.int64)
    return d


def to_system(d: DataFile, center_tol: float = 0.1, a0: float = 0.0) -> CACSystem:
    """CACSystem of a data file with one atom type and one element type.  The
    element centers must lie within center_tol of their node centroids
    (Element::element_size, element.cpp:608-622).  a0 (the fcc lattice
    constant, if known) only sets the atom-list order hint."""
    if d.ntypes != 1 or d.entypes != 1:
        raise DataFileError("the device path runs one atom type and one element type")
    M = d.xc.shape[0]
    mass = float(d.masses.get(1, 0.0))
    if M:
        cen = d.nodex.sum(axis=1) / 8.0
        if np.any(np.linalg.norm(cen - d.xc, axis=1) > center_tol):
            raise DataFileError("element center differs from the node centroid")
        nae = tuple(int(v) for v in d.nae[1])
        ip = d.intp[1]
        ijk = np.ascontiguousarray(ip[:, :3].astype(np.int32))
        w = np.ascontiguousarray(ip[:, 3])
    else:
        nae = (2, 2, 2)
        ijk = np.zeros((0, 3), np.int32)
        w = np.zeros(0)
    s = CACSystem(d.box_lo.copy(), d.box_hi.copy(), mass, np.ascontiguousarray(d.x), np.zeros_like(d.x), nae,
                  np.ascontiguousarray(d.nodex), np.ascontiguousarray(d.nodev) if d.nodev is not None
                  else np.zeros_like(d.nodex), ijk, w, a0=a0)
    s.node_tags = np.ascontiguousarray(d.nodetag) if M else None
    return s
