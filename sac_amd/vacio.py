"""Reader and writers of the VAC ini/out file formats.

ASCII (src/vacio.t:591-676 read, :961-1001 write): a header line;
`it t ndim neqpar nw`; the mesh size; the equation parameters; the variable
names; then one row `x(1..ndim) w(1..nw)` per cell with index 1 fastest.
Binary (src/vacio.t:679-752 read, :1004-1045 write): Fortran unformatted
sequential records (4-byte length markers) holding the header (character
400), `it t ndim neqpar nw` (int32, real64, 3 x int32), nx (int32 per axis),
eqpar (real64), the variable names (character 400), x for all axes, then
one record per variable.  A file may hold several snapshots one after the
other.

Arrays follow the Problem layout: w[nw, G3, G2, G1] (the Fortran column-
major w(G1,G2,G3,nw)), x[ndim, G3, G2, G1].
"""
from __future__ import annotations

import struct
from dataclasses import dataclass

import numpy as np

LENNAME = 400


@dataclass
class Snapshot:
    header: str
    it: int
    t: float
    ndim: int
    neqpar: int
    nw: int
    nx: tuple
    eqpar: np.ndarray
    varnames: str
    x: np.ndarray  # [ndim, n3, n2, n1] of the saved cells
    w: np.ndarray  # [nw, n3, n2, n1]


def _box(prob, fullgrid: bool):
    """(lo, hi) 0-based inclusive per axis (1..3) of the saved cells."""
    lo, hi = [0, 0, 0], [0, 0, 0]
    for d in range(prob.ndim):
        G = prob.nx[d] + 4
        lo[d], hi[d] = (0, G - 1) if fullgrid else (2, G - 3)
    return lo, hi


def _sub(a, lo, hi):
    return a[..., lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1]


def write_ini_ascii(path, prob, header="stratified_mhd33", varnames=None, it=0, t=0.0,
                    fullgrid=False, fmt="%.17e"):
    """Initial-condition file in the ASCII format (full FP64 precision in
    the value rows, which the list-directed reader accepts)."""
    lo, hi = _box(prob, fullgrid)
    names = varnames or default_varnames(prob)
    x = _sub(prob.x, lo, hi)
    w = _sub(prob.w, lo, hi)
    with open(path, "w") as f:
        f.write(header + "\n")
        f.write(f"{it} {t!r} {prob.ndim} {len(prob.eqpar)} {prob.nw}\n")
        f.write(" ".join(str(hi[d] - lo[d] + 1) for d in range(prob.ndim)) + "\n")
        f.write(" ".join(repr(float(v)) for v in prob.eqpar) + "\n")
        f.write(names + "\n")
        rows = np.concatenate([x.reshape(prob.ndim, -1), w.reshape(prob.nw, -1)], axis=0).T
        np.savetxt(f, rows, fmt=fmt)


def _rec(f, payload: bytes):
    n = struct.pack("<i", len(payload))
    f.write(n + payload + n)


def write_ini_binary(path, prob, header="stratified_mhd33", varnames=None, it=0, t=0.0,
                     fullgrid=False):
    lo, hi = _box(prob, fullgrid)
    names = varnames or default_varnames(prob)
    x = _sub(prob.x, lo, hi)
    w = _sub(prob.w, lo, hi)
    with open(path, "wb") as f:
        _rec(f, header.ljust(LENNAME).encode())
        _rec(f, struct.pack("<idiii", it, t, prob.ndim, len(prob.eqpar), prob.nw))
        _rec(f, struct.pack("<" + "i" * prob.ndim,
                            *[hi[d] - lo[d] + 1 for d in range(prob.ndim)]))
        _rec(f, np.asarray(prob.eqpar, dtype="<f8").tobytes())
        _rec(f, names.ljust(LENNAME).encode())
        _rec(f, np.ascontiguousarray(x, dtype="<f8").tobytes())
        for iw in range(prob.nw):
            _rec(f, np.ascontiguousarray(w[iw], dtype="<f8").tobytes())


def default_varnames(prob):
    if prob.ndir == 3:
        core = "rho m1 m2 m3 e b1 b2 b3 eb rhob bg1 bg2 bg3"
    elif prob.ndir == 2:
        core = "rho m1 m2 e b1 b2 eb rhob bg1 bg2"
    else:
        core = "rho m1 e b1 eb rhob bg1"
    coords = " ".join(["x", "y", "z"][:prob.ndim])
    return f"{coords} {core} gamma eta grav1 grav2 grav3 nu"


def read_ascii(path):
    """All snapshots of an ASCII ini/out file."""
    out = []
    with open(path) as f:
        lines = f.read().split("\n")
    i = 0
    while i < len(lines) and lines[i].strip() != "":
        header = lines[i].rstrip()
        it, t, ndim, neq, nw = lines[i + 1].split()
        it, t, ndim, neq, nw = int(it), float(t.replace("D", "E")), int(ndim), int(neq), int(nw)
        nx = tuple(int(v) for v in lines[i + 2].split())[:ndim]
        eq = np.array([float(v.replace("D", "E")) for v in lines[i + 3].split()])
        names = lines[i + 4].rstrip()
        ncell = int(np.prod(nx))
        rows = np.array([[float(v.replace("D", "E")) for v in lines[i + 5 + r].split()]
                         for r in range(ncell)])
        full = [1, 1, 1]
        full[:ndim] = nx
        x = rows[:, :ndim].T.reshape(ndim, full[2], full[1], full[0])
        w = rows[:, ndim:].T.reshape(nw, full[2], full[1], full[0])
        out.append(Snapshot(header, it, t, ndim, neq, nw, nx, eq, names, x, w))
        i += 5 + ncell
    return out


def _readrec(f):
    h = f.read(4)
    if len(h) < 4:
        return None
    (n,) = struct.unpack("<i", h)
    b = f.read(n)
    f.read(4)
    return b


def read_binary(path):
    """All snapshots of a binary ini/out file."""
    out = []
    with open(path, "rb") as f:
        while True:
            hb = _readrec(f)
            if hb is None:
                break
            header = hb.decode(errors="replace").rstrip()
            it, t, ndim, neq, nw = struct.unpack("<idiii", _readrec(f))
            nx = struct.unpack("<" + "i" * abs(ndim), _readrec(f))
            eq = np.frombuffer(_readrec(f), dtype="<f8").copy()
            names = _readrec(f).decode(errors="replace").rstrip()
            full = [1, 1, 1]
            full[:abs(ndim)] = nx
            x = np.frombuffer(_readrec(f), dtype="<f8").reshape(abs(ndim), full[2], full[1],
                                                                full[0]).copy()
            w = np.stack([np.frombuffer(_readrec(f), dtype="<f8").reshape(full[2], full[1],
                                                                          full[0])
                          for _ in range(nw)])
            out.append(Snapshot(header, it, t, ndim, neq, nw, tuple(nx), eq, names, x, w))
    return out


def read_log(path):
    """(header, column line, rows[it, t, dt, wmean...]) of a default log file."""
    with open(path) as f:
        lines = f.read().split("\n")
    rows = [[float(v.replace("D", "E")) for v in l.split()] for l in lines[2:] if l.strip()]
    return lines[0].rstrip(), lines[1].rstrip(), np.array(rows)


def write_par(path, filenameini, fileout, filelog, *, typefileout="auto", itmax=None, tmax=None,
              tmaxexact=False, ditsave=None, dtsave=None, courantpar=0.3, typeadvance="fourstep",
              typeB=None, nw=13, ne=8, ndim=3, wnames=None, dtpar=None, save_initial=True,
              divbfix=False, typepred=None):
    """A par file for the host program: the namelists of src/vacio.t:25-53
    with the method settings of the GPU path (cd4 on the evolving variables,
    dimsplit=F; Powell's divbfix off unless asked)."""
    tB = typeB or ["cont"] * (2 * ndim)
    lines = ["&testlist /", "&filelist",
             f"  filenameini='{filenameini}'",
             f"  filename='{filelog}','{fileout}'",
             f"  typefileout='{typefileout}'", "/", "&savelist"]
    if save_initial:
        lines.append("  tsave(1,1)=0.d0, tsave(1,2)=0.d0")
    if ditsave is not None:
        lines.append(f"  ditsave={ditsave[0]},{ditsave[1]}")
    if dtsave is not None:
        lines.append(f"  dtsave={dtsave[0]!r},{dtsave[1]!r}")
    lines += ["/", "&stoplist"]
    if itmax is not None:
        lines.append(f"  itmax={itmax}")
    if tmax is not None:
        lines.append(f"  tmax={tmax!r}")
    lines.append(f"  tmaxexact={'T' if tmaxexact else 'F'}")
    lines += ["/", "&methodlist",
              f"  wnames='{wnames or 'rho m1 m2 m3 e b1 b2 b3 eb rhob bg1 bg2 bg3'}'",
              f"  typeadvance='{typeadvance}'",
              f"  typefull={ne}*'cd4',{nw - ne}*'nul'"]
    if typepred is not None:
        lines.append(f"  typepred={ne}*'{typepred}',{nw - ne}*'nul'")
    lines += ["  dimsplit=F", "  sourcesplit=F", f"  divbfix={'T' if divbfix else 'F'}",
              "  smallfix=F", "/", "&boundlist"]
    tbs = ",".join(f"{nw}*'{t}'" for t in tB)
    lines += [f"  typeB={tbs}", "/", "&paramlist", f"  courantpar={courantpar!r}"]
    if dtpar is not None:
        lines.append(f"  dtpar={dtpar!r}")
    lines += ["/", ""]
    with open(path, "w") as f:
        f.write("\n".join(lines))
