"""Nastran bulk-data writer for shell meshes (GRID / GRID*, CQUAD4, SPC /
SPC*), the format TACSMeshLoader reads (src/io/TACSMeshLoader.cpp:569-1098).

Element connectivity is given in the shell basis' tensor order (corners
(0,0), (1,0), (0,1), (1,1)); CQUAD4 cards list the corners
counter-clockwise, i.e. tensor corners 0, 1, 3, 2."""
import numpy as np


def _real16(v):
    """A real in at most 16 columns (9 significant digits)."""
    s = f"{v:.9e}" if v != 0.0 else "0.0"
    assert len(s) <= 16
    return s


def _real8(v):
    """A real in at most 8 columns (Nastran implied-exponent form allowed)."""
    if v == 0.0:
        return "0."
    for digits in range(6, 0, -1):
        s = f"{v:.{digits}e}"
        mant, exp = s.split("e")
        e = int(exp)
        t = f"{mant}{'+' if e >= 0 else '-'}{abs(e)}"
        if len(t) <= 8:
            return t
    raise ValueError(v)


def bdf_text(Xpts, conn, spcs=(), node_ids=None, elem_ids=None, comps=None, long_field=True, title="shell mesh"):
    """Bulk-data text.  Xpts [n, 3]; conn [e, 4] tensor order (0-based);
    spcs = iterable of (node, dof_digits, value) with 0-based node indices;
    node_ids / elem_ids map indices to file ids (default 1-based order)."""
    Xpts = np.asarray(Xpts, dtype=float).reshape(-1, 3)
    conn = np.asarray(conn).reshape(-1, 4)
    nn, ne = Xpts.shape[0], conn.shape[0]
    node_ids = np.arange(1, nn + 1) if node_ids is None else np.asarray(node_ids)
    elem_ids = np.arange(1, ne + 1) if elem_ids is None else np.asarray(elem_ids)
    comps = np.ones(ne, dtype=int) if comps is None else np.asarray(comps)
    out = [f"$ {title}", "SOL 103", "CEND", "BEGIN BULK"]
    for k in range(nn):
        x, y, z = Xpts[k]
        if long_field:
            out.append(f"{'GRID*':<8}{node_ids[k]:>16d}{'':16}{_real16(x):>16}{_real16(y):>16}*")
            out.append(f"{'*':<8}{_real16(z):>16}")
        else:
            out.append(f"{'GRID':<8}{node_ids[k]:>8d}{'':8}{_real8(x):>8}{_real8(y):>8}{_real8(z):>8}")
    for e in range(ne):
        a, b, c, d = conn[e]
        ids = [node_ids[a], node_ids[b], node_ids[d], node_ids[c]]  # counter-clockwise
        out.append(f"{'CQUAD4':<8}{elem_ids[e]:>8d}{comps[e]:>8d}" + "".join(f"{i:>8d}" for i in ids))
    for node, dofs, val in spcs:
        if long_field:
            out.append(f"{'SPC*':<8}{1:>16d}{node_ids[node]:>16d}{dofs:>16}{_real16(val):>16}")
        else:
            out.append(f"{'SPC':<8}{1:>8d}{node_ids[node]:>8d}{dofs:>8}{_real8(val):>8}")
    out.append("ENDDATA")
    return "\n".join(out) + "\n"


def write_bdf(path, *args, **kw):
    with open(path, "w") as f:
        f.write(bdf_text(*args, **kw))
