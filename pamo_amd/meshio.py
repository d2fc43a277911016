"""Triangle-mesh files for the command line and the verifier: OBJ and PLY
(ASCII and binary little-endian), geometry only.

The reference reads and writes meshes through trimesh
(example.py:24,40-43: ``trimesh.load(path, force='mesh', process=False)``,
``trimesh.Trimesh(vertices, faces).export(path)``).  trimesh is not a
dependency here, so this module covers what that path needs:

- vertices in file order, none dropped or merged (``process=False``);
- OBJ ``f`` records take the vertex index of each ``v/vt/vn`` token;
  negative indices count back from the last vertex read;
- quads split as (0,1,2),(2,3,0) and larger polygons as the fan (0,i,i+1),
  the splits trimesh's OBJ loader makes;
- texture coordinates, normals, groups and materials are ignored.

Writing keeps every vertex (trimesh's export of a processed Trimesh would
merge coincident ones).  Coordinates are written with 9 significant digits,
which round-trips float32 exactly.
"""
import os

import numpy as np

__all__ = ["load", "save", "load_obj", "save_obj", "load_ply", "save_ply"]


def _tri_split(poly):
    n = len(poly)
    if n == 3:
        return [poly]
    if n == 4:
        return [[poly[0], poly[1], poly[2]], [poly[2], poly[3], poly[0]]]
    return [[poly[0], poly[i], poly[i + 1]] for i in range(1, n - 1)]


def load_obj(path):
    """(vertices float64 [V,3], faces int64 [F,3]) from an OBJ file."""
    vlines = []
    flines = []
    with open(path, "r", errors="replace") as fh:
        for line in fh:
            if line.startswith("v "):
                vlines.append(line)
            elif line.startswith("f "):
                flines.append(line)
    if vlines:
        rows = [ln.split() for ln in vlines]
        if all(len(r) == 4 for r in rows):
            verts = np.array(" ".join(ln[1:] for ln in (l.strip() for l in vlines)).split(),
                             np.float64).reshape(-1, 3)
        else:
            verts = np.array([r[1:4] for r in rows], np.float64)
    else:
        verts = np.zeros((0, 3), np.float64)
    nv = len(verts)
    simple = flines and all("/" not in ln for ln in flines[:64])
    faces = None
    if simple:
        try:
            tok = np.array(" ".join(ln[2:] for ln in flines).split(), np.int64)
            if tok.size == 3 * len(flines):
                faces = tok.reshape(-1, 3)
        except ValueError:
            faces = None
    if faces is None:
        tris = []
        for ln in flines:
            poly = [int(t.split("/")[0]) for t in ln.split()[1:]]
            if len(poly) < 3:
                continue
            tris.extend(_tri_split(poly))
        faces = np.array(tris, np.int64).reshape(-1, 3)
    faces = np.where(faces < 0, faces + nv, faces - 1)
    if faces.size and (faces.min() < 0 or faces.max() >= nv):
        raise ValueError("%s: face index out of range (%d vertices)" % (path, nv))
    return verts, faces


def save_obj(path, verts, faces):
    verts = np.asarray(verts, np.float64).reshape(-1, 3)
    faces = np.asarray(faces, np.int64).reshape(-1, 3)
    with open(path, "w") as fh:
        fh.write("".join("v %.9g %.9g %.9g\n" % tuple(p) for p in verts))
        fh.write("".join("f %d %d %d\n" % tuple(t) for t in (faces + 1)))


_PLY_TYPES = {"char": "i1", "int8": "i1", "uchar": "u1", "uint8": "u1", "short": "i2", "int16": "i2",
              "ushort": "u2", "uint16": "u2", "int": "i4", "int32": "i4", "uint": "u4", "uint32": "u4",
              "float": "f4", "float32": "f4", "double": "f8", "float64": "f8"}


def _ply_header(fh):
    if fh.readline().strip() != b"ply":
        raise ValueError("not a PLY file")
    fmt = None
    elements = []
    while True:
        line = fh.readline()
        if not line:
            raise ValueError("PLY header has no end_header")
        w = line.decode("ascii", "replace").split()
        if not w or w[0] in ("comment", "obj_info"):
            continue
        if w[0] == "format":
            fmt = w[1]
        elif w[0] == "element":
            elements.append([w[1], int(w[2]), []])
        elif w[0] == "property":
            if w[1] == "list":
                elements[-1][2].append((w[4], "list", _PLY_TYPES[w[2]], _PLY_TYPES[w[3]]))
            else:
                elements[-1][2].append((w[2], _PLY_TYPES[w[1]]))
        elif w[0] == "end_header":
            return fmt, elements


def load_ply(path):
    """(vertices float64 [V,3], faces int64 [F,3]) from an ASCII or binary
    little-endian PLY file; polygons are split as in load_obj."""
    with open(path, "rb") as fh:
        fmt, elements = _ply_header(fh)
        if fmt not in ("ascii", "binary_little_endian"):
            raise ValueError("PLY format %s is not supported" % fmt)
        verts = np.zeros((0, 3))
        faces = np.zeros((0, 3), np.int64)
        text = fh.read().decode("ascii", "replace").split() if fmt == "ascii" else None
        pos = 0
        data = None if text is not None else fh
        for name, count, props in elements:
            if text is not None:
                if any(p[1] == "list" for p in props):
                    rows = []
                    for _ in range(count):
                        row = []
                        for p in props:
                            if p[1] == "list":
                                n = int(text[pos]); pos += 1
                                row.append([int(x) for x in text[pos:pos + n]]); pos += n
                            else:
                                row.append(float(text[pos])); pos += 1
                        rows.append(row)
                    table = rows
                else:
                    k = len(props)
                    table = np.array(text[pos:pos + k * count], np.float64).reshape(count, k)
                    pos += k * count
            else:
                if any(p[1] == "list" for p in props):
                    table = []
                    for _ in range(count):
                        row = []
                        for p in props:
                            if p[1] == "list":
                                n = int(np.frombuffer(data.read(np.dtype(p[2]).itemsize), "<" + p[2])[0])
                                row.append(np.frombuffer(data.read(n * np.dtype(p[3]).itemsize),
                                                         "<" + p[3]).tolist())
                            else:
                                row.append(float(np.frombuffer(data.read(np.dtype(p[1]).itemsize),
                                                               "<" + p[1])[0]))
                        table.append(row)
                else:
                    dt = np.dtype([(p[0], "<" + p[1]) for p in props])
                    table = np.frombuffer(data.read(dt.itemsize * count), dt, count)
            if name == "vertex":
                names = [p[0] for p in props]
                idx = [names.index(c) for c in ("x", "y", "z")]
                if isinstance(table, np.ndarray) and table.dtype.names:
                    verts = np.stack([table[c].astype(np.float64) for c in ("x", "y", "z")], axis=1)
                elif isinstance(table, np.ndarray):
                    verts = table[:, idx]
                else:
                    verts = np.array([[r[i] for i in idx] for r in table], np.float64)
            elif name == "face":
                li = [i for i, p in enumerate(props) if p[1] == "list"][0]
                tris = []
                for r in table:
                    if len(r[li]) >= 3:
                        tris.extend(_tri_split(list(r[li])))
                faces = np.array(tris, np.int64).reshape(-1, 3)
    if faces.size and (faces.min() < 0 or faces.max() >= len(verts)):
        raise ValueError("%s: face index out of range (%d vertices)" % (path, len(verts)))
    return verts, faces


def save_ply(path, verts, faces):
    """Binary little-endian PLY: float x,y,z and uchar-counted int faces."""
    verts = np.ascontiguousarray(verts, "<f4").reshape(-1, 3)
    faces = np.asarray(faces, np.int64).reshape(-1, 3)
    rec = np.zeros(len(faces), np.dtype([("n", "u1"), ("i", "<i4", (3,))]))
    rec["n"] = 3
    rec["i"] = faces
    head = ("ply\nformat binary_little_endian 1.0\nelement vertex %d\nproperty float x\nproperty float y\n"
            "property float z\nelement face %d\nproperty list uchar int vertex_indices\nend_header\n"
            % (len(verts), len(faces)))
    with open(path, "wb") as fh:
        fh.write(head.encode("ascii"))
        fh.write(verts.tobytes())
        fh.write(rec.tobytes())


def load(path):
    ext = os.path.splitext(path)[1].lower()
    if ext == ".obj":
        return load_obj(path)
    if ext == ".ply":
        return load_ply(path)
    raise ValueError("unsupported mesh format %r (use .obj or .ply)" % ext)


def save(path, verts, faces):
    ext = os.path.splitext(path)[1].lower()
    if ext == ".obj":
        return save_obj(path, verts, faces)
    if ext == ".ply":
        return save_ply(path, verts, faces)
    raise ValueError("unsupported mesh format %r (use .obj or .ply)" % ext)
