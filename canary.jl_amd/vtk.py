"""VTK output of nodal DG fields: the reference's `writemesh`
(examples/3d_kernels/vtk.jl:1-82, same file under 2d_kernels) as a
dependency-free writer of VTK XML UnstructuredGrid (.vtu) files.

Each element of order N is split into N^d linear sub-cells over its LGL
nodes (VTK_LINE / VTK_PIXEL / VTK_VOXEL, as the reference does), points are
the element nodes in the reference's column-major node order, and every field
is point data.  Data are stored uncompressed in the appended section (raw
little-endian, UInt64 block headers), as WriteVTK does with compress=false.
"""
import struct

import numpy as np

VTK_LINE, VTK_PIXEL, VTK_VOXEL = 3, 8, 11


def _cells(dims, nelem_list):
    """Connectivity (0-based) of the sub-cells of the listed elements."""
    d = len(dims)
    Nq = dims
    npe = int(np.prod(Nq))
    ind = np.arange(npe).reshape(Nq[::-1]).transpose()  # column-major node index
    local = []
    if d == 1:
        for i in range(Nq[0] - 1):
            local.append([ind[i], ind[i + 1]])
        ctype = VTK_LINE
    elif d == 2:
        for j in range(Nq[1] - 1):
            for i in range(Nq[0] - 1):
                local.append(ind[i:i + 2, j:j + 2].flatten(order="F"))
        ctype = VTK_PIXEL
    else:
        for k in range(Nq[2] - 1):
            for j in range(Nq[1] - 1):
                for i in range(Nq[0] - 1):
                    local.append(ind[i:i + 2, j:j + 2, k:k + 2].flatten(order="F"))
        ctype = VTK_VOXEL
    local = np.asarray(local, dtype=np.int64)
    conn = np.concatenate([local + e * npe for e in nelem_list]) if len(nelem_list) else \
        np.zeros((0, local.shape[1]), dtype=np.int64)
    return conn, ctype


def writemesh(base_name, *X, fields=(), realelems=None):
    """writemesh(base_name, x[, y[, z]]; fields=((name, array), ...),
    realelems=1:nelem): x, y, z and every field have shape (Nq, ..., nelem)
    (Fortran order, as the reference reshapes them).  Writes
    base_name + ".vtu" and returns the file name."""
    d = len(X)
    if d not in (1, 2, 3):
        raise ValueError("writemesh takes 1, 2 or 3 coordinate arrays")
    shape = X[0].shape
    if any(x.shape != shape for x in X) or len(shape) != d + 1:
        raise ValueError("coordinate arrays must share the shape (Nq, ..., nelem)")
    nelem = shape[-1]
    elems = list(range(nelem)) if realelems is None else [e - 1 for e in realelems]
    npts = int(np.prod(shape))
    pts = np.zeros((npts, 3), dtype=np.float64)
    for c in range(d):
        pts[:, c] = np.asarray(X[c], dtype=np.float64).reshape(-1, order="F")
    conn, ctype = _cells(shape[:-1], elems)
    ncell = conn.shape[0]
    offsets = np.arange(1, ncell + 1, dtype=np.int64) * conn.shape[1]
    types = np.full(ncell, ctype, dtype=np.uint8)

    blocks = []   # (xml header, bytes)
    off = 0

    def add(arr, name, ncomp, vtype):
        nonlocal off
        raw = np.ascontiguousarray(arr).tobytes()
        hdr = ('<DataArray type="%s" Name="%s" NumberOfComponents="%d" format="appended" '
               'offset="%d"/>' % (vtype, name, ncomp, off))
        data = struct.pack("<Q", len(raw)) + raw
        blocks.append(data)
        off += len(data)
        return hdr

    xml_pts = add(pts.astype("<f8"), "Points", 3, "Float64")
    xml_conn = add(conn.reshape(-1).astype("<i8"), "connectivity", 1, "Int64")
    xml_off = add(offsets.astype("<i8"), "offsets", 1, "Int64")
    xml_typ = add(types, "types", 1, "UInt8")
    xml_fields = []
    for name, v in fields:
        v = np.asarray(v, dtype=np.float64)
        if v.size != npts:
            raise ValueError("field %s has %d values for %d points" % (name, v.size, npts))
        xml_fields.append(add(v.reshape(-1, order="F").astype("<f8"), name, 1, "Float64"))

    head = ['<?xml version="1.0" encoding="utf-8"?>',
            '<VTKFile type="UnstructuredGrid" version="1.0" byte_order="LittleEndian" '
            'header_type="UInt64">',
            '<UnstructuredGrid>',
            '<Piece NumberOfPoints="%d" NumberOfCells="%d">' % (npts, ncell),
            '<Points>', xml_pts, '</Points>',
            '<Cells>', xml_conn, xml_off, xml_typ, '</Cells>',
            '<PointData>', *xml_fields, '</PointData>',
            '</Piece>', '</UnstructuredGrid>',
            '<AppendedData encoding="raw">']
    fname = base_name + ".vtu"
    with open(fname, "wb") as f:
        f.write(("\n".join(head) + "\n_").encode("utf-8"))
        for b in blocks:
            f.write(b)
        f.write(b"\n</AppendedData>\n</VTKFile>\n")
    return fname


def read_vtu(fname):
    """Read back a file written by writemesh: (points (npts,3), connectivity
    (ncell, nodes per cell), cell types, {field name: values})."""
    import re
    data = open(fname, "rb").read()
    start = data.index(b'<AppendedData encoding="raw">')
    body = data.index(b"_", start) + 1
    head = data[:start].decode("utf-8")
    arrays = {}
    for m in re.finditer(r'<DataArray type="(\w+)" Name="([^"]+)" NumberOfComponents="(\d+)" '
                         r'format="appended" offset="(\d+)"/>', head):
        vtype, name, ncomp, off = m.group(1), m.group(2), int(m.group(3)), int(m.group(4))
        n = struct.unpack("<Q", data[body + off:body + off + 8])[0]
        dt = {"Float64": "<f8", "Int64": "<i8", "UInt8": "u1"}[vtype]
        a = np.frombuffer(data[body + off + 8:body + off + 8 + n], dtype=dt)
        arrays[name] = a.reshape(-1, ncomp) if ncomp > 1 else a
    offs = arrays.pop("offsets")
    conn = arrays.pop("connectivity").reshape(len(offs), -1)
    return arrays.pop("Points"), conn, arrays.pop("types"), arrays


def plot_name(eqn_tag, dim, arrtype, rank, step):
    """The drivers' file name: viz/euler3D_set2c_<ArrType>_rank_0000_step_00010
    (euler3d_set2c.jl:879, euler2d_set2c.jl:763) or viz/swe2D_<ArrType>_rank_..
    (shallow_water.jl:738); eqn_tag is "euler" or "swe"."""
    if eqn_tag == "swe":
        return "viz/swe%dD_%s_rank_%04d_step_%05d" % (dim, arrtype, rank, step)
    return "viz/euler%dD_set2c_%s_rank_%04d_step_%05d" % (dim, arrtype, rank, step)


def write_state(base_name, s, Q, theta_ref=300.0, realelems=None):
    """The drivers' plot output on the element nodes.  Euler
    (euler3d_set2c.jl:869-882, euler2d_set2c.jl:753-766): convert set2c to
    set2nc (u = U/rho, ..., theta = E/rho), subtract theta_ref from E, and
    write rho, U, V, (W,) E.  Shallow water (shallow_water.jl:730-741): h, b
    and the velocities U/(h+b), V/(h+b).  realelems: 1-based element numbers
    whose sub-cells are written (the points of every element are)."""
    from ._lib import SWE
    Q = np.asarray(Q)
    dim, Nq = s.dim, s.N + 1
    shp = (Nq,) * dim + (s.nelem,)
    ix = 11 if dim == 3 else 6
    X = tuple(s.vgeo[:, ix + c, :].reshape(shp, order="F") for c in range(dim))
    if s.eqn == SWE:
        h, b = Q[:, 2, :], Q[:, 3, :]
        fields = [("h", h.reshape(shp, order="F")), ("b", b.reshape(shp, order="F")),
                  ("U", (Q[:, 0, :] / (h + b)).reshape(shp, order="F")),
                  ("V", (Q[:, 1, :] / (h + b)).reshape(shp, order="F"))]
        return writemesh(base_name, *X, fields=fields, realelems=realelems)
    rho = Q[:, dim, :]
    names = ["U", "V", "W"][:dim]
    fields = [("ρ", rho.reshape(shp, order="F"))]
    for c, nm in enumerate(names):
        fields.append((nm, (Q[:, c, :] / rho).reshape(shp, order="F")))
    fields.append(("E", (Q[:, dim + 1, :] / rho - theta_ref).reshape(shp, order="F")))
    return writemesh(base_name, *X, fields=fields, realelems=realelems)
