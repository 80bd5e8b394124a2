"""Exodus II mesh ingest for the device mirror.

The reference reads its meshes through STK (src/mesh/meshIO.cpp:24) and then
walks the element buckets of each block and the side buckets of each side set
(src/mesh/meshGeometry.cpp:1277-1528, meshNodeGraph.cpp:54-330).  This module
flattens an Exodus II file into the arrays oagpu_mesh_create takes:

* coordinates [n_nodes][3] in the file's node order (Exodus ids - 1),
* element blocks concatenated in file order (Exodus global element numbering):
  conn [n_elems][8] (trailing entries -1 for smaller elements) and a
  topology code per element; the Exodus node orders of HEX8, TETRA4, WEDGE6
  and PYRAMID5 are the nalu master-element orders (Hex8CVFEM.C,
  Tet4CVFEM.C, Wed6CVFEM.C, Pyr5CVFEM.C),
* side sets as (parent element, side ordinal, boundary id): the Exodus side
  number minus one is the nalu side ordinal, the boundary id the side set's
  position in the file.

The file is read with a small reader of the NetCDF classic formats (CDF-1,
CDF-2 64-bit offset and CDF-5 64-bit data), which hold the Exodus meshes of
the reference's examples; NetCDF-4 / HDF5 files are rejected.
"""
from __future__ import annotations

import struct

import numpy as np

HEX8, TET4, WED6, PYR5 = 0, 1, 2, 3
_TOPOLOGY = {"HEX": HEX8, "HEX8": HEX8, "HEXAHEDRON": HEX8, "TETRA": TET4, "TETRA4": TET4,
             "TET4": TET4, "TET": TET4, "WEDGE": WED6, "WEDGE6": WED6, "PYRAMID": PYR5,
             "PYRAMID5": PYR5}
_NPE = {HEX8: 8, TET4: 4, WED6: 6, PYR5: 5}

# NetCDF classic type codes -> (numpy big-endian dtype)
_NC_TYPES = {1: ">i1", 2: "S1", 3: ">i2", 4: ">i4", 5: ">f4", 6: ">f8", 7: ">u1", 8: ">u2",
             9: ">u4", 10: ">i8", 11: ">u8"}
_NC_DIMENSION, _NC_VARIABLE, _NC_ATTRIBUTE = 0x0A, 0x0B, 0x0C


class NetCDFClassic:
    """Header and non-record variables of a NetCDF classic file."""

    def __init__(self, path: str):
        with open(path, "rb") as fh:
            self._buf = fh.read()
        b = self._buf
        if b[:4] == b"\x89HDF":
            raise ValueError(f"{path}: NetCDF-4 / HDF5 file; only the classic formats are read")
        if b[:3] != b"CDF" or b[3] not in (1, 2, 5):
            raise ValueError(f"{path}: not a NetCDF classic file")
        self.version = b[3]
        self._pos = 4
        self._size_fmt = ">q" if self.version == 5 else ">i"     # NON_NEG counts
        self.numrecs = self._nonneg()
        self.dims = []          # (name, length), length 0: the record dimension
        self.attrs = {}
        self.vars = {}          # name -> (dimids, attrs, nc_type, vsize, begin)
        self._dim_list()
        self.attrs = self._att_list()
        self._var_list()

    # ---- primitive readers
    def _take(self, fmt):
        v = struct.unpack_from(fmt, self._buf, self._pos)[0]
        self._pos += struct.calcsize(fmt)
        return v

    def _nonneg(self):
        return self._take(self._size_fmt)

    def _tag(self):
        return self._take(">i")

    def _name(self):
        n = self._nonneg()
        s = self._buf[self._pos:self._pos + n].decode("ascii", "replace")
        self._pos += (n + 3) // 4 * 4
        return s

    def _values(self, nc_type, n):
        dt = np.dtype(_NC_TYPES[nc_type])
        nbytes = dt.itemsize * n
        raw = self._buf[self._pos:self._pos + nbytes]
        self._pos += (nbytes + 3) // 4 * 4
        if nc_type == 2:
            return raw.rstrip(b"\0").decode("ascii", "replace")
        return np.frombuffer(raw, dtype=dt).astype(dt.newbyteorder("="))

    def _absent_or(self, tag_expected):
        tag = self._tag()
        n = self._nonneg()
        if tag == 0:
            if n != 0:
                raise ValueError("malformed NetCDF header (ABSENT with a count)")
            return 0
        if tag != tag_expected:
            raise ValueError(f"malformed NetCDF header (tag {tag:#x})")
        return n

    # ---- header lists
    def _dim_list(self):
        for _ in range(self._absent_or(_NC_DIMENSION)):
            name = self._name()
            self.dims.append((name, self._nonneg()))

    def _att_list(self):
        out = {}
        for _ in range(self._absent_or(_NC_ATTRIBUTE)):
            name = self._name()
            t = self._tag()
            out[name] = self._values(t, self._nonneg())
        return out

    def _var_list(self):
        for _ in range(self._absent_or(_NC_VARIABLE)):
            name = self._name()
            dimids = [self._nonneg() for _ in range(self._nonneg())]
            attrs = self._att_list()
            t = self._tag()
            vsize = self._nonneg()
            begin = self._take(">i" if self.version == 1 else ">q")
            self.vars[name] = (dimids, attrs, t, vsize, begin)

    # ---- data
    def dim(self, name, default=None):
        for n, v in self.dims:
            if n == name:
                return v
        if default is not None:
            return default
        raise KeyError(name)

    def has(self, name):
        return name in self.vars

    def var_attrs(self, name):
        return self.vars[name][1]

    def read(self, name):
        dimids, _, t, _, begin = self.vars[name]
        shape = [self.dims[d][1] for d in dimids]
        if any(self.dims[d][1] == 0 for d in dimids):
            raise ValueError(f"{name}: record variables are not read")
        dt = np.dtype(_NC_TYPES[t])
        n = int(np.prod(shape)) if shape else 1
        a = np.frombuffer(self._buf, dtype=dt, count=n, offset=begin)
        a = a.astype(dt.newbyteorder("=")).reshape(shape)
        return a


def _names(nc, var, count):
    if not nc.has(var):
        return [""] * count
    raw = nc.read(var)
    return [bytes(r.tobytes()).split(b"\0", 1)[0].decode("ascii", "replace").strip() for r in raw]


def read_exodus(path: str) -> dict:
    """Mesh dict of an Exodus II file (the keys of meshgen's generators):
    xyz [nn][3], conn [ne][8] int32 (-1 padded), etype [ne], bf_elem, bf_ord,
    bf_bc (boundary id = side-set position), plus block_names, block_ids,
    block_of_elem [ne], sideset_names, sideset_ids."""
    nc = NetCDFClassic(path)
    nn = nc.dim("num_nodes")
    ndim = nc.dim("num_dim")
    if ndim != 3:
        raise ValueError(f"{path}: {ndim}-dimensional mesh; only 3-D meshes are supported")
    if nc.has("coordx"):
        xyz = np.stack([nc.read("coordx"), nc.read("coordy"), nc.read("coordz")], axis=1)
    else:
        xyz = nc.read("coord").T
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    assert xyz.shape == (nn, 3)
    nblk = nc.dim("num_el_blk")
    blk_names = _names(nc, "eb_names", nblk)
    blk_ids = nc.read("eb_prop1").astype(np.int64).tolist() if nc.has("eb_prop1") else list(range(1, nblk + 1))
    conns, types, owner = [], [], []
    for b in range(1, nblk + 1):
        cname = f"connect{b}"
        if not nc.has(cname):                     # empty block
            continue
        et = nc.var_attrs(cname).get("elem_type", "")
        et = et if isinstance(et, str) else ""
        key = et.strip().upper()
        if key not in _TOPOLOGY:
            raise ValueError(f"{path}: block {b} has element type {et!r}; hex8, tet4, wedge6 and "
                             f"pyramid5 are supported")
        t = _TOPOLOGY[key]
        c = nc.read(cname).astype(np.int64) - 1
        if c.shape[1] != _NPE[t]:
            raise ValueError(f"{path}: block {b} ({et}) has {c.shape[1]} nodes per element")
        full = np.full((c.shape[0], 8), -1, dtype=np.int64)
        full[:, :c.shape[1]] = c
        conns.append(full)
        types.append(np.full(c.shape[0], t, dtype=np.int32))
        owner.append(np.full(c.shape[0], b - 1, dtype=np.int32))
    conn = np.concatenate(conns).astype(np.int32)
    etype = np.concatenate(types)
    if conn.min(initial=0) < -1 or conn.max(initial=0) >= nn:
        raise ValueError(f"{path}: connectivity out of range")
    nss = nc.dim("num_side_sets", 0)
    ss_names = _names(nc, "ss_names", nss)
    ss_ids = nc.read("ss_prop1").astype(np.int64).tolist() if nss and nc.has("ss_prop1") else list(range(1, nss + 1))
    bfe, bfo, bfb = [], [], []
    for s in range(1, nss + 1):
        if not nc.has(f"elem_ss{s}"):
            continue
        e = nc.read(f"elem_ss{s}").astype(np.int64) - 1
        o = nc.read(f"side_ss{s}").astype(np.int64) - 1
        bfe.append(e)
        bfo.append(o)
        bfb.append(np.full(e.shape, s - 1, dtype=np.int64))
    if bfe:
        bf_elem, bf_ord, bf_bc = (np.concatenate(v).astype(np.int32) for v in (bfe, bfo, bfb))
    else:
        bf_elem = bf_ord = bf_bc = np.zeros(0, dtype=np.int32)
    if bf_elem.size and (bf_elem.min() < 0 or bf_elem.max() >= conn.shape[0]):
        raise ValueError(f"{path}: side-set element out of range")
    return {
        "xyz": xyz,
        "conn": np.ascontiguousarray(conn),
        "etype": np.ascontiguousarray(etype),
        "bf_elem": np.ascontiguousarray(bf_elem),
        "bf_ord": np.ascontiguousarray(bf_ord),
        "bf_bc": np.ascontiguousarray(bf_bc),
        "block_names": blk_names,
        "block_ids": blk_ids,
        "block_of_elem": np.concatenate(owner),
        "sideset_names": ss_names,
        "sideset_ids": ss_ids,
    }


def nearest_node(xyz: np.ndarray, point) -> int:
    """Node closest to a point: the pressure-level node of
    pressure_level_information option cartesian_coordinates."""
    d = np.einsum("ij,ij->i", xyz - np.asarray(point), xyz - np.asarray(point))
    return int(np.argmin(d))
