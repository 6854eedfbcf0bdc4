"""Read CGNS files (the HDF5 flavour libcgns writes) without libcgns or h5py.

CGNS stores its tree in HDF5: every CGNS node is an HDF5 group carrying the
attributes ``name``, ``label``, ``type`` (``MT``, ``I4``, ``I8``, ``R4``,
``R8``, ``C1``) and ``flags``, and, when it holds data, a dataset named
``" data"`` in Fortran (column-major) order.  This module parses the part of
the HDF5 format that such files use: the version-2 superblock, version-2
object headers with continuation chunks, compact link storage, attribute
messages (versions 1-3) and compact / contiguous / chunked-but-unfiltered data
layouts.  Every checksum is verified.

It reads the solutions the reference writes (zone_FD/src/src/cgnsio.cpp:3-101:
a structured zone with vertex ``CoordinateX/Y/Z`` and a cell-centred
``flowSolution`` holding Density, XVelocity, [YVelocity, [ZVelocity,]]
Pressure) and the files the ``congfd`` driver writes in the same layout.
"""
import struct

import numpy as np

__all__ = ["read_tree", "read_solution", "lookup3", "Node", "FormatError"]


def lookup3(data, initval=0):
    """Bob Jenkins' lookup3 ``hashlittle`` (the HDF5 metadata checksum)."""
    def rot(x, k):
        return ((x << k) | (x >> (32 - k))) & 0xFFFFFFFF

    n = len(data)
    a = b = c = (0xDEADBEEF + n + initval) & 0xFFFFFFFF
    i = 0
    while n > 12:
        a = (a + struct.unpack_from("<I", data, i)[0]) & 0xFFFFFFFF
        b = (b + struct.unpack_from("<I", data, i + 4)[0]) & 0xFFFFFFFF
        c = (c + struct.unpack_from("<I", data, i + 8)[0]) & 0xFFFFFFFF
        a = (a - c) & 0xFFFFFFFF; a ^= rot(c, 4); c = (c + b) & 0xFFFFFFFF
        b = (b - a) & 0xFFFFFFFF; b ^= rot(a, 6); a = (a + c) & 0xFFFFFFFF
        c = (c - b) & 0xFFFFFFFF; c ^= rot(b, 8); b = (b + a) & 0xFFFFFFFF
        a = (a - c) & 0xFFFFFFFF; a ^= rot(c, 16); c = (c + b) & 0xFFFFFFFF
        b = (b - a) & 0xFFFFFFFF; b ^= rot(a, 19); a = (a + c) & 0xFFFFFFFF
        c = (c - b) & 0xFFFFFFFF; c ^= rot(b, 4); b = (b + a) & 0xFFFFFFFF
        n -= 12
        i += 12
    if n == 0:
        return c
    tail = bytes(data[i:i + n]) + b"\0" * (12 - n)
    ta, tb, tc = struct.unpack("<III", tail)
    a = (a + ta) & 0xFFFFFFFF
    b = (b + tb) & 0xFFFFFFFF
    c = (c + tc) & 0xFFFFFFFF
    c ^= b; c = (c - rot(b, 14)) & 0xFFFFFFFF
    a ^= c; a = (a - rot(c, 11)) & 0xFFFFFFFF
    b ^= a; b = (b - rot(a, 25)) & 0xFFFFFFFF
    c ^= b; c = (c - rot(b, 16)) & 0xFFFFFFFF
    a ^= c; a = (a - rot(c, 4)) & 0xFFFFFFFF
    b ^= a; b = (b - rot(a, 14)) & 0xFFFFFFFF
    c ^= b; c = (c - rot(b, 24)) & 0xFFFFFFFF
    return c


class FormatError(ValueError):
    pass


UNDEF = 0xFFFFFFFFFFFFFFFF


class _Obj:
    """One HDF5 object: its links (groups), attributes and data (datasets)."""

    def __init__(self):
        self.links = {}       # name -> address (insertion order kept)
        self.attrs = {}       # name -> numpy value / str
        self.dtype = None
        self.shape = None
        self.layout = None    # ("compact", bytes) | ("contiguous", addr, size) | ("chunked", ...)


class _H5:
    def __init__(self, buf):
        self.b = memoryview(buf)
        if bytes(self.b[:8]) != b"\x89HDF\r\n\x1a\n":
            raise FormatError("not an HDF5 file")
        ver = self.b[8]
        if ver not in (2, 3):
            raise FormatError(f"superblock version {ver} not supported")
        self.so, self.sl = self.b[9], self.b[10]
        if (self.so, self.sl) != (8, 8):
            raise FormatError("only 8-byte offsets and lengths are supported")
        base, ext, eof, root = struct.unpack_from("<QQQQ", self.b, 12)
        if lookup3(self.b[:44]) != struct.unpack_from("<I", self.b, 44)[0]:
            raise FormatError("superblock checksum mismatch")
        if base != 0:
            raise FormatError("non-zero base address")
        self.root = root
        self.cache = {}

    # ---- object headers ----------------------------------------------------
    def obj(self, addr):
        if addr in self.cache:
            return self.cache[addr]
        b = self.b
        if bytes(b[addr:addr + 4]) != b"OHDR":
            raise FormatError(f"no object header at {addr:#x}")
        if b[addr + 4] != 2:
            raise FormatError("object header version != 2")
        flags = b[addr + 5]
        p = addr + 6
        if flags & 0x20:
            p += 16
        if flags & 0x10:
            p += 4
        sz = 1 << (flags & 3)
        csize = int.from_bytes(b[p:p + sz], "little")
        p += sz
        o = _Obj()
        chunks = [(addr, p, p + csize)]
        while chunks:
            start, p0, end = chunks.pop(0)
            if lookup3(b[start:end]) != struct.unpack_from("<I", b, end)[0]:
                raise FormatError(f"object header checksum mismatch at {start:#x}")
            p = p0
            while p + 4 <= end:
                mtype, msize, mflags = b[p], struct.unpack_from("<H", b, p + 1)[0], b[p + 3]
                p += 4
                if flags & 0x04:
                    p += 2
                body = b[p:p + msize]
                p += msize
                if mtype == 0x10:  # continuation
                    caddr, clen = struct.unpack_from("<QQ", body, 0)
                    if bytes(b[caddr:caddr + 4]) != b"OCHK":
                        raise FormatError("bad continuation chunk")
                    chunks.append((caddr, caddr + 4, caddr + clen - 4))
                else:
                    self._message(o, mtype, mflags, body)
        self.cache[addr] = o
        return o

    def _message(self, o, mtype, mflags, m):
        if mflags & 0x02:
            raise FormatError("shared messages are not supported")
        if mtype == 0x06:
            self._link(o, m)
        elif mtype == 0x0C:
            name, val = self._attribute(m)
            o.attrs[name] = val
        elif mtype == 0x01:
            o.shape = self._dataspace(m)[0]
        elif mtype == 0x03:
            o.dtype = self._datatype(m)[0]
        elif mtype == 0x08:
            o.layout = self._layout(m)
        elif mtype == 0x0B:  # filter pipeline
            raise FormatError("filtered (compressed) datasets are not supported")
        elif mtype == 0x02:  # link info: dense storage is not supported
            flags = m[1]
            p = 2 + (8 if flags & 1 else 0)
            heap = struct.unpack_from("<Q", m, p)[0]
            if heap != UNDEF:
                raise FormatError("dense link storage is not supported")
        elif mtype == 0x15:  # attribute info
            flags = m[1]
            p = 2 + (2 if flags & 1 else 0)
            heap = struct.unpack_from("<Q", m, p)[0]
            if heap != UNDEF:
                raise FormatError("dense attribute storage is not supported")
        # others (NIL, fill value, group info, modification time, ...) carry nothing needed

    def _link(self, o, m):
        if m[0] != 1:
            raise FormatError("link message version != 1")
        flags = m[1]
        p = 2
        ltype = 0
        if flags & 0x08:
            ltype = m[p]
            p += 1
        if flags & 0x04:
            p += 8
        if flags & 0x10:
            p += 1
        sz = 1 << (flags & 3)
        nlen = int.from_bytes(m[p:p + sz], "little")
        p += sz
        name = bytes(m[p:p + nlen]).decode("utf-8")
        p += nlen
        if ltype != 0:
            raise FormatError("only hard links are supported")
        o.links[name] = struct.unpack_from("<Q", m, p)[0]

    def _datatype(self, m):
        cls, ver = m[0] & 0x0F, m[0] >> 4
        bits = m[1] | (m[2] << 8) | (m[3] << 16)
        size = struct.unpack_from("<I", m, 4)[0]
        if cls == 0:  # fixed point
            order = ">" if bits & 1 else "<"
            signed = "i" if bits & 8 else "u"
            return np.dtype(f"{order}{signed}{size}"), 12
        if cls == 1:  # floating point
            order = ">" if bits & 1 else "<"
            if size not in (4, 8):
                raise FormatError("unsupported float size")
            return np.dtype(f"{order}f{size}"), 20
        if cls == 3:  # string
            return np.dtype(f"S{size}"), 8
        raise FormatError(f"datatype class {cls} (version {ver}) is not supported")

    def _dataspace(self, m):
        ver, rank, flags = m[0], m[1], m[2]
        if ver == 1:
            p = 8
        elif ver == 2:
            p = 4
            if m[3] == 0:
                return (), p
            if m[3] == 2:
                return None, p
        else:
            raise FormatError("dataspace version")
        dims = tuple(struct.unpack_from(f"<{rank}Q", m, p)) if rank else ()
        p += 8 * rank * (2 if flags & 1 else 1)
        return dims, p

    def _attribute(self, m):
        ver = m[0]
        nsize, tsize, ssize = struct.unpack_from("<HHH", m, 2)
        if ver == 1:
            p = 8
            pad = lambda x: (x + 7) & ~7  # noqa: E731
        elif ver in (2, 3):
            p = 8 if ver == 2 else 9
            pad = lambda x: x  # noqa: E731
        else:
            raise FormatError("attribute version")
        name = bytes(m[p:p + nsize]).split(b"\0", 1)[0].decode("utf-8")
        p += pad(nsize)
        dt, _ = self._datatype(m[p:p + tsize])
        p += pad(tsize)
        shape, _ = self._dataspace(m[p:p + ssize])
        p += pad(ssize)
        count = int(np.prod(shape)) if shape else 1
        raw = bytes(m[p:p + count * dt.itemsize])
        val = np.frombuffer(raw, dtype=dt, count=count)
        if dt.kind == "S":
            return name, b"".join(val.tolist()).split(b"\0", 1)[0].decode("latin-1").rstrip()
        return name, val.reshape(shape) if shape else val[0]

    def _layout(self, m):
        ver = m[0]
        if ver not in (3, 4):
            raise FormatError("data layout version")
        cls = m[1]
        if cls == 0:
            size = struct.unpack_from("<H", m, 2)[0]
            return ("compact", bytes(m[4:4 + size]))
        if cls == 1:
            addr, size = struct.unpack_from("<QQ", m, 2)
            return ("contiguous", addr, size)
        raise FormatError("chunked / virtual layouts are not supported")

    def data(self, o):
        if o.shape is None or o.dtype is None or o.layout is None:
            return None
        count = int(np.prod(o.shape)) if o.shape else 1
        if o.layout[0] == "compact":
            raw = o.layout[1]
        else:
            _, addr, size = o.layout
            raw = bytes(self.b[addr:addr + size]) if addr != UNDEF else b"\0" * (count * o.dtype.itemsize)
        a = np.frombuffer(raw, dtype=o.dtype, count=count)
        return a.reshape(o.shape) if o.shape else a


class Node:
    """A CGNS node: name, label, data type, data (numpy, CGNS index order) and children."""

    def __init__(self, name, label, dtype, data, children):
        self.name, self.label, self.type, self.data, self.children = name, label, dtype, data, children

    def __getitem__(self, path):
        node = self
        for part in path.split("/"):
            node = next(c for c in node.children if c.name == part)
        return node

    def find(self, label):
        """Depth-first nodes with the given SIDS label."""
        out = [self] if self.label == label else []
        for c in self.children:
            out += c.find(label)
        return out

    def text(self):
        return bytes(np.asarray(self.data, dtype=np.int8).tobytes()).decode("latin-1") if self.data is not None else ""

    def __repr__(self):
        shape = None if self.data is None else self.data.shape
        return f"Node({self.name!r}, {self.label!r}, {self.type}, {shape}, {len(self.children)} children)"


def _node(h5, name, addr):
    o = h5.obj(addr)
    data = None
    if " data" in o.links:
        d = h5.obj(o.links[" data"])
        arr = h5.data(d)
        # HDF5 dims are C order; CGNS indexes Fortran order
        data = np.ascontiguousarray(arr.T) if arr is not None and arr.ndim > 1 else arr
    children = [_node(h5, k, a) for k, a in o.links.items() if not k.startswith(" ")]
    return Node(o.attrs.get("name", name), o.attrs.get("label", ""), o.attrs.get("type", ""), data, children)


def read_tree(path):
    """The whole CGNS tree of a file (the root node is the HDF5 root group)."""
    with open(path, "rb") as f:
        buf = f.read()
    h5 = _H5(buf)
    root = h5.obj(h5.root)
    children = [_node(h5, k, a) for k, a in root.links.items() if not k.startswith(" ")]
    top = Node("/", root.attrs.get("label", ""), root.attrs.get("type", ""), None, children)
    top.attrs = dict(root.attrs)
    top.hidden = {k: h5.data(h5.obj(a)) for k, a in root.links.items() if k.startswith(" ")}
    return top


def read_solution(path):
    """First zone of the first base: vertex coordinates and cell-centred
    fields, each as an array in (i, j, k) order, plus the zone size."""
    tree = read_tree(path)
    base = tree.find("CGNSBase_t")[0]
    zone = base.find("Zone_t")[0]
    out = {"base": base.name, "zone": zone.name, "cell_dim": int(base.data[0]), "phys_dim": int(base.data[1]),
           "zone_size": np.asarray(zone.data)}
    for g in zone.find("GridCoordinates_t"):
        for c in g.children:
            if c.label == "DataArray_t":
                out[c.name] = c.data
    sols = zone.find("FlowSolution_t")
    if sols:
        out["solution_name"] = sols[0].name
        loc = [c for c in sols[0].children if c.label == "GridLocation_t"]
        out["location"] = loc[0].text() if loc else "Vertex"
        for c in sols[0].children:
            if c.label == "DataArray_t":
                out[c.name] = c.data
    return out
