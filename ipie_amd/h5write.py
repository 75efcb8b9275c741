"""A small HDF5 writer (no h5py in this environment) for estimator output.

Writes a whole file from a nested dict: dict -> group, numpy array / number /
str / bytes -> dataset.  It follows the public HDF5 File Format
Specification (version 3.0) in its oldest layout, the one every HDF5 library
reads: superblock version 0, version-1 object headers, groups as symbol
tables (one version-1 B-tree node over symbol-table nodes, names in a local
heap), contiguous datasets with a version-1 dataspace, a version-2 fill-value
message and a version-3 layout message.  Element types: IEEE float64 / float32
(little endian), signed integers, fixed-length byte strings (null padded), and
complex128 as the compound {"r": float64, "i": float64} that h5py writes.

    write_file("estimates.0.h5", {"metadata": json_string,
                                  "block_size_1": {"data": {"000000000": arr}}})

``ipie_amd.h5read.H5File`` reads the result back.
"""

import struct

import numpy

UNDEF = 0xFFFFFFFFFFFFFFFF
LEAF_K = 32      # symbol-table nodes hold 2 LEAF_K names
INTERNAL_K = 16  # one B-tree node holds 2 INTERNAL_K symbol-table nodes
SNOD_CAP = 2 * LEAF_K
MAX_MEMBERS = SNOD_CAP * 2 * INTERNAL_K


def _pad8(b):
    return b + b"\0" * (-len(b) % 8)


def _dt_float(size):
    if size == 8:
        props = struct.pack("<HHBBBBI", 0, 64, 52, 11, 0, 52, 1023)
        return bytes([0x11, 0x20, 63, 0]) + struct.pack("<I", 8) + props
    if size == 4:
        props = struct.pack("<HHBBBBI", 0, 32, 23, 8, 0, 23, 127)
        return bytes([0x11, 0x20, 31, 0]) + struct.pack("<I", 4) + props
    raise NotImplementedError(f"float{8 * size}")


def _dt_int(size, signed):
    return bytes([0x10, 0x08 if signed else 0x00, 0, 0]) + struct.pack("<I", size) + struct.pack("<HH", 0, 8 * size)


def _dt_string(size):
    return bytes([0x13, 0x01, 0, 0]) + struct.pack("<I", size)  # null padded, ASCII


def _dt_complex128():
    out = bytes([0x16, 2, 0, 0]) + struct.pack("<I", 16)
    for name, off in (("r", 0), ("i", 8)):
        out += _pad8(name.encode() + b"\0")
        out += struct.pack("<IB3xI4x16x", off, 0, 0)
        out += _dt_float(8)
    return out


def _as_array(value):
    """(numpy array, datatype message) for a dataset value."""
    if isinstance(value, str):
        value = value.encode()
    if isinstance(value, bytes):
        a = numpy.array(value, dtype=f"S{max(len(value), 1)}")
        return a, _dt_string(a.dtype.itemsize)
    a = numpy.asarray(value)
    if a.dtype.kind == "U":
        a = numpy.char.encode(a, "ascii")
    if a.dtype.kind == "S":
        return a, _dt_string(max(a.dtype.itemsize, 1))
    if a.dtype.kind == "b":
        a = a.astype(numpy.int64)
    if a.dtype.kind == "c":
        a = a.astype(numpy.complex128)
        return a, _dt_complex128()
    if a.dtype.kind == "f":
        a = a.astype("<f8" if a.dtype.itemsize == 8 else "<f4")
        return a, _dt_float(a.dtype.itemsize)
    if a.dtype.kind in "iu":
        a = a.astype(f"<{a.dtype.kind}{a.dtype.itemsize}")
        return a, _dt_int(a.dtype.itemsize, a.dtype.kind == "i")
    raise NotImplementedError(f"dataset of dtype {a.dtype}")


class _Writer:
    def __init__(self):
        self.buf = bytearray(96)  # superblock, filled in last

    def alloc(self, nbytes):
        off = len(self.buf)
        self.buf += b"\0" * (nbytes + (-nbytes % 8))
        return off

    def put(self, off, data):
        self.buf[off:off + len(data)] = data

    @staticmethod
    def header(messages):
        """Version-1 object header: messages = [(type, data)]."""
        body = b""
        for mtype, data in messages:
            data = _pad8(data)
            body += struct.pack("<HHB3x", mtype, len(data), 0) + data
        return struct.pack("<BBHII4x", 1, 0, len(messages), 1, len(body)) + body

    def dataset(self, value):
        a, dt = _as_array(value)
        raw = numpy.ascontiguousarray(a).tobytes()
        space = struct.pack("<BBBB4x", 1, a.ndim, 0, 0) + b"".join(struct.pack("<Q", d) for d in a.shape)
        fill = struct.pack("<BBBB", 2, 2, 2, 0)
        # header first (its size is known before the data address)
        hdr_len = len(self.header([(1, space), (3, dt), (5, fill), (8, b"\0" * 18)]))
        hoff = self.alloc(hdr_len)
        doff = self.alloc(len(raw)) if raw else UNDEF
        if raw:
            self.put(doff, raw)
        layout = struct.pack("<BBQQ", 3, 1, doff, len(raw))
        self.put(hoff, self.header([(1, space), (3, dt), (5, fill), (8, layout)]))
        return hoff, None

    def group(self, members):
        """members: dict name -> value; returns (header address, (btree, heap))."""
        names = sorted(members)
        if len(names) > MAX_MEMBERS:
            raise NotImplementedError(f"groups of more than {MAX_MEMBERS} members")
        entries = []
        for name in names:
            v = members[name]
            if "/" in name or not name:
                raise ValueError(f"bad member name {name!r}")
            entries.append((name,) + (self.group(v) if isinstance(v, dict) else self.dataset(v)))
        # local heap: "" at offset 0, then the names
        heap_data = bytearray(8)
        name_off = {}
        for name in names:
            name_off[name] = len(heap_data)
            heap_data += _pad8(name.encode() + b"\0")
        heap_data += b"\0" * 16  # room at the end, unused
        hdata = self.alloc(len(heap_data))
        self.put(hdata, bytes(heap_data))
        heap = self.alloc(32)
        self.put(heap, b"HEAP" + bytes([0, 0, 0, 0]) + struct.pack("<QQQ", len(heap_data), UNDEF, hdata))
        # symbol-table nodes of up to SNOD_CAP names each
        chunks = [entries[i:i + SNOD_CAP] for i in range(0, len(entries), SNOD_CAP)] or [[]]
        snods = []
        for chunk in chunks:
            node = bytearray(b"SNOD" + bytes([1, 0]) + struct.pack("<H", len(chunk)))
            for name, haddr, cache in chunk:
                if cache is None:
                    node += struct.pack("<QQI4x16x", name_off[name], haddr, 0)
                else:
                    node += struct.pack("<QQI4xQQ", name_off[name], haddr, 1, cache[0], cache[1])
            node += b"\0" * (8 + SNOD_CAP * 40 - len(node))
            off = self.alloc(len(node))
            self.put(off, bytes(node))
            snods.append((off, name_off[chunk[-1][0]] if chunk else 0))
        # one level-0 B-tree node over the symbol-table nodes
        bt = bytearray(b"TREE" + bytes([0, 0]) + struct.pack("<H", len(snods) if entries else 0))
        bt += struct.pack("<QQ", UNDEF, UNDEF)
        keys = [0] + [k for _, k in snods]
        for i in range(2 * INTERNAL_K + 1):
            bt += struct.pack("<Q", keys[i] if i < len(keys) else 0)
            if i < 2 * INTERNAL_K:
                bt += struct.pack("<Q", snods[i][0] if i < len(snods) and entries else 0)
        btree = self.alloc(len(bt))
        self.put(btree, bytes(bt))
        hdr = self.header([(0x11, struct.pack("<QQ", btree, heap))])
        hoff = self.alloc(len(hdr))
        self.put(hoff, hdr)
        return hoff, (btree, heap)

    def finish(self, root):
        haddr, (btree, heap) = self.group(root)
        sb = b"\x89HDF\r\n\x1a\n" + bytes([0, 0, 0, 0, 0, 8, 8, 0]) + struct.pack("<HHI", LEAF_K, INTERNAL_K, 0)
        sb += struct.pack("<QQQQ", 0, UNDEF, len(self.buf), UNDEF)
        sb += struct.pack("<QQI4xQQ", 0, haddr, 1, btree, heap)
        assert len(sb) == 96
        self.put(0, sb)
        return bytes(self.buf)


def write_file(path, tree):
    """Write the nested dict `tree` as an HDF5 file (replacing `path`)."""
    data = _Writer().finish(tree)
    tmp = path + ".tmp"
    with open(tmp, "wb") as fh:
        fh.write(data)
    import os

    os.replace(tmp, path)
