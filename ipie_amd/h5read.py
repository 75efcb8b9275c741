"""A small read-only HDF5 reader (no h5py in this environment).

Covers what the integral and wavefunction files ipie reads use: superblock
versions 0 and 1, groups as symbol tables (v1 B-trees of symbol nodes with a
local heap) and as link messages, object headers version 1 and 2, datasets
with contiguous or compact layout (chunked storage without filters too), and
element types IEEE float, integer, fixed-length string and compound (read as
numpy structured types; complex pairs as written by h5py become complex128).  It follows the
public HDF5 File Format Specification (version 3.0); everything it does not
cover raises NotImplementedError.  Nothing in a file is executed.

    with H5File(path) as f:
        hcore = f["Hamiltonian/hcore"]          # numpy array
        names = f.keys("Hamiltonian")
"""

import struct

import numpy

_SIG = b"\x89HDF\r\n\x1a\n"
UNDEF = 0xFFFFFFFFFFFFFFFF


class H5File:
    def __init__(self, path):
        with open(path, "rb") as fh:
            self.buf = fh.read()
        b = self.buf
        base = b.find(_SIG)
        if base != 0:
            raise ValueError(f"{path}: not an HDF5 file (superblock at offset 0 expected)")
        ver = b[8]
        if ver in (0, 1):
            self.so, self.sl = b[13], b[14]
            p = 24 if ver == 0 else 28
            p += 4 * self.so  # base, free-space, EOF, driver addresses
            # root group symbol table entry
            self.root = self._symbol_entry(p)["header"]
        elif ver in (2, 3):
            self.so, self.sl = b[9], b[10]
            p = 12 + 3 * self.so
            self.root = self._off(p)
        else:
            raise NotImplementedError(f"HDF5 superblock version {ver}")

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.buf = None

    # ------------------------------------------------------------------ primitives
    def _u(self, p, n):
        return int.from_bytes(self.buf[p:p + n], "little")

    def _off(self, p):
        return self._u(p, self.so)

    def _symbol_entry(self, p):
        return {"name": self._off(p), "header": self._off(p + self.so), "cache": self._u(p + 2 * self.so, 4),
                "scratch": p + 2 * self.so + 8}

    def _heap_data(self, heap):
        b = self.buf
        if b[heap:heap + 4] != b"HEAP":
            raise ValueError("local heap signature")
        p = heap + 8 + 2 * self.sl
        return self._off(p)

    def _cstr(self, p):
        e = self.buf.index(b"\0", p)
        return self.buf[p:e].decode()

    # ------------------------------------------------------------------ object headers
    def _messages(self, addr):
        b = self.buf
        out = []
        if b[addr:addr + 4] == b"OHDR":
            ver = b[addr + 4]
            flags = b[addr + 5]
            p = addr + 6
            if flags & 0x20:
                p += 16
            if flags & 0x10:
                p += 4
            nsz = 1 << (flags & 3)
            size = self._u(p, nsz)
            p += nsz
            blocks = [(p, p + size)]
            while blocks:
                s, e = blocks.pop(0)
                q = s
                while q + 4 <= e:
                    mt = b[q]
                    ms = self._u(q + 1, 2)
                    q += 4
                    if flags & 0x04:
                        q += 2
                    if mt == 0x10:  # continuation
                        ca, cl = self._off(q), self._u(q + self.so, self.sl)
                        blocks.append((ca + 4, ca + cl - 4))  # OCHK signature ... checksum
                    elif mt:
                        out.append((mt, q, ms))
                    q += ms
            del ver
            return out
        ver = b[addr]
        if ver != 1:
            raise NotImplementedError(f"object header version {ver}")
        nmsg = self._u(addr + 2, 2)
        size = self._u(addr + 8, 4)
        blocks = [(addr + 16, addr + 16 + size)]
        while blocks and len(out) < nmsg + 64:
            s, e = blocks.pop(0)
            q = s
            while q + 8 <= e:
                mt = self._u(q, 2)
                ms = self._u(q + 2, 2)
                q += 8
                if mt == 0x10:
                    ca, cl = self._off(q), self._u(q + self.so, self.sl)
                    blocks.append((ca, ca + cl))
                elif mt:
                    out.append((mt, q, ms))
                q += ms
        return out

    # ------------------------------------------------------------------ groups
    def _group_links(self, addr):
        msgs = self._messages(addr)
        links = {}
        for mt, p, ms in msgs:
            if mt == 0x11:  # symbol table: v1 B-tree + local heap
                btree, heap = self._off(p), self._off(p + self.so)
                self._walk_btree(btree, self._heap_data(heap), links)
            elif mt == 0x06:  # link message
                self._link_message(p, links)
            elif mt == 0x02:  # link info: dense (fractal heap) storage
                fheap = self._off(p + 2 + (8 if self.buf[p + 1] & 1 else 0))
                if fheap != UNDEF:
                    raise NotImplementedError("groups stored in fractal heaps")
        return links

    def _walk_btree(self, node, heap_data, links):
        b = self.buf
        if b[node:node + 4] != b"TREE":
            raise ValueError("B-tree signature")
        level = b[node + 5]
        used = self._u(node + 6, 2)
        p = node + 8 + 2 * self.so
        p += self.sl  # key 0
        for _ in range(used):
            child = self._off(p)
            p += self.so + self.sl
            if level > 0:
                self._walk_btree(child, heap_data, links)
            else:
                self._symbol_node(child, heap_data, links)

    def _symbol_node(self, addr, heap_data, links):
        b = self.buf
        if b[addr:addr + 4] != b"SNOD":
            raise ValueError("symbol node signature")
        n = self._u(addr + 6, 2)
        p = addr + 8
        esz = 2 * self.so + 24
        for _ in range(n):
            e = self._symbol_entry(p)
            links[self._cstr(heap_data + e["name"])] = e["header"]
            p += esz

    def _link_message(self, p, links):
        b = self.buf
        flags = b[p + 1]
        q = p + 2
        ltype = 0
        if flags & 0x08:
            ltype = b[q]
            q += 1
        if flags & 0x04:
            q += 8
        if flags & 0x10:
            q += 1
        lsz = 1 << (flags & 3)
        nlen = self._u(q, lsz)
        q += lsz
        name = b[q:q + nlen].decode()
        q += nlen
        if ltype != 0:
            return  # soft / external links are not followed
        links[name] = self._off(q)

    def _resolve(self, path):
        addr = self.root
        for part in [s for s in path.split("/") if s]:
            links = self._group_links(addr)
            if part not in links:
                raise KeyError(path)
            addr = links[part]
        return addr

    def keys(self, path="/"):
        return sorted(self._group_links(self._resolve(path)))

    def __contains__(self, path):
        try:
            self._resolve(path)
            return True
        except KeyError:
            return False

    # ------------------------------------------------------------------ datasets
    def _dtype(self, p):
        b = self.buf
        cls = b[p] & 0x0F
        bits = b[p + 1] | (b[p + 2] << 8) | (b[p + 3] << 16)
        size = self._u(p + 4, 4)
        order = ">" if bits & 1 else "<"
        if cls == 0:
            return numpy.dtype(f"{order}{'i' if bits & 0x08 else 'u'}{size}")
        if cls == 1:
            return numpy.dtype(f"{order}f{size}")
        if cls == 3:  # fixed-length string
            return numpy.dtype(f"S{size}")
        if cls == 6:
            ver = b[p] >> 4
            nmem = bits & 0xFFFF
            q = p + 8
            names, formats, offsets = [], [], []
            for _ in range(nmem):
                e = b.index(b"\0", q)
                name = b[q:e].decode()
                if ver < 3:
                    q = q + ((e - q) // 8 + 1) * 8
                else:
                    q = e + 1
                if ver == 1:
                    off = self._u(q, 4)
                    q += 4 + 1 + 3 + 4 + 4 + 16
                elif ver == 2:
                    off = self._u(q, 4)
                    q += 4
                else:
                    osz = max(1, (size.bit_length() + 7) // 8)
                    off = self._u(q, osz)
                    q += osz
                mdt, mlen = self._dtype_len(q)
                names.append(name)
                formats.append(mdt)
                offsets.append(off)
                q += mlen
            return numpy.dtype({"names": names, "formats": formats, "offsets": offsets, "itemsize": size})
        raise NotImplementedError(f"HDF5 datatype class {cls}")

    def _dtype_len(self, p):
        dt = self._dtype(p)
        cls = self.buf[p] & 0x0F
        if cls == 0:
            return dt, 8 + 4
        if cls == 1:
            return dt, 8 + 12
        if cls == 3:
            return dt, 8
        raise NotImplementedError("nested compound members")

    def _read_dataset(self, addr):
        b = self.buf
        shape, dtype, layout, filters = None, None, None, False
        for mt, p, ms in self._messages(addr):
            if mt == 0x01:  # dataspace
                ver, nd, fl = b[p], b[p + 1], b[p + 2]
                q = p + (8 if ver == 1 else 4)
                shape = tuple(self._u(q + self.sl * i, self.sl) for i in range(nd))
            elif mt == 0x03:
                dtype = self._dtype(p)
            elif mt == 0x08:
                layout = p
            elif mt == 0x0B:
                filters = True
        if shape is None or dtype is None or layout is None:
            raise KeyError("not a dataset")
        n = int(numpy.prod(shape)) if shape else 1
        nbytes = n * dtype.itemsize
        ver, cls = b[layout], b[layout + 1]
        if ver < 3:
            raise NotImplementedError(f"data layout message version {ver}")
        if cls == 0:  # compact
            sz = self._u(layout + 2, 2)
            raw = b[layout + 4:layout + 4 + sz]
        elif cls == 1:  # contiguous
            a = self._off(layout + 2)
            raw = b"\0" * nbytes if a == UNDEF else b[a:a + nbytes]
        elif cls == 2:
            if filters:
                raise NotImplementedError("filtered (compressed) chunked datasets")
            raw = self._read_chunked(layout, shape, dtype)
        else:
            raise NotImplementedError(f"data layout class {cls}")
        arr = numpy.frombuffer(raw[:nbytes], dtype=dtype).reshape(shape)
        return _complexify(arr.copy())

    def _read_chunked(self, layout, shape, dtype):
        b = self.buf
        nd = b[layout + 2] - 1
        btree = self._off(layout + 3)
        q = layout + 3 + self.so
        cdims = [self._u(q + 4 * i, 4) for i in range(nd)]
        out = numpy.zeros(shape, dtype=dtype)
        self._walk_chunks(btree, nd, cdims, out)
        return out.tobytes()

    def _walk_chunks(self, node, nd, cdims, out):
        b = self.buf
        if b[node:node + 4] != b"TREE" or b[node + 4] != 1:
            raise ValueError("chunk B-tree")
        level = b[node + 5]
        used = self._u(node + 6, 2)
        ksz = 8 + 8 * (nd + 1)
        p = node + 8 + 2 * self.so
        for _ in range(used):
            csize = self._u(p, 4)
            offs = [self._u(p + 8 + 8 * i, 8) for i in range(nd)]
            child = self._off(p + ksz)
            if level > 0:
                self._walk_chunks(child, nd, cdims, out)
            else:
                chunk = numpy.frombuffer(b[child:child + csize], dtype=out.dtype)
                chunk = chunk[:int(numpy.prod(cdims))].reshape(cdims)
                sl = tuple(slice(o, min(o + c, s)) for o, c, s in zip(offs, cdims, out.shape))
                out[sl] = chunk[tuple(slice(0, s.stop - s.start) for s in sl)]
            p += ksz + self.so

    def __getitem__(self, path):
        return self._read_dataset(self._resolve(path))


def _complexify(arr):
    """Compound (r, i) pairs as written by h5py for complex numbers -> complex."""
    names = arr.dtype.names
    if names and len(names) == 2 and set(n.lower() for n in names) == {"r", "i"}:
        rn = [n for n in names if n.lower() == "r"][0]
        im = [n for n in names if n.lower() == "i"][0]
        return arr[rn].astype(numpy.float64) + 1j * arr[im].astype(numpy.float64)
    return arr
