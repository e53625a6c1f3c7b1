"""A small HDF5 writer and reader for the files MaxwellFDTD exchanges: the solve! output
(src/outputs.jl:46-65,71-485, src/monitors.jl:72-78,362-373) and the TFSF record files
(src/sources.jl:264-317,387-466,592-710,779-877).

The reference writes them through HDF5.jl (libhdf5).  This module writes the classic
HDF5 format that libhdf5 produces by default -- superblock version 0, version-1
object headers, groups held in a symbol table (a version-1 B-tree, one symbol table
node and a local heap), contiguous datasets -- with the datatypes these files use:
IEEE little-endian floats (32/64 bit), little-endian integers (8-64 bit), and complex
numbers as the {r, i} compound HDF5.jl and h5py write.  Arrays are stored the way
HDF5.jl stores a Julia array: the dataset's dimensions are the array's dimensions
reversed, and the bytes are its column-major (Fortran) order.

The reader handles the same subset plus what libhdf5's defaults add to such files
(continuation blocks, version-2 dataspaces, compact layouts), enough to load record
files written by the reference or by this module.
"""
import struct

import numpy as np

UNDEF = 0xFFFFFFFFFFFFFFFF
SIG = b"\x89HDF\r\n\x1a\n"
LEAF_K = 32          # symbol table node: up to 2*LEAF_K = 64 links per group
INTERNAL_K = 16


def _pad8(b):
    return b + b"\0" * (-len(b) % 8)


# ---- datatypes ---------------------------------------------------------------------------
def _dtype_message(dt):
    """Datatype message (version 1) for a numpy dtype."""
    dt = np.dtype(dt)
    if dt.kind == "f":
        size = dt.itemsize
        bits = 8 * size
        sign, eloc, esize, mloc, msize, bias = ((63, 52, 11, 0, 52, 1023) if size == 8 else
                                                (31, 23, 8, 0, 23, 127))
        head = struct.pack("<BBBBI", (1 << 4) | 1, 0x20, sign, 0, size)
        return head + struct.pack("<HHBBBBI", 0, bits, eloc, esize, mloc, msize, bias)
    if dt.kind in "iub":
        size = dt.itemsize
        signed = 0x08 if dt.kind == "i" else 0
        head = struct.pack("<BBBBI", (1 << 4) | 0, signed, 0, 0, size)
        return head + struct.pack("<HH", 0, 8 * size)
    if dt.kind == "c":
        half = np.dtype("<f%d" % (dt.itemsize // 2))
        body = b""
        for k, name in enumerate(("r", "i")):
            body += _pad8(name.encode() + b"\0")
            body += struct.pack("<IB3xI4x4I", k * half.itemsize, 0, 0, 0, 0, 0, 0)
            body += _dtype_message(half)
        return struct.pack("<BBBBI", (1 << 4) | 6, 2, 0, 0, dt.itemsize) + body
    raise TypeError(f"no HDF5 datatype for {dt}")


def _parse_dtype(b, off):
    """(numpy dtype, bytes consumed) of a datatype message at b[off:]."""
    cv, f0, f1, f2, size = struct.unpack_from("<BBBBI", b, off)
    cls, ver = cv & 0x0F, cv >> 4
    if cls == 0:
        order = ">" if f0 & 1 else "<"
        return np.dtype(f"{order}{'i' if f0 & 0x08 else 'u'}{size}"), 12
    if cls == 1:
        order = ">" if f0 & 1 else "<"
        return np.dtype(f"{order}f{size}"), 20
    if cls == 6:
        nmemb = f0 | (f1 << 8)
        p = off + 8
        fields = []
        for _ in range(nmemb):
            end = b.index(b"\0", p)
            name = b[p:end].decode()
            if ver < 3:
                p += (end - p + 1 + 7) // 8 * 8
            else:
                p = end + 1
            if ver == 1:
                moff = struct.unpack_from("<I", b, p)[0]
                p += 28 + 4
            else:
                moff = struct.unpack_from("<I", b, p)[0]
                p += 4
            mdt, n = _parse_dtype(b, p)
            p += n
            fields.append((name, mdt, moff))
        names = [f[0] for f in fields]
        if names == ["r", "i"] and fields[0][1] == fields[1][1] and fields[0][1].kind == "f":
            return np.dtype(f"<c{2 * fields[0][1].itemsize}"), p - off
        return np.dtype({"names": names, "formats": [f[1] for f in fields],
                         "offsets": [f[2] for f in fields], "itemsize": size}), p - off
    if cls == 8:        # enumeration (HDF5.jl writes Bool this way): its base type
        base, n = _parse_dtype(b, off + 8)
        return base, 8 + n
    raise ValueError(f"unsupported HDF5 datatype class {cls}")


# ---- writer ------------------------------------------------------------------------------
class _Group:
    def __init__(self):
        self.items = {}


class H5Writer:
    """Writes an HDF5 file: dataset bytes go to the file as the datasets are created (or
    their slices written), groups and object headers when the writer is closed.

        w = H5Writer(path)
        w.create("x", array)                         # whole array at once
        w.create_dataset("fields/Ex", "f4", (Nx, Ny, Nz, Nt))
        w.write_last("fields/Ex", k, frame)          # [:, :, :, k] = frame
        w.close()

    Shapes are Julia shapes: a dataset is stored as HDF5.jl stores the Julia array of
    that shape (dimensions reversed, column-major bytes), so a slice along the last
    dimension is one contiguous block.  Missing parent groups are created."""

    def __init__(self, path):
        self.path = path
        self.root = _Group()
        self.f = open(path, "wb")
        self.f.write(b"\0" * 96)          # the superblock, written on close
        self.end = 96
        self.sets = {}

    def group(self, path):
        g = self.root
        for part in [p for p in path.split("/") if p]:
            g = g.items.setdefault(part, _Group())
            if not isinstance(g, _Group):
                raise ValueError(f"{path}: {part} is a dataset")
        return g

    def _alloc(self, b):
        off = self.end
        b = _pad8(b)
        self.f.seek(off)
        self.f.write(b)
        self.end += len(b)
        return off

    def create_dataset(self, path, dtype, shape):
        """An all-zero dataset of a Julia shape; fill it with write_last."""
        parts = [p for p in path.split("/") if p]
        g = self.group("/".join(parts[:-1]))
        dt = np.dtype(dtype)
        if dt == bool:
            dt = np.dtype(np.uint8)
        dt = dt.newbyteorder("<")
        shape = tuple(int(n) for n in shape)
        nbytes = int(np.prod(shape)) * dt.itemsize if shape else dt.itemsize
        addr = self._alloc(b"\0" * nbytes) if nbytes else UNDEF
        entry = (tuple(reversed(shape)), dt, addr, nbytes)
        g.items[parts[-1]] = entry
        self.sets[path.strip("/")] = (shape, dt, addr)
        return entry

    def write_last(self, path, k, values):
        """dataset[..., k] = values (0-based k along the last Julia dimension)."""
        shape, dt, addr = self.sets[path.strip("/")]
        frame = int(np.prod(shape[:-1])) * dt.itemsize
        v = np.asarray(values, dtype=dt)
        if v.size * dt.itemsize != frame:
            raise ValueError(f"{path}: slice of {v.size} values, expected {frame // dt.itemsize}")
        self.f.seek(addr + k * frame)
        self.f.write(v.ravel(order="F").tobytes())

    def create(self, path, value):
        """A dataset holding `value` (an array or scalar; bool as uint8)."""
        a = np.asarray(value)
        if a.dtype == bool:
            a = a.astype(np.uint8)
        if a.dtype.kind == "c":
            a = a.astype(np.complex128)
        a = a.astype(a.dtype.newbyteorder("<"))
        parts = [p for p in path.split("/") if p]
        g = self.group("/".join(parts[:-1]))
        data = a.ravel(order="F").tobytes()
        addr = self._alloc(data) if data else UNDEF
        g.items[parts[-1]] = (tuple(reversed(a.shape)), a.dtype, addr, len(data))

    def close(self):
        if self.f is None:
            return
        alloc = self._alloc

        def write_dataset(shape, dt, daddr, nbytes):
            msgs = []
            space = struct.pack("<BBBB4x", 1, len(shape), 0, 0)
            space += b"".join(struct.pack("<Q", n) for n in shape)
            msgs.append((0x0001, space))
            msgs.append((0x0003, _dtype_message(dt)))
            msgs.append((0x0005, struct.pack("<BBBB", 2, 2, 2, 0)))
            msgs.append((0x0008, struct.pack("<BBQQ", 3, 1, daddr, nbytes)))
            return alloc(_object_header(msgs))

        def write_group(g):
            names = sorted(g.items)
            if len(names) > 2 * LEAF_K:
                raise ValueError(f"a group holds at most {2 * LEAF_K} links here")
            children = {}
            for n in names:
                v = g.items[n]
                children[n] = write_group(v)[0] if isinstance(v, _Group) else write_dataset(*v)
            heap = bytearray(b"\0" * 8)          # offset 0: the empty name
            offs = {}
            for n in names:
                offs[n] = len(heap)
                heap.extend(_pad8(n.encode() + b"\0"))
            heap_data = alloc(bytes(heap))
            # free-list offset 1: no free block (offsets of blocks are multiples of 8)
            heap_addr = alloc(b"HEAP" + struct.pack("<B3xQQQ", 0, len(heap), 1, heap_data))
            entries = b"".join(struct.pack("<QQI4x16x", offs[n], children[n], 0) for n in names)
            entries += b"\0" * (40 * (2 * LEAF_K - len(names)))
            snod = alloc(b"SNOD" + struct.pack("<BBH", 1, 0, len(names)) + entries)
            keys = [0, offs[names[-1]] if names else 0]
            node = b"TREE" + struct.pack("<BBHQQ", 0, 0, 1, UNDEF, UNDEF)
            node += struct.pack("<QQQ", keys[0], snod, keys[1])
            node += b"\0" * (8 * (2 * (2 * INTERNAL_K) + 1) - 24)
            btree = alloc(node)
            oh = alloc(_object_header([(0x0011, struct.pack("<QQ", btree, heap_addr))]))
            return oh, btree, heap_addr

        root_oh, root_bt, root_heap = write_group(self.root)
        sb = SIG + struct.pack("<BBBBBBBBHHI", 0, 0, 0, 0, 0, 8, 8, 0, LEAF_K, INTERNAL_K, 0)
        sb += struct.pack("<QQQQ", 0, UNDEF, self.end, UNDEF)
        sb += struct.pack("<QQI4xQQ", 0, root_oh, 1, root_bt, root_heap)
        assert len(sb) == 96
        self.f.seek(0)
        self.f.write(sb)
        self.f.close()
        self.f = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def _object_header(msgs):
    """Version-1 object header holding the given (type, data) messages."""
    body = b""
    for mtype, data in msgs:
        data = _pad8(data)
        body += struct.pack("<HHB3x", mtype, len(data), 0) + data
    return struct.pack("<BBHII4x", 1, 0, len(msgs), 1, len(body)) + body


# ---- reader ------------------------------------------------------------------------------
class H5File:
    """Reads groups and datasets: f["fields/Ex"] returns the dataset as the Julia array
    HDF5.jl would (dimensions reversed back, column-major), f.keys("monitors") the
    names in a group, "name" in f whether a path exists."""

    def __init__(self, path):
        with open(path, "rb") as fh:
            self.b = fh.read()
        b = self.b
        base = b.find(SIG)
        if base != 0:
            raise ValueError(f"{path}: not an HDF5 file")
        ver = b[8]
        if ver in (0, 1):
            so, sl = b[13], b[14]
            if so != 8 or sl != 8:
                raise ValueError("only 8-byte offsets and lengths are supported")
            p = 24 + (4 if ver == 1 else 0) + 32
            self.root = self._entry_oh(p)
        elif ver in (2, 3):
            self.root = struct.unpack_from("<Q", b, 12 + 8 * 3)[0]
        else:
            raise ValueError(f"unsupported superblock version {ver}")

    def _entry_oh(self, p):
        return struct.unpack_from("<Q", self.b, p + 8)[0]

    def _messages(self, addr):
        """(type, data) messages of the object header at addr (v1 or v2)."""
        b = self.b
        out = []
        if b[addr:addr + 4] == b"OHDR":
            flags = b[addr + 5]
            p = addr + 6
            if flags & 0x20:
                p += 16
            if flags & 0x10:
                p += 4
            szn = 1 << (flags & 3)
            size = int.from_bytes(b[p:p + szn], "little")
            p += szn
            blocks = [(p, p + size)]
            track = 2 if flags & 0x04 else 0
            while blocks:
                s, e = blocks.pop(0)
                q = s
                while q + 4 + track <= e - 4 + 4:
                    mtype = b[q]
                    msize = struct.unpack_from("<H", b, q + 1)[0]
                    q += 4 + track
                    if mtype == 0x10:
                        a, ln = struct.unpack_from("<QQ", b, q)
                        blocks.append((a + 4, a + ln - 4))
                    elif mtype != 0:
                        out.append((mtype, b[q:q + msize]))
                    q += msize
                    if e - q < 4:
                        break
            return out
        nmsg, _, size = struct.unpack_from("<HII", b, addr + 2)
        blocks = [(addr + 16, addr + 16 + size)]
        while blocks and len(out) < nmsg + 64:
            s, e = blocks.pop(0)
            q = s
            while q + 8 <= e:
                mtype, msize = struct.unpack_from("<HH", b, q)
                data = b[q + 8:q + 8 + msize]
                if mtype == 0x10:
                    a, ln = struct.unpack_from("<QQ", data)
                    blocks.append((a, a + ln))
                elif mtype != 0:
                    out.append((mtype, data))
                q += 8 + msize
        return out

    def _group_links(self, addr):
        links = {}
        for mtype, data in self._messages(addr):
            if mtype == 0x11:
                btree, heap = struct.unpack_from("<QQ", data)
                self._walk_btree(btree, self._heap_data(heap), links)
            elif mtype == 0x06:
                self._link_message(data, links)
        return links

    def _heap_data(self, heap):
        b = self.b
        if b[heap:heap + 4] != b"HEAP":
            raise ValueError("bad local heap")
        size, _, daddr = struct.unpack_from("<QQQ", b, heap + 8)
        return b[daddr:daddr + size]

    def _walk_btree(self, addr, heap, links):
        b = self.b
        if b[addr:addr + 4] != b"TREE":
            raise ValueError("bad B-tree node")
        level = b[addr + 5]
        used = struct.unpack_from("<H", b, addr + 6)[0]
        p = addr + 24
        for k in range(used):
            child = struct.unpack_from("<Q", b, p + 8 + 16 * k)[0]
            if level > 0:
                self._walk_btree(child, heap, links)
            else:
                if b[child:child + 4] != b"SNOD":
                    raise ValueError("bad symbol table node")
                n = struct.unpack_from("<H", b, child + 6)[0]
                for j in range(n):
                    e = child + 8 + 40 * j
                    noff, oh = struct.unpack_from("<QQ", b, e)
                    name = heap[noff:heap.index(b"\0", noff)].decode()
                    links[name] = oh

    def _link_message(self, data, links):
        flags = data[1]
        p = 2
        ltype = 0
        if flags & 0x08:
            ltype = data[p]; p += 1
        if flags & 0x04:
            p += 8
        if flags & 0x10:
            p += 1
        ln = 1 << (flags & 3)
        nlen = int.from_bytes(data[p:p + ln], "little")
        p += ln
        name = data[p:p + nlen].decode()
        p += nlen
        if ltype == 0:
            links[name] = struct.unpack_from("<Q", data, p)[0]

    def _resolve(self, path):
        addr = self.root
        for part in [q for q in path.split("/") if q]:
            links = self._group_links(addr)
            if part not in links:
                raise KeyError(path)
            addr = links[part]
        return addr

    def __contains__(self, path):
        try:
            self._resolve(path)
            return True
        except KeyError:
            return False

    def keys(self, path="/"):
        return sorted(self._group_links(self._resolve(path)))

    def __getitem__(self, path):
        return self.read(path)

    def read(self, path, fortran=True):
        msgs = dict()
        for mtype, data in self._messages(self._resolve(path)):
            msgs.setdefault(mtype, data)
        if 0x0001 not in msgs:
            raise KeyError(f"{path} is not a dataset")
        sp = msgs[0x0001]
        ver, rank, flags = sp[0], sp[1], sp[2]
        p = 8 if ver == 1 else 4
        shape = tuple(struct.unpack_from("<Q", sp, p + 8 * k)[0] for k in range(rank))
        if ver >= 2 and sp[3] == 2:
            shape = (0,)
        dt, _ = _parse_dtype(msgs[0x0003], 0)
        lay = msgs[0x0008]
        count = int(np.prod(shape)) if shape else 1
        nbytes = count * dt.itemsize
        if lay[0] == 3 and lay[1] == 1:
            addr = struct.unpack_from("<Q", lay, 2)[0]
            raw = b"" if addr == UNDEF else self.b[addr:addr + nbytes]
        elif lay[0] == 3 and lay[1] == 0:
            n = struct.unpack_from("<H", lay, 2)[0]
            raw = lay[4:4 + n]
        else:
            raise ValueError(f"{path}: only contiguous and compact datasets are supported")
        if len(raw) < nbytes:
            raw = raw + b"\0" * (nbytes - len(raw))
        a = np.frombuffer(raw, dtype=dt, count=count)
        if dt.kind in "fiuc":
            a = a.astype(dt.newbyteorder("="))
        if not shape:
            return a[0]
        if fortran:
            return a.reshape(tuple(reversed(shape)), order="F")
        return a.reshape(shape)
