"""Read-only access to HDF5 files without h5py.

forcepho keeps its Sersic mixture tables (``sources.py:623-645``:
``nsersic``, ``rh``, ``amplitudes``, ``radii``) and its PSF stores
(``patches/storage.py:436-489``: ``<band>/parameters`` with the attribute
``n_psf_per_source``) in HDF5 files.  h5py is not part of this stack, so this
module reads the subset of the HDF5 file format (version 0/1 superblocks, as
written by h5py's default settings) that such files use:

* groups stored as symbol tables (version 1 B-trees of symbol-table nodes and a
  local heap), and groups stored as link messages;
* version 1 and version 2 object headers, with continuation blocks;
* datasets with compact, contiguous or chunked layout (version 1 B-tree chunk
  index), the deflate, shuffle and fletcher32 filters;
* fixed-point, floating-point, fixed-length string, compound and array
  datatypes, which map onto numpy dtypes;
* attributes with those datatypes (scalar or simple dataspaces).

The structures follow the public "HDF5 File Format Specification" (version 3.0).
Interface: ``File(path)`` behaves like a read-only ``h5py.File`` for these
uses: ``f["group/dataset"][:]``, ``.attrs``, ``.keys()``, ``.shape``,
``.dtype``, context manager.
"""

import mmap
import struct
import zlib

import numpy as np

__all__ = ["File", "Group", "Dataset", "HDF5FormatError"]

SIGNATURE = b"\x89HDF\r\n\x1a\n"
UNDEFINED = 0xFFFFFFFFFFFFFFFF


class HDF5FormatError(ValueError):
    """The file uses a part of the HDF5 format this reader does not handle."""


def _align8(n):
    return (n + 7) & ~7


class _Reader:
    """Byte access with the file's offset and length sizes."""

    def __init__(self, buf, base, so, sl):
        self.buf, self.base, self.so, self.sl = buf, base, so, sl

    def u(self, pos, n):
        return int.from_bytes(self.buf[pos:pos + n], "little")

    def off(self, pos):
        return self.u(pos, self.so)

    def length(self, pos):
        return self.u(pos, self.sl)

    def addr(self, a):
        """Absolute position of a file address."""
        return self.base + a

    def bytes(self, pos, n):
        return bytes(self.buf[pos:pos + n])

    def cstring(self, pos):
        end = self.buf.find(b"\x00", pos)
        return bytes(self.buf[pos:end]).decode("ascii", "replace")


# --------------------------------------------------------------------------
# Datatypes

def _parse_datatype(r, pos):
    """Datatype message at pos -> (numpy dtype, bytes consumed)."""
    cv = r.u(pos, 1)
    cls, version = cv & 0x0F, cv >> 4
    bits = r.u(pos + 1, 3)
    size = r.u(pos + 4, 4)
    p = pos + 8
    if cls == 0:      # fixed point
        order = ">" if bits & 1 else "<"
        signed = bool(bits & 0x08)
        dt = np.dtype("%s%s%d" % (order, "i" if signed else "u", size))
        return dt, 8 + 4
    if cls == 1:      # floating point
        order = ">" if bits & 1 else "<"
        if size not in (2, 4, 8):
            raise HDF5FormatError("float of %d bytes" % size)
        return np.dtype("%sf%d" % (order, size)), 8 + 12
    if cls == 3:      # fixed-length string
        return np.dtype("S%d" % size), 8
    if cls == 4:      # bit field
        order = ">" if bits & 1 else "<"
        return np.dtype("%su%d" % (order, size)), 8 + 4
    if cls == 6:      # compound
        n_members = bits & 0xFFFF
        names, formats, offsets = [], [], []
        for _ in range(n_members):
            name = r.cstring(p)
            if version < 3:
                p += _align8(len(name) + 1)
            else:
                p += len(name) + 1
            if version == 1:
                moff = r.u(p, 4)
                ndims = r.u(p + 4, 1)
                dims = [r.u(p + 16 + 4 * k, 4) for k in range(4)][:ndims]
                p += 32
            elif version == 2:
                moff = r.u(p, 4)
                dims = []
                p += 4
            else:
                nb = max(1, (size.bit_length() + 7) // 8)
                moff = r.u(p, nb)
                dims = []
                p += nb
            mdt, used = _parse_datatype(r, p)
            p += used
            if dims:
                mdt = np.dtype((mdt, tuple(dims)))
            names.append(name)
            formats.append(mdt)
            offsets.append(moff)
        dt = np.dtype({"names": names, "formats": formats, "offsets": offsets,
                       "itemsize": size})
        return dt, p - pos
    if cls == 10:     # array
        ndims = r.u(p, 1)
        if version < 3:
            dims = [r.u(p + 4 + 4 * k, 4) for k in range(ndims)]
            p += 4 + 4 * ndims + 4 * ndims   # dimension sizes then permutation
        else:
            dims = [r.u(p + 1 + 4 * k, 4) for k in range(ndims)]
            p += 1 + 4 * ndims
        base, used = _parse_datatype(r, p)
        p += used
        return np.dtype((base, tuple(dims))), p - pos
    raise HDF5FormatError("datatype class %d is not supported" % cls)


def _parse_dataspace(r, pos):
    """Dataspace message -> shape tuple (() for a scalar)."""
    version = r.u(pos, 1)
    ndims = r.u(pos + 1, 1)
    if version == 1:
        p = pos + 8
    elif version == 2:
        if r.u(pos + 3, 1) == 2:    # null dataspace
            return None
        p = pos + 4
    else:
        raise HDF5FormatError("dataspace version %d" % version)
    return tuple(r.length(p + r.sl * k) for k in range(ndims))


# --------------------------------------------------------------------------
# Object headers

class _Message:
    __slots__ = ("mtype", "pos", "size", "flags")

    def __init__(self, mtype, pos, size, flags):
        self.mtype, self.pos, self.size, self.flags = mtype, pos, size, flags


def _read_messages(r, addr):
    """All header messages of the object header at file address addr."""
    pos = r.addr(addr)
    msgs = []
    if r.bytes(pos, 4) == b"OHDR":
        version = r.u(pos + 4, 1)
        flags = r.u(pos + 5, 1)
        p = pos + 6
        if flags & 0x20:
            p += 16                     # access/modification/change/birth times
        if flags & 0x10:
            p += 4                      # attribute phase-change values
        size_bytes = 1 << (flags & 0x03)
        chunk0 = r.u(p, size_bytes)
        p += size_bytes
        blocks = [(p, chunk0)]
        track_order = bool(flags & 0x04)
        while blocks:
            start, length = blocks.pop(0)
            q, end = start, start + length - 4   # ends in a checksum
            if r.bytes(start - 4, 4) == b"OCHK":
                pass
            while q + 4 <= end:
                mtype = r.u(q, 1)
                msize = r.u(q + 1, 2)
                mflags = r.u(q + 3, 1)
                q += 4 + (2 if track_order else 0)
                if mtype == 0x10:
                    caddr, clen = r.off(q), r.length(q + r.so)
                    cpos = r.addr(caddr)
                    if r.bytes(cpos, 4) != b"OCHK":
                        raise HDF5FormatError("bad continuation block")
                    blocks.append((cpos + 4, clen - 4))
                elif mtype != 0:
                    msgs.append(_Message(mtype, q, msize, mflags))
                q += msize
        return msgs
    version = r.u(pos, 1)
    if version != 1:
        raise HDF5FormatError("object header version %d" % version)
    n_msgs = r.u(pos + 2, 2)
    hsize = r.u(pos + 8, 4)
    blocks = [(pos + 16, hsize)]
    while blocks and len(msgs) < n_msgs:
        start, length = blocks.pop(0)
        q, end = start, start + length
        while q + 8 <= end:
            mtype = r.u(q, 2)
            msize = r.u(q + 2, 2)
            mflags = r.u(q + 4, 1)
            q += 8
            if mtype == 0x10:
                blocks.append((r.addr(r.off(q)), r.length(q + r.so)))
            elif mtype != 0:
                msgs.append(_Message(mtype, q, msize, mflags))
            q += _align8(msize)
    return msgs


def _parse_attribute(r, m):
    """Attribute message -> (name, value)."""
    pos = m.pos
    version = r.u(pos, 1)
    name_size = r.u(pos + 2, 2)
    dt_size = r.u(pos + 4, 2)
    ds_size = r.u(pos + 6, 2)
    if version == 1:
        p = pos + 8
        name = r.bytes(p, name_size).rstrip(b"\x00").decode("ascii", "replace")
        p += _align8(name_size)
        dt, _ = _parse_datatype(r, p)
        p += _align8(dt_size)
        shape = _parse_dataspace(r, p)
        p += _align8(ds_size)
    elif version in (2, 3):
        p = pos + 8 + (1 if version == 3 else 0)
        name = r.bytes(p, name_size).rstrip(b"\x00").decode("utf-8", "replace")
        p += name_size
        dt, _ = _parse_datatype(r, p)
        p += dt_size
        shape = _parse_dataspace(r, p)
        p += ds_size
    else:
        raise HDF5FormatError("attribute version %d" % version)
    if shape is None:
        return name, None
    n = int(np.prod(shape)) if shape else 1
    raw = r.bytes(p, n * dt.itemsize)
    value = np.frombuffer(raw, dtype=dt, count=n).reshape(shape)
    value = value.astype(value.dtype.newbyteorder("="))
    if value.shape == ():
        value = value[()]
        if isinstance(value, bytes):
            value = value.rstrip(b"\x00").decode("utf-8", "replace")
    return name, value


# --------------------------------------------------------------------------
# Groups

def _symbol_table_links(r, btree_addr, heap_addr):
    """name -> object header address of a symbol-table group."""
    hpos = r.addr(heap_addr)
    if r.bytes(hpos, 4) != b"HEAP":
        raise HDF5FormatError("bad local heap")
    data_addr = r.off(hpos + 8 + 2 * r.sl)
    heap = r.addr(data_addr)
    links = {}
    stack = [btree_addr]
    entry = 2 * r.so + 8 + 16
    while stack:
        npos = r.addr(stack.pop())
        sig = r.bytes(npos, 4)
        if sig == b"SNOD":
            n = r.u(npos + 6, 2)
            for k in range(n):
                e = npos + 8 + k * entry
                name = r.cstring(heap + r.off(e))
                links[name] = r.off(e + r.so)
            continue
        if sig != b"TREE" or r.u(npos + 4, 1) != 0:
            raise HDF5FormatError("bad group B-tree node")
        used = r.u(npos + 6, 2)
        p = npos + 8 + 2 * r.so + r.sl         # after the first key
        for _ in range(used):
            stack.append(r.off(p))
            p += r.so + r.sl
    return links


def _link_message(r, pos):
    """Link message -> (name, object header address) for hard links."""
    flags = r.u(pos + 1, 1)
    p = pos + 2
    ltype = 0
    if flags & 0x08:
        ltype = r.u(p, 1)
        p += 1
    if flags & 0x04:
        p += 8
    if flags & 0x10:
        p += 1
    nlen_size = 1 << (flags & 0x03)
    nlen = r.u(p, nlen_size)
    p += nlen_size
    name = r.bytes(p, nlen).decode("utf-8", "replace")
    p += nlen
    if ltype != 0:
        return name, None
    return name, r.off(p)


class _Object:
    def __init__(self, f, addr, name):
        self._f, self._addr, self.name = f, addr, name
        self._msgs = _read_messages(f._r, addr)
        self._attrs = None

    def _first(self, mtype):
        for m in self._msgs:
            if m.mtype == mtype:
                return m
        return None

    @property
    def attrs(self):
        if self._attrs is None:
            r = self._f._r
            self._attrs = dict(_parse_attribute(r, m) for m in self._msgs if m.mtype == 0x0C)
        return self._attrs


class Group(_Object):
    """A group: mapping of names to Groups and Datasets."""

    def _links(self):
        r = self._f._r
        st = self._first(0x11)
        if st is not None:
            return _symbol_table_links(r, r.off(st.pos), r.off(st.pos + r.so))
        links = {}
        for m in self._msgs:
            if m.mtype == 0x06:
                name, addr = _link_message(r, m.pos)
                if addr is not None:
                    links[name] = addr
        if not links and self._first(0x02) is not None:
            raise HDF5FormatError("dense link storage is not supported")
        return links

    def keys(self):
        return list(self._links().keys())

    def __iter__(self):
        return iter(self.keys())

    def __contains__(self, name):
        try:
            self[name]
            return True
        except KeyError:
            return False

    def __getitem__(self, path):
        node = self
        for part in [p for p in path.split("/") if p]:
            if not isinstance(node, Group):
                raise KeyError(path)
            links = node._links()
            if part not in links:
                raise KeyError(path)
            node = self._f._open(links[part], node.name.rstrip("/") + "/" + part)
        return node


class Dataset(_Object):
    """A dataset: ``shape``, ``dtype``, ``attrs``; ``ds[...]`` reads it."""

    def __init__(self, f, addr, name):
        super().__init__(f, addr, name)
        r = f._r
        self.shape = _parse_dataspace(r, self._first(0x01).pos)
        self._dtype, _ = _parse_datatype(r, self._first(0x03).pos)
        self.dtype = self._dtype.newbyteorder("=") if self._dtype.fields is None else self._dtype

    @property
    def size(self):
        return int(np.prod(self.shape)) if self.shape else 1

    def __len__(self):
        return self.shape[0]

    def _filters(self):
        m = self._first(0x0B)
        if m is None:
            return []
        r = self._f._r
        pos = m.pos
        version = r.u(pos, 1)
        n = r.u(pos + 1, 1)
        p = pos + (8 if version == 1 else 2)
        out = []
        for _ in range(n):
            fid = r.u(p, 2)
            if version == 1 or fid >= 256:
                nlen = r.u(p + 2, 2)
                flags, nval = r.u(p + 4, 2), r.u(p + 6, 2)
                p += 8
            else:
                nlen = 0
                flags, nval = r.u(p + 2, 2), r.u(p + 4, 2)
                p += 6
            p += _align8(nlen) if version == 1 else nlen
            vals = [r.u(p + 4 * k, 4) for k in range(nval)]
            p += 4 * nval
            if version == 1 and nval % 2:
                p += 4
            out.append((fid, flags, vals))
        return out

    def _unfilter(self, raw, mask, filters):
        for k in reversed(range(len(filters))):
            if mask & (1 << k):
                continue
            fid, _, vals = filters[k]
            if fid == 1:
                raw = zlib.decompress(raw)
            elif fid == 2:
                item = vals[0] if vals else self._dtype.itemsize
                a = np.frombuffer(raw, dtype=np.uint8)
                n = len(a) // item
                body = a[:n * item].reshape(item, n).T.reshape(-1)
                raw = body.tobytes() + a[n * item:].tobytes()
            elif fid == 3:
                raw = raw[:-4]
            else:
                raise HDF5FormatError("filter %d is not supported" % fid)
        return raw

    def _read_all(self):
        r = self._f._r
        dt = self._dtype
        n = self.size
        lay = self._first(0x08)
        if lay is None:
            raise HDF5FormatError("dataset without a layout message")
        pos = lay.pos
        version = r.u(pos, 1)
        if version in (1, 2):
            ndims = r.u(pos + 1, 1)
            cls = r.u(pos + 2, 1)
            p = pos + 8
            if cls == 0:
                dims = [r.u(p + 4 * k, 4) for k in range(ndims)]
                p += 4 * ndims
                size = r.u(p, 4)
                raw = r.bytes(p + 4, size)
                return self._shape_out(raw, dt, n)
            addr = r.off(p)
            p += r.so
            dims = [r.u(p + 4 * k, 4) for k in range(ndims)]
            if cls == 1:
                return self._contiguous(addr, dt, n)
            return self._chunked(addr, dims[:-1], dt)
        if version in (3, 4):
            cls = r.u(pos + 1, 1)
            if cls == 0:
                size = r.u(pos + 2, 2)
                return self._shape_out(r.bytes(pos + 4, size), dt, n)
            if cls == 1:
                return self._contiguous(r.off(pos + 2), dt, n)
            if cls == 2 and version == 3:
                ndims = r.u(pos + 2, 1)
                addr = r.off(pos + 3)
                p = pos + 3 + r.so
                dims = [r.u(p + 4 * k, 4) for k in range(ndims)]
                return self._chunked(addr, dims[:-1], dt)
        raise HDF5FormatError("layout version %d is not supported" % version)

    def _shape_out(self, raw, dt, n):
        a = np.frombuffer(raw, dtype=dt, count=n).reshape(self.shape)
        if dt.fields is None:
            a = a.astype(dt.newbyteorder("="))
        else:
            a = a.copy()
        return a

    def _contiguous(self, addr, dt, n):
        r = self._f._r
        if addr == UNDEFINED:
            return np.zeros(self.shape, dtype=self.dtype)
        return self._shape_out(r.bytes(r.addr(addr), n * dt.itemsize), dt, n)

    def _chunked(self, btree, chunk, dt):
        r = self._f._r
        out = np.zeros(self.shape, dtype=dt)
        if btree == UNDEFINED:
            return self._shape_out(out.tobytes(), dt, self.size)
        filters = self._filters()
        ndims = len(chunk)
        key = 8 + 8 * (ndims + 1)
        stack = [btree]
        while stack:
            npos = r.addr(stack.pop())
            if r.bytes(npos, 4) != b"TREE" or r.u(npos + 4, 1) != 1:
                raise HDF5FormatError("bad chunk B-tree node")
            level = r.u(npos + 5, 1)
            used = r.u(npos + 6, 2)
            p = npos + 8 + 2 * r.so
            for _ in range(used):
                csize, mask = r.u(p, 4), r.u(p + 4, 4)
                origin = [r.u(p + 8 + 8 * k, 8) for k in range(ndims)]
                child = r.off(p + key)
                if level > 0:
                    stack.append(child)
                else:
                    raw = self._unfilter(r.bytes(r.addr(child), csize), mask, filters)
                    block = np.frombuffer(raw, dtype=dt, count=int(np.prod(chunk)))
                    block = block.reshape(chunk)
                    sl_out, sl_in = [], []
                    for o, c, s in zip(origin, chunk, self.shape):
                        m = min(c, s - o)
                        sl_out.append(slice(o, o + m))
                        sl_in.append(slice(0, m))
                    out[tuple(sl_out)] = block[tuple(sl_in)]
                p += key + r.so
        return self._shape_out(out.tobytes(), dt, self.size)

    def __getitem__(self, index):
        return self._read_all()[index]

    def __array__(self, dtype=None, copy=None):
        a = self._read_all()
        return a if dtype is None else a.astype(dtype)


class File(Group):
    """A read-only HDF5 file (``with File(path) as f: f["a/b"][:]``)."""

    def __init__(self, path, mode="r"):
        if mode != "r":
            raise ValueError("this reader opens files read-only")
        self._fh = open(path, "rb")
        self._buf = mmap.mmap(self._fh.fileno(), 0, access=mmap.ACCESS_READ)
        start = self._find_superblock()
        version = self._buf[start + 8]
        if version in (0, 1):
            so, sl = self._buf[start + 13], self._buf[start + 14]
            r = _Reader(self._buf, 0, so, sl)
            p = start + 24 + (4 if version == 1 else 0)
            base = r.off(p)
            root_entry = p + 4 * so
            root = r.off(root_entry + so)
        elif version in (2, 3):
            so, sl = self._buf[start + 9], self._buf[start + 10]
            r = _Reader(self._buf, 0, so, sl)
            base = r.off(start + 12)
            root = r.off(start + 12 + 3 * so)
        else:
            raise HDF5FormatError("superblock version %d" % version)
        self._r = _Reader(self._buf, base, so, sl)
        self.filename = path
        super().__init__(self, root, "/")

    def _find_superblock(self):
        for start in [0] + [512 << k for k in range(20)]:
            if start + 8 > len(self._buf):
                break
            if self._buf[start:start + 8] == SIGNATURE:
                return start
        raise HDF5FormatError("not an HDF5 file")

    def _open(self, addr, name):
        msgs = _read_messages(self._r, addr)
        kinds = {m.mtype for m in msgs}
        if 0x08 in kinds:
            return Dataset(self, addr, name)
        return Group(self, addr, name)

    def close(self):
        if self._buf is not None:
            self._buf.close()
            self._fh.close()
            self._buf = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
