"""Minimal HDF5 reader for the files ABLATE's Hdf5Serializer writes through PETSc (src/io/hdf5Serializer.cpp,
PetscViewerHDF5): superblock version 0 or 1 with 8-byte addresses, version-1 object headers, old-style groups
(symbol tables: v1 B-tree + local heap + symbol nodes), datasets with contiguous or chunked (v1 B-tree) storage and
no filters, fixed-point (int32 / int64) and IEEE floating-point (float32 / float64) element types.  That is the
subset those files use; anything else raises ``NotImplementedError`` naming what was met.

No HDF5 library is available in this image, so the format is read directly from the published HDF5 File Format
Specification (version 3.0: superblock, object header, B-tree, local heap, dataspace, datatype, data layout and
attribute messages).
"""
from __future__ import annotations

import struct

import numpy as np

_SIG = b"\x89HDF\r\n\x1a\n"


class H5File:
    def __init__(self, path):
        self.path = str(path)
        with open(path, "rb") as f:
            self.d = f.read()
        d = self.d
        if d[:8] != _SIG:
            raise ValueError(f"{path}: not an HDF5 file")
        ver = d[8]
        if ver not in (0, 1):
            raise NotImplementedError(f"{path}: superblock version {ver} (only 0 and 1 are read)")
        if d[13] != 8 or d[14] != 8:
            raise NotImplementedError(f"{path}: {d[13]}-byte offsets / {d[14]}-byte lengths (only 8 are read)")
        p = 24 + (4 if ver == 1 else 0)  # version 1 adds the indexed-storage K and reserved bytes
        self.base = struct.unpack_from("<Q", d, p)[0]
        root = p + 32  # root group symbol table entry
        self.root = struct.unpack_from("<Q", d, root + 8)[0]

    # ---------------------------------------------------------------------------------------- low level
    def _messages(self, addr):
        d = self.d
        if d[addr] != 1:
            raise NotImplementedError(f"object header version {d[addr]} at {addr} (only version 1 is read)")
        hsize = struct.unpack_from("<I", d, addr + 8)[0]
        blocks, msgs = [(addr + 16, hsize)], []
        while blocks:
            start, size = blocks.pop(0)
            p = start
            while p + 8 <= start + size:
                mt, ms = struct.unpack_from("<HH", d, p)
                body = d[p + 8 : p + 8 + ms]
                if mt == 0x10:  # header continuation
                    blocks.append(struct.unpack_from("<QQ", body))
                msgs.append((mt, body))
                p += 8 + ms
        return msgs

    def _heap_string(self, heap, off):
        d = self.d
        if d[heap : heap + 4] != b"HEAP":
            raise ValueError("bad local heap")
        data = struct.unpack_from("<Q", d, heap + 24)[0]
        s = data + off
        return d[s : d.index(b"\0", s)].decode()

    def _group(self, addr):
        """name -> object header address of the members of the group whose header is at addr."""
        st = [b for t, b in self._messages(addr) if t == 0x11]
        if not st:
            raise NotImplementedError("group without a symbol table message (new-style groups are not read)")
        btree, heap = struct.unpack_from("<QQ", st[0])
        out = {}

        def walk(node):
            d = self.d
            if d[node : node + 4] != b"TREE" or d[node + 4] != 0:
                raise ValueError("bad group B-tree node")
            level, n = d[node + 5], struct.unpack_from("<H", d, node + 6)[0]
            p = node + 24 + 8  # header + siblings, then key 0
            for _ in range(n):
                child = struct.unpack_from("<Q", d, p)[0]
                p += 16  # child, next key
                if level:
                    walk(child)
                    continue
                if d[child : child + 4] != b"SNOD":
                    raise ValueError("bad symbol node")
                q = child + 8
                for _ in range(struct.unpack_from("<H", d, child + 6)[0]):
                    name_off, ohdr = struct.unpack_from("<QQ", d, q)
                    out[self._heap_string(heap, name_off)] = ohdr
                    q += 40

        walk(btree)
        return out

    def _find(self, path):
        node = self.root
        for part in [p for p in str(path).split("/") if p]:
            members = self._group(node)
            if part not in members:
                raise KeyError(f"{self.path}: no object {path!r}")
            node = members[part]
        return node

    # ---------------------------------------------------------------------------------------- public
    def keys(self, path="/"):
        return sorted(self._group(self._find(path)))

    def has(self, path) -> bool:
        try:
            self._find(path)
            return True
        except KeyError:
            return False

    def read(self, path) -> np.ndarray:
        msgs = {}
        for t, b in self._messages(self._find(path)):
            msgs.setdefault(t, b)
        if 0x0B in msgs:
            raise NotImplementedError(f"{path}: filtered (compressed) dataset")
        if 1 not in msgs or 3 not in msgs or 8 not in msgs:
            raise ValueError(f"{path}: not a dataset")
        ds = msgs[1]
        ver, rank = ds[0], ds[1]
        dims = tuple(int(x) for x in struct.unpack_from("<" + "Q" * rank, ds, 8 if ver == 1 else 4))
        dt = msgs[3]
        cls, size = dt[0] & 0x0F, struct.unpack_from("<I", dt, 4)[0]
        big = bool(dt[1] & 1)
        kind = {0: "i", 1: "f"}.get(cls)
        if kind is None or size not in (4, 8):
            raise NotImplementedError(f"{path}: datatype class {cls} size {size}")
        if cls == 0 and not dt[1] & 0x08:
            kind = "u"
        dtype = np.dtype(("<" if not big else ">") + kind + str(size))
        lay = msgs[8]
        if lay[0] != 3:
            raise NotImplementedError(f"{path}: data layout message version {lay[0]}")
        count = int(np.prod(dims)) if dims else 1
        if lay[1] == 1:  # contiguous
            addr = struct.unpack_from("<Q", lay, 2)[0]
            if addr == 0xFFFFFFFFFFFFFFFF:
                return np.zeros(dims, dtype=dtype.newbyteorder("="))
            return np.frombuffer(self.d, dtype=dtype, count=count, offset=addr).reshape(dims).astype(dtype.newbyteorder("="))
        if lay[1] == 0:  # compact
            sz = struct.unpack_from("<H", lay, 2)[0]
            return np.frombuffer(lay[4 : 4 + sz], dtype=dtype, count=count).reshape(dims).astype(dtype.newbyteorder("="))
        if lay[1] != 2:
            raise NotImplementedError(f"{path}: layout class {lay[1]}")
        nd = lay[2]
        btree = struct.unpack_from("<Q", lay, 3)[0]
        cdims = struct.unpack_from("<" + "I" * nd, lay, 11)[:-1]  # the last entry is the element size
        full = np.zeros(dims, dtype=dtype.newbyteorder("="))
        ksz = 8 + 8 * nd

        def walk(node):
            d = self.d
            if d[node : node + 4] != b"TREE" or d[node + 4] != 1:
                raise ValueError("bad chunk B-tree node")
            level, n = d[node + 5], struct.unpack_from("<H", d, node + 6)[0]
            p = node + 24
            for _ in range(n):
                offs = struct.unpack_from("<" + "Q" * nd, d, p + 8)[:-1]
                child = struct.unpack_from("<Q", d, p + ksz)[0]
                if level:
                    walk(child)
                else:
                    chunk = np.frombuffer(d, dtype=dtype, count=int(np.prod(cdims)), offset=child).reshape(cdims)
                    dst = tuple(slice(o, min(o + c, D)) for o, c, D in zip(offs, cdims, dims))
                    full[dst] = chunk[tuple(slice(0, s.stop - s.start) for s in dst)]
                p += ksz + 8

        walk(btree)
        return full
