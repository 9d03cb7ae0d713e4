"""Minimal read-only HDF5 reader for the reference's results file.

h5py is not installed in this image, so the golden fixture
``dda_results/pcas_ocbs_simulated_data.hdf5`` of the reference (written by
``dda_results/create_h5py.ipynb`` and filled by ``run_dda.py:131-141``) is read
with this small parser instead.  It supports exactly what that file uses:

* superblock version 0, 8-byte offsets and lengths;
* groups stored as symbol tables (v1 B-tree of group nodes + local heap);
* version-1 object headers (with continuation blocks);
* dataspace messages (v1/v2), datatype messages of class float (IEEE LE) and
  compound {r, i} of two floats (h5py's complex layout);
* contiguous data layout (layout message version 3) or an undefined address
  (never-written dataset -> zeros / fill).

Attributes are not decoded.  This is test infrastructure only.
"""
import struct

import numpy as np

_SIG = b"\x89HDF\r\n\x1a\n"


class _File:
    def __init__(self, path):
        with open(path, "rb") as fh:
            self.d = fh.read()
        d = self.d
        if d[:8] != _SIG:
            raise ValueError("not an HDF5 file")
        if d[8] != 0:
            raise ValueError("only superblock version 0 is supported")
        self.so = d[13]
        self.sl = d[14]
        if self.so != 8 or self.sl != 8:
            raise ValueError("only 8-byte offsets/lengths are supported")
        # superblock v0: 8 sig + 8 versions/sizes + 2+2 K + 4 flags = 24
        p = 24
        self.base, _fs, self.eof, _drv = struct.unpack_from("<4Q", d, p)
        p += 32
        # root group symbol table entry
        self.root = self._symbol_entry(p)

    # -- low level ---------------------------------------------------------
    def _u(self, fmt, p):
        return struct.unpack_from(fmt, self.d, p)

    def _symbol_entry(self, p):
        name_off, ohdr, cache = self._u("<QQI", p)
        scratch = self.d[p + 24 : p + 40]
        btree = heap = None
        if cache == 1:
            btree, heap = struct.unpack_from("<QQ", scratch, 0)
        return dict(name_off=name_off, ohdr=ohdr, cache=cache, btree=btree, heap=heap)

    def _heap_data(self, heap_addr):
        if self.d[heap_addr : heap_addr + 4] != b"HEAP":
            raise ValueError("bad local heap")
        _size, _free, data_addr = self._u("<QQQ", heap_addr + 8)
        return data_addr

    def _cstr(self, p):
        e = self.d.index(b"\x00", p)
        return self.d[p:e].decode()

    def _messages(self, ohdr):
        d = self.d
        ver = d[ohdr]
        if ver != 1:
            raise ValueError("only v1 object headers are supported")
        nmsg, _rc, hsize = self._u("<HII", ohdr + 2)
        blocks = [(ohdr + 16, hsize)]
        out = []
        while blocks:
            start, size = blocks.pop(0)
            p = start
            end = start + size
            while p + 8 <= end and len(out) < nmsg:
                mtype, msize, _flags = self._u("<HHB", p)
                body = p + 8
                if mtype == 0x0010:  # continuation
                    off, length = self._u("<QQ", body)
                    blocks.append((off, length))
                out.append((mtype, body, msize))
                p = body + msize
        return out

    # -- groups ------------------------------------------------------------
    def _group_children(self, btree, heap):
        data_addr = self._heap_data(heap)
        out = {}
        stack = [btree]
        while stack:
            node = stack.pop()
            if self.d[node : node + 4] != b"TREE":
                raise ValueError("bad B-tree node")
            ntype, level, used = self._u("<BBH", node + 4)
            p = node + 8 + 16  # skip siblings
            p += 8  # key 0
            children = []
            for _ in range(used):
                (child,) = self._u("<Q", p)
                children.append(child)
                p += 16  # child + next key
            if level > 0:
                stack.extend(children)
                continue
            for snod in children:
                if self.d[snod : snod + 4] != b"SNOD":
                    raise ValueError("bad symbol node")
                (nsym,) = self._u("<H", snod + 6)
                q = snod + 8
                for _ in range(nsym):
                    ent = self._symbol_entry(q)
                    name = self._cstr(data_addr + ent["name_off"])
                    out[name] = ent
                    q += 40
        return out

    def _object_children(self, ent):
        if ent["cache"] == 1:
            return self._group_children(ent["btree"], ent["heap"])
        for mtype, body, _ in self._messages(ent["ohdr"]):
            if mtype == 0x0011:
                btree, heap = self._u("<QQ", body)
                return self._group_children(btree, heap)
        return None

    # -- datasets ----------------------------------------------------------
    def _dtype(self, body):
        cv = self.d[body]
        cls = cv & 0x0F
        (size,) = self._u("<I", body + 4)
        if cls == 1:
            if size == 8:
                return np.dtype("<f8")
            if size == 4:
                return np.dtype("<f4")
            raise ValueError("unsupported float size %d" % size)
        if cls == 0:
            return np.dtype("<i%d" % size)
        if cls == 6:
            # compound of two floats named r, i  -> complex
            if size == 16:
                return np.dtype("<c16")
            if size == 8:
                return np.dtype("<c8")
        raise ValueError("unsupported datatype class %d" % cls)

    def _dataset(self, ent):
        shape = dtype = addr = None
        for mtype, body, _ in self._messages(ent["ohdr"]):
            if mtype == 0x0001:
                ver, ndim, flags = self._u("<BBB", body)
                p = body + (8 if ver == 1 else 4)
                shape = tuple(self._u("<%dQ" % ndim, p)) if ndim else ()
            elif mtype == 0x0003:
                dtype = self._dtype(body)
            elif mtype == 0x0008:
                ver, cls = self._u("<BB", body)
                if ver != 3 or cls != 1:
                    raise ValueError("only contiguous layout v3 is supported")
                addr, size = self._u("<QQ", body + 2)
        if shape is None or dtype is None:
            return None
        count = int(np.prod(shape)) if shape else 1
        if addr is None or addr == 0xFFFFFFFFFFFFFFFF:
            return np.zeros(shape, dtype)
        arr = np.frombuffer(self.d, dtype=dtype, count=count, offset=addr)
        return arr.reshape(shape).copy()

    def walk(self):
        """Return {path: ndarray} for every dataset in the file."""
        out = {}
        stack = [("", self.root)]
        while stack:
            path, ent = stack.pop()
            kids = self._object_children(ent)
            if kids is not None:
                for name, e in kids.items():
                    stack.append((path + "/" + name, e))
                continue
            arr = self._dataset(ent)
            if arr is not None:
                out[path] = arr
        return out


def read_all(path):
    return _File(path).walk()
