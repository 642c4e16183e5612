"""Staged-input HDF5 reader: the reference's staged cache (Hamiltonian + trial) without h5py or PySCF.

The reference stages a molecule once (PySCF integrals, frozen core, Cholesky, trial orbitals) and caches
it with h5py (trot/staging.py:1353-1423, ``_dump_h5`` / ``_load_h5``):

    /            attr meta_json (str)
    /ham         datasets h0 (scalar), h1 (N, N), chol (G, N, N), nelec (2,) int64;
                 attrs norb, chol_cut, source_kind, basis; frozen as an int attr or an int64 dataset
    /trial       attrs kind, source_kind; frozen as above
    /trial/data  one dataset per trial array ("mo", or "mo_a" / "mo_b")

h5py is not installed here, so this module reads that subset of the HDF5 file format (the HDF Group's
published File Format Specification, versions 0-3 of the superblock) directly: superblocks 0-3, object
headers 1 and 2 with continuation blocks, groups stored as symbol tables (v1 B-tree + local heap) or as
link messages, attributes (messages 1-3) and datasets with compact or contiguous layout; datatypes
fixed-point, IEEE float, fixed-length and variable-length strings (global heap).  Anything else
(chunked or filtered data, dense link / attribute storage, shared messages) raises
NotImplementedError rather than guessing.  ``load_staged`` returns the reference's StagedInputs shape.
"""

from __future__ import annotations

import json
from dataclasses import dataclass
import mmap
from pathlib import Path
from typing import Any, Dict

import numpy as np

_SIG = b"\x89HDF\r\n\x1a\n"


class _Reader:
    def __init__(self, data):
        """``data``: the file as bytes or a read-only mmap; sliced through one memoryview (no copies)."""
        self.src = data  # for find()
        self.b = memoryview(data)
        base = None
        for off in (0, 512, 1024, 2048, 4096, 8192):
            if data[off:off + 8] == _SIG:
                base = off
                break
        if base is None:
            raise ValueError("not an HDF5 file (no superblock signature)")
        self.sb = base
        ver = data[base + 8]
        self.ver = ver
        if ver in (0, 1):
            self.O = data[base + 13]
            self.L = data[base + 14]
            p = base + 24 + (4 if ver == 1 else 0)
            self.base_addr = self.u(p, self.O)
            p += 4 * self.O  # base, free-space info, end of file, driver info
            # root group symbol table entry: name offset, object header address, cache type, ...
            self.root = self.u(p + self.O, self.O)
        elif ver in (2, 3):
            self.O = data[base + 9]
            self.L = data[base + 10]
            p = base + 12
            self.base_addr = self.u(p, self.O)
            self.root = self.u(p + 3 * self.O, self.O)
        else:
            raise NotImplementedError(f"HDF5 superblock version {ver}")
        self.undef = (1 << (8 * self.O)) - 1

    # -- primitive reads ------------------------------------------------------------------------
    def u(self, off: int, n: int) -> int:
        return int.from_bytes(self.b[off:off + n], "little")

    def addr(self, a: int) -> int:
        return self.base_addr + a

    def cstr(self, off: int) -> str:
        end = self.src.find(b"\0", off)
        if end < 0:
            raise ValueError("unterminated string")
        return bytes(self.b[off:end]).decode("utf-8")

    # -- object headers ---------------------------------------------------------------------------
    def messages(self, a: int):
        """[(type, data offset, size, flags)] of the object header at address a."""
        b = self.b
        p = self.addr(a)
        out = []
        if b[p:p + 4] == b"OHDR":
            if b[p + 4] != 2:
                raise NotImplementedError("object header version")
            flags = b[p + 5]
            q = p + 6
            if flags & 0x20:
                q += 16
            if flags & 0x10:
                q += 4
            nsz = 1 << (flags & 3)
            size0 = self.u(q, nsz)
            q += nsz
            blocks = [(q, q + size0)]
            order = 2 if flags & 0x04 else 0
            while blocks:
                s, e = blocks.pop(0)
                while s + 4 + order <= e:
                    mtype, msize, mflags = b[s], self.u(s + 1, 2), b[s + 3]
                    d = s + 4 + order
                    if mtype == 0x10:
                        c = self.addr(self.u(d, self.O))
                        clen = self.u(d + self.O, self.L)
                        if b[c:c + 4] != b"OCHK":
                            raise ValueError("bad continuation block")
                        blocks.append((c + 4, c + clen - 4))
                    elif mtype != 0:
                        out.append((mtype, d, msize, mflags))
                    s = d + msize
            return out
        if b[p] != 1:
            raise NotImplementedError(f"object header version {b[p]}")
        nmsg = self.u(p + 2, 2)
        hsize = self.u(p + 8, 4)
        blocks = [(p + 16, p + 16 + hsize)]
        seen = 0
        while blocks and seen < nmsg:
            s, e = blocks.pop(0)
            while s + 8 <= e and seen < nmsg:
                mtype, msize, mflags = self.u(s, 2), self.u(s + 2, 2), b[s + 4]
                d = s + 8
                seen += 1
                if mtype == 0x10:
                    c = self.addr(self.u(d, self.O))
                    clen = self.u(d + self.O, self.L)
                    blocks.append((c, c + clen))
                elif mtype != 0:
                    out.append((mtype, d, msize, mflags))
                s = d + msize
        return out

    # -- datatypes, dataspaces, data ----------------------------------------------------------------
    def datatype(self, off: int) -> dict:
        b = self.b
        cls, ver = b[off] & 0x0F, b[off] >> 4
        bits = self.u(off + 1, 3)
        size = self.u(off + 4, 4)
        if cls in (0, 1):
            if bits & 1:
                raise NotImplementedError("big-endian data")
            if cls == 0:
                kind = "i" if bits & 0x08 else "u"
            else:
                kind = "f"
            return {"kind": "num", "dtype": np.dtype(f"<{kind}{size}"), "size": size}
        if cls == 3:
            return {"kind": "str", "size": size}
        if cls == 9:
            if bits & 0x0F != 1:
                raise NotImplementedError("variable-length sequences")
            return {"kind": "vlstr", "size": size}
        raise NotImplementedError(f"HDF5 datatype class {cls} (version {ver})")

    def dataspace(self, off: int) -> tuple:
        b = self.b
        ver, rank, flags = b[off], b[off + 1], b[off + 2]
        if ver == 1:
            p = off + 8
        elif ver == 2:
            if b[off + 3] == 2:
                return None  # null dataspace
            p = off + 4
        else:
            raise NotImplementedError(f"dataspace version {ver}")
        return tuple(self.u(p + i * self.L, self.L) for i in range(rank))

    def gheap_object(self, coll: int, index: int) -> bytes:
        b = self.b
        p = self.addr(coll)
        if b[p:p + 4] != b"GCOL":
            raise ValueError("bad global heap collection")
        csize = self.u(p + 8, self.L)
        q, end = p + 8 + self.L, p + csize
        while q + 8 + self.L <= end:
            idx = self.u(q, 2)
            osize = self.u(q + 8, self.L)
            d = q + 8 + self.L
            if idx == index:
                return bytes(b[d:d + osize])
            if idx == 0:
                break
            q = d + ((osize + 7) // 8) * 8
        raise KeyError(f"global heap object {index} not found")

    def decode(self, raw: bytes, dt: dict, shape: tuple):
        n = int(np.prod(shape)) if shape else 1
        if dt["kind"] == "num":
            arr = np.frombuffer(raw, dtype=dt["dtype"], count=n).copy()
            return arr.reshape(shape) if shape else arr[0]
        if dt["kind"] == "str":
            s = dt["size"]
            vals = [bytes(raw[i * s:(i + 1) * s]).split(b"\0", 1)[0].decode("utf-8") for i in range(n)]
            return vals[0] if not shape else np.array(vals, dtype=object).reshape(shape)
        if dt["kind"] == "vlstr":
            vals = []
            for i in range(n):
                r = raw[i * dt["size"]:(i + 1) * dt["size"]]
                length = int.from_bytes(r[0:4], "little")
                coll = int.from_bytes(r[4:4 + self.O], "little")
                idx = int.from_bytes(r[4 + self.O:8 + self.O], "little")
                vals.append(self.gheap_object(coll, idx)[:length].decode("utf-8"))
            return vals[0] if not shape else np.array(vals, dtype=object).reshape(shape)
        raise NotImplementedError(dt["kind"])

    # -- objects ------------------------------------------------------------------------------------
    def attrs(self, a: int) -> Dict[str, Any]:
        b = self.b
        out = {}
        for mtype, d, size, mflags in self.messages(a):
            if mtype == 0x15:
                if self.u(d + 2 + (2 if b[d + 1] & 1 else 0), self.O) != self.undef:
                    raise NotImplementedError("dense attribute storage")
                continue
            if mtype != 0x0C:
                continue
            if mflags & 0x02:
                raise NotImplementedError("shared attribute message")
            ver = b[d]
            nlen, tlen, slen = self.u(d + 2, 2), self.u(d + 4, 2), self.u(d + 6, 2)
            pad = (lambda x: (x + 7) // 8 * 8) if ver == 1 else (lambda x: x)
            p = d + 8 + (1 if ver == 3 else 0)
            name = bytes(b[p:p + nlen]).split(b"\0", 1)[0].decode("utf-8")
            p += pad(nlen)
            dt = self.datatype(p)
            p += pad(tlen)
            shape = self.dataspace(p)
            p += pad(slen)
            n = int(np.prod(shape)) if shape else 1
            out[name] = self.decode(b[p:p + n * dt["size"]], dt, shape)
        return out

    def members(self, a: int) -> Dict[str, int]:
        """name -> object header address of a group's links (hard links only)."""
        b = self.b
        out = {}
        for mtype, d, size, _ in self.messages(a):
            if mtype == 0x11:  # symbol table: v1 B-tree of symbol nodes + local heap of names
                btree, heap = self.u(d, self.O), self.u(d + self.O, self.O)
                hp = self.addr(heap)
                if b[hp:hp + 4] != b"HEAP":
                    raise ValueError("bad local heap")
                names_at = self.addr(self.u(hp + 8 + 2 * self.L, self.O))
                self._walk_group_btree(btree, names_at, out)
            elif mtype == 0x06:  # link message
                flags = b[d + 1]
                p = d + 2
                ltype = 0
                if flags & 0x08:
                    ltype = b[p]
                    p += 1
                if flags & 0x04:
                    p += 8
                if flags & 0x10:
                    p += 1
                nsz = 1 << (flags & 3)
                nlen = self.u(p, nsz)
                p += nsz
                name = bytes(b[p:p + nlen]).decode("utf-8")
                p += nlen
                if ltype == 0:
                    out[name] = self.u(p, self.O)
            elif mtype == 0x02:  # link info: dense storage when the fractal heap address is defined
                flags = b[d + 1]
                p = d + 2 + (8 if flags & 1 else 0)
                if self.u(p, self.O) != self.undef:
                    raise NotImplementedError("dense link storage")
        return out

    def _walk_group_btree(self, node: int, names_at: int, out: dict) -> None:
        b = self.b
        p = self.addr(node)
        if b[p:p + 4] != b"TREE" or b[p + 4] != 0:
            raise ValueError("bad group B-tree node")
        level, used = b[p + 5], self.u(p + 6, 2)
        q = p + 8 + 2 * self.O + self.L  # past siblings and key 0
        for _ in range(used):
            child = self.u(q, self.O)
            q += self.O + self.L
            if level > 0:
                self._walk_group_btree(child, names_at, out)
                continue
            s = self.addr(child)
            if b[s:s + 4] != b"SNOD":
                raise ValueError("bad symbol table node")
            nsym = self.u(s + 6, 2)
            e = s + 8
            for _ in range(nsym):
                name = self.cstr(names_at + self.u(e, self.O))
                out[name] = self.u(e + self.O, self.O)
                e += 2 * self.O + 24

    def dataset(self, a: int) -> np.ndarray:
        b = self.b
        dt = shape = layout = None
        for mtype, d, size, mflags in self.messages(a):
            if mflags & 0x02 and mtype in (0x01, 0x03):
                raise NotImplementedError("shared datatype or dataspace")
            if mtype == 0x01:
                shape = self.dataspace(d)
            elif mtype == 0x03:
                dt = self.datatype(d)
            elif mtype == 0x0B:
                raise NotImplementedError("filtered (compressed) datasets")
            elif mtype == 0x08:
                layout = d
        if dt is None or layout is None:
            raise ValueError("not a dataset")
        shape = shape or ()
        n = int(np.prod(shape)) if shape else 1
        nbytes = n * dt["size"]
        ver, p = b[layout], layout
        if ver in (3, 4):
            cls = b[p + 1]
            if cls == 0:
                raw = b[p + 4:p + 4 + self.u(p + 2, 2)]
            elif cls == 1:
                raw = self.contiguous(self.u(p + 2, self.O), nbytes)
            else:
                raise NotImplementedError("chunked or virtual dataset layout")
        elif ver in (1, 2):
            rank, cls = b[p + 1], b[p + 2]
            q = p + 8
            if cls == 1:
                raw = self.contiguous(self.u(q, self.O), nbytes)
            elif cls == 0:
                q += 4 * rank
                raw = b[q + 4:q + 4 + self.u(q, 4)]
            else:
                raise NotImplementedError("chunked dataset layout")
        else:
            raise NotImplementedError(f"data layout version {ver}")
        if len(raw) < nbytes:
            raw = bytes(raw) + b"\0" * (nbytes - len(raw))  # unallocated storage reads as the zero fill value
        return self.decode(raw, dt, shape)

    def contiguous(self, start: int, nbytes: int):
        """Bytes of contiguous storage at file address `start`; the undefined address is unallocated storage
        (read as the zero fill value), and storage that runs past the end of the file is a truncated file."""
        if start == self.undef:
            return b""
        a = self.addr(start)
        if a + nbytes > len(self.b):
            raise ValueError(f"dataset storage [{a}, {a + nbytes}) runs past the end of the file ({len(self.b)} "
                             "bytes): truncated file")
        return self.b[a:a + nbytes]

    def is_dataset(self, a: int) -> bool:
        return any(m[0] == 0x08 for m in self.messages(a))


class H5File:
    """Read-only view of an HDF5 file: ``f["ham/h1"]`` (numpy array), ``f.attrs(path)``, ``f.keys(path)``."""

    def __init__(self, path):
        # memory-mapped: datasets are sliced from the mapping without copies and copied once into the
        # returned arrays (a 5 GB Cholesky array never sits in memory twice)
        with open(Path(path), "rb") as fh:
            try:
                self._mm = mmap.mmap(fh.fileno(), 0, access=mmap.ACCESS_READ)
            except ValueError as e:  # empty file
                raise ValueError(f"{path}: not an HDF5 file") from e
        self._r = _Reader(self._mm)

    def _resolve(self, path: str) -> int:
        a = self._r.root
        for part in [p for p in path.split("/") if p]:
            m = self._r.members(a)
            if part not in m:
                raise KeyError(path)
            a = m[part]
        return a

    def keys(self, path: str = "/") -> list:
        return list(self._r.members(self._resolve(path)))

    def attrs(self, path: str = "/") -> Dict[str, Any]:
        return self._r.attrs(self._resolve(path))

    def __contains__(self, path: str) -> bool:
        try:
            self._resolve(path)
            return True
        except KeyError:
            return False

    def __getitem__(self, path: str):
        a = self._resolve(path)
        if not self._r.is_dataset(a):
            raise TypeError(f"{path} is a group")
        return self._r.dataset(a)


# ---------------------------------------------------------------------------
# the reference's staged inputs (trot/staging.py:471-500, 1388-1423)
# ---------------------------------------------------------------------------


@dataclass(frozen=True)
class HamInput:
    """trot/staging.py:471-483."""

    h0: float
    h1: np.ndarray
    chol: np.ndarray
    nelec: tuple
    norb: int
    chol_cut: float
    frozen: Any
    source_kind: str
    basis: str


@dataclass(frozen=True)
class TrialInput:
    """trot/staging.py:486-493."""

    kind: str
    data: Dict[str, np.ndarray]
    frozen: Any
    source_kind: str


@dataclass(frozen=True)
class StagedInputs:
    """trot/staging.py:496-500."""

    ham: HamInput
    trial: TrialInput
    meta: Dict[str, Any]


def _frozen(f: H5File, group: str):
    """trot/staging.py:456-468: an int64 dataset named "frozen", else the int attribute."""
    if f"{group}/frozen" in f:
        return np.asarray(f[f"{group}/frozen"], dtype=np.int64)
    v = f.attrs(group)["frozen"]
    return int(np.asarray(v).item())


def _meta_frozen(v):
    """trot/staging.py:437-446."""
    if v is None:
        return None
    if isinstance(v, (list, np.ndarray)):
        return np.asarray(v, dtype=np.int64)
    if isinstance(v, (int, np.integer)):
        return int(v)
    raise TypeError(f"Unsupported frozen metadata type: {type(v)}")


def load_staged(path) -> StagedInputs:
    """trot/staging.py:1388-1423 (_load_h5): the staged Hamiltonian and trial inputs of a cache file."""
    f = H5File(path)
    meta = json.loads(str(f.attrs("/")["meta_json"]))
    if "frozen" in meta:
        meta["frozen"] = _meta_frozen(meta["frozen"])
    ga = f.attrs("ham")
    nelec = np.asarray(f["ham/nelec"])
    ham = HamInput(
        h0=float(np.asarray(f["ham/h0"]).item()),
        h1=np.asarray(f["ham/h1"]),
        chol=np.asarray(f["ham/chol"]),
        nelec=(int(nelec[0]), int(nelec[1])),
        norb=int(np.asarray(ga["norb"]).item()),
        chol_cut=float(np.asarray(ga["chol_cut"]).item()),
        frozen=_frozen(f, "ham"),
        source_kind=str(ga["source_kind"]),
        basis=str(ga["basis"]),
    )
    gt = f.attrs("trial")
    trial = TrialInput(
        kind=str(gt["kind"]),
        data={k: np.asarray(f[f"trial/data/{k}"]) for k in f.keys("trial/data")},
        frozen=_frozen(f, "trial"),
        source_kind=str(gt["source_kind"]),
    )
    return StagedInputs(ham=ham, trial=trial, meta=meta)


def ham_chol(staged: StagedInputs):
    """The staged Hamiltonian as trot_amd.core.HamChol (host arrays; moved to the GPU on first use)."""
    from .core import HamChol

    return HamChol(staged.ham.h0, staged.ham.h1, staged.ham.chol, basis=staged.ham.basis)
