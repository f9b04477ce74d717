"""A small HDF5 writer for measurement files (no HDF5 library needed).

It writes the subset of the HDF5 file format (HDF5 File Format Specification,
version 3.0, The HDF Group) that carlo-style measurement files use:

* superblock version 0, 8-byte offsets and lengths, the library's default
  B-tree ranks (group leaf K = 4, group internal K = 16, chunk index K = 32);
* groups in the original form: a version-1 object header with a symbol-table
  message, a local heap for the member names, a version-1 B-tree (node type 0)
  whose leaf points at symbol-table nodes of at most 2K = 8 entries, members
  sorted by name;
* datasets with a version-1 object header holding dataspace (version 1),
  datatype (version 1), fill value (version 2: default value, write time
  "if set") and layout (version 3) messages.  Scalars and strings are
  contiguous; a chunked dataset stores full chunks listed by a version-1
  B-tree (node type 1) whose keys are (chunk bytes, filter mask, offsets),
  the last key being the end of the last chunk.

Datatypes: little-endian int64, IEEE float64, fixed-length ASCII strings
(null-terminated padding).  Files are built in memory and written at once.
"""
from __future__ import annotations

import struct
from typing import Dict, Union

import numpy as np

UNDEF = 0xFFFFFFFFFFFFFFFF   # undefined address; also H5S_UNLIMITED
_HEAP_FREE_NONE = 1          # free-list head of a local heap without free blocks
_LEAF_K, _GROUP_K, _CHUNK_K = 4, 16, 32


def _pad8(n: int) -> int:
    return (n + 7) & ~7


class Dataset:
    """A dataset node: `dtype` is 'i8', 'f8' or 'S' (string)."""

    def __init__(self, dtype: str, data, shape=None, maxshape=None, chunks=None):
        self.dtype = dtype
        if dtype == "S":
            raw = data.encode("ascii") if isinstance(data, str) else bytes(data)
            if not raw:
                raise ValueError("string datasets need at least one byte")
            self.raw, self.shape, self.elem = raw, (), len(raw)
        else:
            arr = np.ascontiguousarray(data, dtype="<i8" if dtype == "i8" else "<f8")
            self.shape = tuple(arr.shape) if shape is None else tuple(shape)
            if int(np.prod(self.shape, dtype=np.int64)) != arr.size:
                raise ValueError("data does not fill the shape")
            self.raw, self.elem = arr.tobytes(), 8
        self.maxshape = None if maxshape is None else tuple(maxshape)
        self.chunks = None if chunks is None else tuple(chunks)
        if self.chunks is not None and len(self.chunks) != len(self.shape):
            raise ValueError("chunk rank must equal the dataset rank")
        if self.maxshape is not None and len(self.maxshape) != len(self.shape):
            raise ValueError("maxshape rank must equal the dataset rank")
        if self.maxshape is not None and self.chunks is None and UNDEF in self.maxshape:
            raise ValueError("unlimited dimensions need a chunked layout")


Group = Dict[str, Union["Group", Dataset]]


class _Out:
    def __init__(self):
        self.buf = bytearray()

    def alloc(self, n: int) -> int:
        addr = _pad8(len(self.buf))
        self.buf.extend(b"\0" * (addr + _pad8(n) - len(self.buf)))
        return addr

    def put(self, addr: int, data: bytes) -> None:
        self.buf[addr:addr + len(data)] = data


def _message(mtype: int, body: bytes, flags: int = 0) -> bytes:
    body = body + b"\0" * (_pad8(len(body)) - len(body))
    return struct.pack("<HHB3x", mtype, len(body), flags) + body


def _object_header(out: _Out, messages) -> int:
    data = b"".join(messages)
    addr = out.alloc(16 + len(data))
    out.put(addr, struct.pack("<BBHII4x", 1, 0, len(messages), 1, len(data)) + data)
    return addr


def _dataspace(shape, maxshape) -> bytes:
    flags = 1 if maxshape is not None else 0
    body = struct.pack("<BBBB4x", 1, len(shape), flags, 0)
    body += b"".join(struct.pack("<Q", d) for d in shape)
    if maxshape is not None:
        body += b"".join(struct.pack("<Q", d) for d in maxshape)
    return _message(0x0001, body)


def _datatype(ds: Dataset) -> bytes:
    if ds.dtype == "i8":     # fixed point, version 1: little endian, signed
        body = struct.pack("<B3BI", 0x10, 0x08, 0, 0, 8) + struct.pack("<HH", 0, 64)
    elif ds.dtype == "f8":   # floating point: LE, implied mantissa msb, sign bit 63
        body = (struct.pack("<B3BI", 0x11, 0x20, 63, 0, 8)
                + struct.pack("<HHBBBBI", 0, 64, 52, 11, 0, 52, 1023))
    else:                    # string: null-terminated padding, ASCII
        body = struct.pack("<B3BI", 0x13, 0x00, 0, 0, ds.elem)
    return _message(0x0003, body, flags=1)


def _fill(alloc_time: int) -> bytes:
    # version 2; fill time "if set"; default fill value (defined, size 0)
    return _message(0x0005, struct.pack("<BBBBI", 2, alloc_time, 2, 1, 0), flags=1)


def _write_dataset(out: _Out, ds: Dataset) -> int:
    if ds.chunks is None:
        data_addr = out.alloc(len(ds.raw))
        out.put(data_addr, ds.raw)
        layout = struct.pack("<BBQQ", 3, 1, data_addr, len(ds.raw))
        alloc_time = 2   # late
    else:
        if len(ds.shape) != 1:
            raise ValueError("chunked datasets are one-dimensional here")
        n, c = ds.shape[0], ds.chunks[0]
        if c <= 0:
            raise ValueError("chunk size must be positive")
        nchunks = -(-n // c)
        if nchunks > 2 * _CHUNK_K:
            raise ValueError(f"at most {2 * _CHUNK_K * c} elements in a chunked dataset")
        chunk_bytes = c * ds.elem
        node_size = 24 + 2 * _CHUNK_K * 8 + (2 * _CHUNK_K + 1) * 24
        tree = out.alloc(node_size)
        node = bytearray(b"TREE" + struct.pack("<BBHQQ", 1, 0, nchunks, UNDEF, UNDEF))
        for i in range(nchunks):
            piece = ds.raw[i * chunk_bytes:(i + 1) * chunk_bytes]
            addr = out.alloc(chunk_bytes)   # full chunk; the tail stays zero (the fill)
            out.put(addr, piece)
            node += struct.pack("<IIQQ", chunk_bytes, 0, i * c, 0) + struct.pack("<Q", addr)
        node += struct.pack("<IIQQ", 0, 0, nchunks * c, ds.elem)
        out.put(tree, bytes(node))
        layout = struct.pack("<BBBQII", 3, 2, 2, tree, c, ds.elem)
        alloc_time = 3   # incremental
    msgs = [_dataspace(ds.shape, ds.maxshape), _datatype(ds), _fill(alloc_time),
            _message(0x0008, layout)]
    return _object_header(out, msgs)


def _write_group(out: _Out, members: Group):
    """Returns (object header, B-tree, local heap) addresses of the group."""
    names = sorted(members, key=lambda s: s.encode("utf-8"))
    entries = []
    for name in names:
        node = members[name]
        if isinstance(node, Dataset):
            entries.append((name, _write_dataset(out, node), None))
        else:
            oh, tree, heap = _write_group(out, node)
            entries.append((name, oh, (tree, heap)))
    # local heap: offset 0 holds the empty string, then the names
    data = bytearray(8)
    offsets = []
    for name, _, _ in entries:
        raw = name.encode("utf-8")
        if b"\0" in raw or b"/" in raw or not raw:
            raise ValueError(f"invalid member name {name!r}")
        offsets.append(len(data))
        data += raw + b"\0" * (_pad8(len(raw) + 1) - len(raw))
    heap = out.alloc(32 + len(data))
    out.put(heap, b"HEAP" + struct.pack("<B3xQQQ", 0, len(data), _HEAP_FREE_NONE, heap + 32)
            + bytes(data))
    # symbol-table nodes of at most 2 * leaf K entries each
    per = 2 * _LEAF_K
    snods = []
    for s in range(0, len(entries), per):
        part = entries[s:s + per]
        addr = out.alloc(8 + per * 40)
        body = bytearray(b"SNOD" + struct.pack("<BBH", 1, 0, len(part)))
        for k, (name, oh, cache) in enumerate(part):
            off = offsets[s + k]
            if cache is None:
                body += struct.pack("<QQII16x", off, oh, 0, 0)
            else:
                body += struct.pack("<QQIIQQ", off, oh, 1, 0, cache[0], cache[1])
        out.put(addr, bytes(body))
        snods.append((addr, offsets[s + len(part) - 1]))
    if len(snods) > 2 * _GROUP_K:
        raise ValueError(f"at most {2 * _GROUP_K * per} members in a group")
    tree = out.alloc(24 + 2 * _GROUP_K * 8 + (2 * _GROUP_K + 1) * 8)
    node = bytearray(b"TREE" + struct.pack("<BBHQQ", 0, 0, len(snods), UNDEF, UNDEF))
    node += struct.pack("<Q", 0)   # key 0: the empty name
    for addr, last in snods:
        node += struct.pack("<QQ", addr, last)
    out.put(tree, bytes(node))
    oh = _object_header(out, [_message(0x0011, struct.pack("<QQ", tree, heap))])
    return oh, tree, heap


def file_bytes(root: Group) -> bytes:
    """The bytes of an HDF5 file holding the group tree `root`."""
    out = _Out()
    out.alloc(96)   # superblock
    oh, tree, heap = _write_group(out, root)
    eof = len(out.buf)
    sb = b"\x89HDF\r\n\x1a\n" + struct.pack("<BBBBBBBBHHI", 0, 0, 0, 0, 0, 8, 8, 0,
                                              _LEAF_K, _GROUP_K, 0)
    sb += struct.pack("<QQQQ", 0, UNDEF, eof, UNDEF)
    sb += struct.pack("<QQII", 0, oh, 1, 0) + struct.pack("<QQ", tree, heap)
    assert len(sb) == 96
    out.put(0, sb)
    return bytes(out.buf)


def write_file(path: str, root: Group) -> None:
    with open(path, "wb") as f:
        f.write(file_bytes(root))
