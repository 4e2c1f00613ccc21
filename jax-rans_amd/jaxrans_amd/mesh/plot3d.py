"""3D Plot3D grid files: multi-block, formatted (text) or unformatted
(Fortran sequential records, either byte order), double or single
precision, with or without iblank.

The reference reads the same variants (``src/jaxrans/mesh/plot3d.py:1-11``,
layout deduced from the sizes); this module reads the 3D ones and writes the
unformatted double-precision multi-block form.  Node order in a block is the
Plot3D one, ``i`` fastest; blocks come back as ``(3, ni, nj, nk)`` node
arrays.
"""

from __future__ import annotations

import gzip
from pathlib import Path

import numpy as np


def _raw(path: Path) -> bytes:
    data = path.read_bytes()
    return gzip.decompress(data) if data[:2] == b"\x1f\x8b" else data


def _binary(raw: bytes) -> bool:
    return any(b < 9 or 13 < b < 32 for b in raw[:4096])


def _blocks_from(values: np.ndarray, dims: np.ndarray, iblank: bool):
    out, at = [], 0
    for shape in dims:
        n = int(np.prod(shape))
        coords = []
        for _ in range(3):
            coords.append(values[at:at + n].reshape(tuple(shape[::-1])).T)
            at += n
        if iblank:
            at += n
        out.append(np.stack(coords))
    return out if at == values.size else None


def _formatted(raw: bytes):
    text = raw.decode(errors="replace").replace("D", "E").replace("d", "e")
    tokens = text.split()
    try:
        numbers = np.array([float(t) for t in tokens])
    except ValueError:
        return None
    for multi in (True, False):
        head = 1 if multi else 0
        if numbers.size <= head:
            continue
        nb = int(numbers[0]) if multi else 1
        if multi and (numbers[0] != nb or nb < 1):
            continue
        dims = numbers[head:head + 3 * nb]
        if dims.size < 3 * nb or np.any(dims != np.round(dims)) or np.any(dims < 1):
            continue
        dims = dims.astype(np.int64).reshape(nb, 3)
        values = numbers[head + 3 * nb:]
        n = int(dims.prod(axis=1).sum())
        for iblank in (False, True):
            if values.size == n * (3 + iblank):
                return _blocks_from(values, dims, iblank)
    return None


def _records(raw: bytes, order: str):
    out, at = [], 0
    while at < len(raw):
        if at + 4 > len(raw):
            return None
        n = int(np.frombuffer(raw, order + "i4", 1, at)[0])
        if n < 0 or at + 8 + n > len(raw) or int(np.frombuffer(raw, order + "i4", 1, at + 4 + n)[0]) != n:
            return None
        out.append(raw[at + 4:at + 4 + n])
        at += n + 8
    return out


def _unformatted(raw: bytes):
    for order in ("<", ">"):
        recs = _records(raw, order)
        if not recs:
            continue
        for multi in (True, False):
            k = 0
            if multi:
                if len(recs[0]) != 4:
                    continue
                nb, k = int(np.frombuffer(recs[0], order + "i4")[0]), 1
            else:
                nb = 1
            if k >= len(recs):
                continue
            dims = np.frombuffer(recs[k], order + "i4")
            if dims.size != 3 * nb or np.any(dims < 1) or len(recs) != k + 1 + nb:
                continue
            shapes = dims.reshape(nb, 3)
            blocks = []
            for shape, rec in zip(shapes, recs[k + 1:]):
                n = int(np.prod(shape))
                for prec in ("f8", "f4"):
                    size = np.dtype(prec).itemsize
                    if len(rec) in (3 * n * size, n * (3 * size + 4)):
                        v = np.frombuffer(rec, order + prec, 3 * n).astype(np.float64)
                        blocks.append(np.stack([v[c * n:(c + 1) * n].reshape(tuple(shape[::-1])).T
                                                for c in range(3)]))
                        break
                else:
                    break
            if len(blocks) == nb:
                return blocks
    return None


def read_plot3d_3d(path: str | Path) -> list[np.ndarray]:
    """Every block of a 3D Plot3D file, nodes ``(3, ni, nj, nk)``."""
    path = Path(path)
    raw = _raw(path)
    blocks = _unformatted(raw) if _binary(raw) else _formatted(raw)
    if blocks is None:
        raise ValueError(f"{path.name}: not a 3D Plot3D grid this reader understands (formatted or unformatted, "
                         f"single or multi-block, double or single precision, with or without iblank)")
    return blocks


def write_plot3d_3d(path: str | Path, blocks) -> Path:
    """Multi-block unformatted (little-endian, double precision) Plot3D."""
    path = Path(path)

    def record(payload: bytes) -> bytes:
        n = np.array([len(payload)], "<i4").tobytes()
        return n + payload + n

    parts = [record(np.array([len(blocks)], "<i4").tobytes()),
             record(np.array([b.shape[1:] for b in blocks], "<i4").ravel().tobytes())]
    for b in blocks:
        parts.append(record(np.concatenate([np.asarray(b[c], np.float64).T.ravel() for c in range(3)])
                            .astype("<f8").tobytes()))
    path.write_bytes(b"".join(parts))
    return path
