"""
Structure files: the 3D ``.raw`` format and the 2D GIF / PPM / DAT inputs.

3D ``.raw`` (scripts/structure_io3d.py:40-67, LMB3d/lbm3d-perm.cpp:174-193)::

    bytes 0..11   three little-endian int32: nx, ny, nz
    bytes 12..    nx*ny*nz uint8, 1 = solid, 0 = pore, x varying fastest

Arrays are indexed ``[z, y, x]``.  The checks of the reference reader are kept:
bad dimensions, truncated payload, trailing bytes, non-binary voxels and a
mask with no pore space are all errors (``StructureError``).

2D inputs follow LMB2d/lbm.cpp:412-483: a GIF pixel is pore when its red
channel is > 0 (``read_from_gif``), a PPM cell is pore when atoi of its R token
is non-zero, a DAT cell is pore when atoi of its token is 0.  Arrays are indexed ``[y, x]``
(row = solver y, column = solver x, scripts/structure_io.py:44-54).  The GIF
decode uses PIL instead of Magick++.
"""
from __future__ import annotations

import re
from pathlib import Path

import numpy as np

__all__ = ["StructureError", "save_raw", "load_raw", "check_mask3d", "porosity",
           "load_structure2d", "save_gif2d", "atoi"]

HEADER_BYTES = 12


class StructureError(ValueError):
    """A structure the reference reader would reject (its exit code 2)."""


def save_raw(path, solid) -> Path:
    """Write a (nz, ny, nx) {0,1} array as .raw (structure_io3d.py:40-52)."""
    solid = np.ascontiguousarray(solid)
    if solid.ndim != 3:
        raise ValueError(f"expected a 3D array, got shape {solid.shape}")
    solid = solid.astype(np.uint8)
    if solid.size and solid.max() > 1:
        raise ValueError("structure must be binary (0 = pore, 1 = solid)")
    nz, ny, nx = solid.shape
    path = Path(path)
    with open(path, "wb") as fh:
        fh.write(np.array([nx, ny, nz], dtype="<i4").tobytes())
        fh.write(solid.tobytes())
    return path


def load_raw(path, check: bool = True) -> np.ndarray:
    """Read a .raw as uint8 (nz, ny, nx), applying lbm3d-perm.cpp:174-193's checks."""
    path = Path(path)
    try:
        raw = path.read_bytes()
    except OSError as e:
        raise StructureError(f"cannot open {path}") from e
    if len(raw) < HEADER_BYTES:
        raise StructureError(f"cannot read header of {path}")
    nx, ny, nz = (int(v) for v in np.frombuffer(raw, dtype="<i4", count=3))
    if nx <= 0 or ny <= 0 or nz <= 0:
        raise StructureError(f"bad dimensions in {path}")
    n = nx * ny * nz
    if len(raw) < HEADER_BYTES + n:
        raise StructureError(f"truncated payload in {path}")
    if len(raw) > HEADER_BYTES + n:
        raise StructureError(f"trailing bytes in {path}")
    arr = np.frombuffer(raw, dtype=np.uint8, offset=HEADER_BYTES).reshape(nz, ny, nx)
    if check:
        check_mask3d(arr)
    return arr


def check_mask3d(solid: np.ndarray) -> int:
    """Return the pore count; raise like read_structure (:190-192) does."""
    if solid.ndim != 3 or min(solid.shape) <= 0:
        raise StructureError("bad dimensions")
    if solid.dtype != np.uint8:
        raise StructureError("mask must be uint8")
    if solid.size and solid.max() > 1:
        raise StructureError("non-binary voxel")
    n_pore = int(solid.size - np.count_nonzero(solid))
    if n_pore == 0:
        raise StructureError("structure has no pore space")
    return n_pore


def porosity(solid) -> float:
    return float(1.0 - np.asarray(solid).mean())


_ATOI = re.compile(r"[+-]?[0-9]+")


def atoi(token: str) -> int:
    """C atoi of one token as the reference's ``>>`` extraction yields it (no
    whitespace): an optional sign and the leading digits, 0 when there are none
    ("12abc" -> 12, "abc" -> 0, "0.5" -> 0, "-3" -> -3)."""
    m = _ATOI.match(token)
    return int(m.group(0)) if m else 0


def _ppm_dims(lines):
    """(lx, ly) from the first of header lines 2-3 that holds two integers"""
    for line in lines:
        tok = line.split()
        if len(tok) >= 2 and _ATOI.fullmatch(tok[0]) and _ATOI.fullmatch(tok[1]):
            return int(tok[0]), int(tok[1])
    return None


def load_structure2d(path, shape=None) -> np.ndarray:
    """Read a 2D structure as uint8 [y, x], 1 = solid (LMB2d/lbm.cpp:412-491).

    GIF (read_from_gif, :412-426): pore where the pixel's red channel is > 0.
    PPM (read_from_ppm, :427-454): four header lines are skipped, then LY rows
    of LX (R, G, B) tokens; pore where atoi(R) != 0.  DAT (read_from_dat,
    :456-470): LY rows of LX tokens, pore where atoi(token) == 0.  The file
    extension picks the reader (:472-483: ".ppm", ".gif", anything else DAT).

    The reference reads exactly LX*LY values of its compiled 256x256 lattice;
    here the PPM shape comes from its header (line 2, or line 3 when line 2 is
    a comment) and a DAT file is taken as square unless ``shape=(ly, lx)`` is
    given.  A file with fewer values than the lattice needs is an error here
    (the reference would read the missing tokens as empty strings: solid in a
    PPM, pore in a DAT).
    """
    path = str(path)
    if path.endswith(".gif"):
        from PIL import Image
        try:
            with Image.open(path) as im:
                red = np.array(im.convert("RGB"))[:, :, 0]
        except OSError as e:
            raise StructureError(f"cannot read {path}") from e
        return (red == 0).astype(np.uint8)
    try:
        with open(path) as fh:
            if path.endswith(".ppm"):
                head = [fh.readline() for _ in range(4)]
                toks = fh.read().split()
            else:
                head = None
                toks = fh.read().split()
    except OSError as e:
        raise StructureError(f"cannot open {path}") from e
    if head is not None:
        if shape is None:
            dims = _ppm_dims(head[1:3])
            if dims is None:
                raise StructureError(f"{path}: no dimensions in the PPM header")
            lx, ly = dims
        else:
            ly, lx = shape
        if lx <= 0 or ly <= 0:
            raise StructureError(f"bad dimensions in {path}")
        if len(toks) < 3 * lx * ly:
            raise StructureError(f"truncated payload in {path}")
        red = np.array([atoi(t) for t in toks[0:3 * lx * ly:3]], dtype=np.int64)
        return (red == 0).astype(np.uint8).reshape(ly, lx)
    if shape is None:
        n = int(round(np.sqrt(len(toks))))
        if n * n != len(toks):
            raise StructureError(f"{path}: DAT payload is not square; pass shape=(ly, lx)")
        shape = (n, n)
    if len(toks) < shape[0] * shape[1]:
        raise StructureError(f"truncated payload in {path}")
    vals = np.array([atoi(t) for t in toks[:shape[0] * shape[1]]], dtype=np.int64)
    return (vals != 0).astype(np.uint8).reshape(shape)


def save_gif2d(path, solid) -> Path:
    """Write a [y, x] {0,1} mask as a GIF the reference reader accepts (pore = 255)."""
    from PIL import Image
    solid = np.asarray(solid, dtype=np.uint8)
    img = np.where(solid == 0, 255, 0).astype(np.uint8)
    Image.fromarray(img, mode="L").save(path)
    return Path(path)
