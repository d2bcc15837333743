"""
Structure files: the 3D ``.raw`` format and the 2D GIF / PPM / DAT inputs.

3D ``.raw`` (scripts/structure_io3d.py:40-67, LMB3d/lbm3d-perm.cpp:174-193)::

    bytes 0..11   three little-endian int32: nx, ny, nz
    bytes 12..    nx*ny*nz uint8, 1 = solid, 0 = pore, x varying fastest

Arrays are indexed ``[z, y, x]``.  The checks of the reference reader are kept:
bad dimensions, truncated payload, trailing bytes, non-binary voxels and a
mask with no pore space are all errors (``StructureError``).

2D inputs follow LMB2d/lbm.cpp:412-483: a GIF pixel is pore when its red
channel is > 0 (``read_from_gif``), a PPM cell is pore when its R value is
non-zero, a DAT cell is pore when it is 0.  Arrays are indexed ``[y, x]``
(row = solver y, column = solver x, scripts/structure_io.py:44-54).  The GIF
decode uses PIL instead of Magick++.
"""
from __future__ import annotations

from pathlib import Path

import numpy as np

__all__ = ["StructureError", "save_raw", "load_raw", "check_mask3d", "porosity",
           "load_structure2d", "save_gif2d"]

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


def load_structure2d(path, shape=None) -> np.ndarray:
    """Read a 2D structure as uint8 [y, x], 1 = solid (lbm.cpp:412-491 rules).

    GIF: pore where red > 0.  PPM (P3): four header lines are skipped as the
    reference does, then LY rows of LX (R, G, B) triples; pore where R != 0.
    DAT: LY rows of LX integers, pore where the value is 0.  The reference
    reads exactly LX*LY values of its compiled 256x256 lattice; here the PPM
    shape comes from its header and a DAT file is taken as square unless
    ``shape=(ly, lx)`` is given.
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
    if path.endswith(".ppm"):
        with open(path) as fh:
            fh.readline()
            dims = fh.readline().split()
            fh.readline()
            fh.readline()
            vals = fh.read().split()
        lx, ly = (int(dims[0]), int(dims[1])) if shape is None else (shape[1], shape[0])
        if len(vals) < 3 * lx * ly:
            raise StructureError(f"truncated payload in {path}")
        r = np.array([int(v) for v in vals[0:3 * lx * ly:3]], dtype=np.int64)
        return (r == 0).astype(np.uint8).reshape(ly, lx)
    with open(path) as fh:
        vals = np.array([int(v) for v in fh.read().split()], dtype=np.int64)
    if shape is None:
        n = int(round(np.sqrt(vals.size)))
        if n * n != vals.size:
            raise StructureError(f"{path}: DAT payload is not square; pass shape=(ly, lx)")
        shape = (n, n)
    if vals.size < shape[0] * shape[1]:
        raise StructureError(f"truncated payload in {path}")
    return (vals[:shape[0] * shape[1]] != 0).astype(np.uint8).reshape(shape)


def save_gif2d(path, solid) -> Path:
    """Write a [y, x] {0,1} mask as a GIF the reference reader accepts (pore = 255)."""
    from PIL import Image
    solid = np.asarray(solid, dtype=np.uint8)
    img = np.where(solid == 0, 255, 0).astype(np.uint8)
    Image.fromarray(img, mode="L").save(path)
    return Path(path)
