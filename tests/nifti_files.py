"""Writes NIfTI-1 / NIfTI-2 / Analyze test files from the published header
layouts (nifti1.h / nifti2.h field offsets), independently of both readers
under test.  Test helper only."""
import gzip
import struct

import numpy as np

DTYPES = {2: np.uint8, 4: np.int16, 8: np.int32, 16: np.float32, 64: np.float64,
          256: np.int8, 512: np.uint16, 768: np.uint32, 1024: np.int64, 1280: np.uint64}


def _nifti1_header(dim, datatype, pixdim, vox_offset, scl, qform, sform, quat, qoff, srow,
                   magic, endian):
    e = endian
    h = bytearray(348)
    struct.pack_into(e + "i", h, 0, 348)
    struct.pack_into(e + "8h", h, 40, *dim)
    bitpix = np.dtype(DTYPES[datatype]).itemsize * 8
    struct.pack_into(e + "hh", h, 70, datatype, bitpix)
    struct.pack_into(e + "8f", h, 76, *pixdim)
    struct.pack_into(e + "fff", h, 108, vox_offset, scl[0], scl[1])
    struct.pack_into(e + "hh", h, 252, qform, sform)
    struct.pack_into(e + "6f", h, 256, *quat, *qoff)
    struct.pack_into(e + "12f", h, 280, *srow)
    h[344:348] = magic
    return bytes(h)


def _nifti2_header(dim, datatype, pixdim, vox_offset, scl, qform, sform, quat, qoff, srow,
                   endian):
    e = endian
    h = bytearray(540)
    struct.pack_into(e + "i", h, 0, 540)
    h[4:12] = b"n+2\0\r\n\x1a\n"
    bitpix = np.dtype(DTYPES[datatype]).itemsize * 8
    struct.pack_into(e + "hh", h, 12, datatype, bitpix)
    struct.pack_into(e + "8q", h, 16, *dim)
    struct.pack_into(e + "8d", h, 104, *pixdim)
    struct.pack_into(e + "q", h, 168, vox_offset)
    struct.pack_into(e + "dd", h, 176, scl[0], scl[1])
    struct.pack_into(e + "ii", h, 344, qform, sform)
    struct.pack_into(e + "6d", h, 352, *quat, *qoff)
    struct.pack_into(e + "12d", h, 400, *srow)
    return bytes(h)


def write(path, vol, datatype=16, vs=(1.0, 1.0, 1.0), srow=None, version=1, endian="<",
          scl=(0.0, 0.0), qform=0, sform=1, quat=(0.0, 0.0, 0.0), qoff=(0.0, 0.0, 0.0),
          qfac=1.0, gz=False):
    """vol: (t, z, y, x) or (z, y, x) array of voxel values; written as
    DTYPES[datatype] in the given byte order.  version 0 writes an Analyze
    .hdr/.img pair (path is the .hdr)."""
    vol = np.asarray(vol)
    if vol.ndim == 3:
        vol = vol[None]
    t, z, y, x = vol.shape
    dim = [4 if t > 1 else 3, x, y, z, t, 1, 1, 1]
    pixdim = [qfac, vs[0], vs[1], vs[2], 1.0, 1.0, 1.0, 1.0]
    if srow is None:
        srow = [-vs[0], 0, 0, vs[0] * (x - 1), 0, -vs[1], 0, vs[1] * (y - 1), 0, 0, vs[2], 0]
    data = vol.astype(np.dtype(DTYPES[datatype]).newbyteorder(endian)).tobytes()
    if version == 2:
        head = _nifti2_header(dim, datatype, pixdim, 544, scl, qform, sform, quat, qoff, srow,
                              endian) + b"\0" * 4
    elif version == 1:
        head = _nifti1_header(dim, datatype, pixdim, 352.0, scl, qform, sform, quat, qoff, srow,
                              b"n+1\0", endian) + b"\0" * 4
    else:
        hdr = _nifti1_header(dim, datatype, pixdim, 0.0, scl, qform, sform, quat, qoff, srow,
                             b"ni1\0", endian)
        with open(path, "wb") as f:
            f.write(hdr)
        with open(str(path)[:-4] + ".img", "wb") as f:
            f.write(data)
        return
    blob = head + data
    if gz:
        with gzip.open(path, "wb") as f:
            f.write(blob)
    else:
        with open(path, "wb") as f:
            f.write(blob)
