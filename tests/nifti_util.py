"""A minimal NIfTI-1 writer for the tests (the single-file "n+1" layout: 348-byte header, 4 bytes
of extension flags, data from byte 352)."""
import gzip
import struct

import numpy as np

CODES = {np.dtype(np.uint8): 2, np.dtype(np.int16): 4, np.dtype(np.int32): 8, np.dtype(np.float32): 16,
         np.dtype(np.float64): 64, np.dtype(np.int8): 256, np.dtype(np.uint16): 512}


def write(path, data, vox=(1.0, 1.0, 1.0), sform=None, quatern=None, qoffset=(0.0, 0.0, 0.0), qfac=1.0,
          slope=0.0, inter=0.0, sizeof_hdr=348):
    """data: array [x, y, z, ...] (axis 0 fastest in the file); sform: 3x4 rows (sets sform_code 1);
    quatern: (b, c, d) (sets qform_code 1)."""
    data = np.asarray(data)
    hdr = bytearray(348)
    struct.pack_into("<i", hdr, 0, sizeof_hdr)
    dims = [data.ndim] + list(data.shape) + [1] * (7 - data.ndim)
    struct.pack_into("<8h", hdr, 40, *dims)
    struct.pack_into("<h", hdr, 70, CODES[data.dtype])
    struct.pack_into("<h", hdr, 72, data.dtype.itemsize * 8)
    pix = [qfac] + list(vox) + [1.0] * (7 - len(vox))
    struct.pack_into("<8f", hdr, 76, *pix)
    struct.pack_into("<f", hdr, 108, 352.0)
    struct.pack_into("<f", hdr, 112, slope)
    struct.pack_into("<f", hdr, 116, inter)
    if quatern is not None:
        struct.pack_into("<h", hdr, 252, 1)
        struct.pack_into("<3f", hdr, 256, *quatern)
        struct.pack_into("<3f", hdr, 268, *qoffset)
    if sform is not None:
        struct.pack_into("<h", hdr, 254, 1)
        for r in range(3):
            struct.pack_into("<4f", hdr, 280 + 16 * r, *[float(v) for v in sform[r]])
    hdr[344:348] = b"n+1\0"
    raw = bytes(hdr) + b"\0\0\0\0" + np.asfortranarray(data).tobytes(order="F")
    opener = gzip.open if path.endswith(".gz") else open
    with opener(path, "wb") as f:
        f.write(raw)
