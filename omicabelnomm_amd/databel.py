"""DatABEL filevector files (.fvi header + names, .fvd column-major data).

Header layout as the reference reads it (src/AIOwrapper.h:158-168): u16
type, 2 pad bytes, u32 nelements, numObservations, numVariables,
bytesPerRecord, bitsPerRecord, namelength, 5 reserved u32; then
(numObservations + numVariables) NUL-padded names, observations first.
Type codes: 1..8 = u16, i16, u32, i32, float, double, i8, u8.
"""
from __future__ import annotations

import struct

import numpy as np

TYPES = {1: "<u2", 2: "<i2", 3: "<u4", 4: "<i4", 5: "<f4", 6: "<f8", 7: "i1", 8: "u1"}
CODES = {np.dtype("<f4"): 5, np.dtype("<f8"): 6}


def write(prefix, data, obs_names=None, var_names=None, namelength=32):
    """data: (nobs, nvars) matrix; written column by column (.fvd)."""
    a = np.asarray(data)
    if a.dtype not in (np.float32, np.float64):
        a = a.astype(np.float32)
    a = a.astype(a.dtype.newbyteorder("<"))
    nobs, nvars = a.shape
    obs_names = obs_names if obs_names is not None else ["id%d" % (i + 1) for i in range(nobs)]
    var_names = var_names if var_names is not None else ["v%d" % (j + 1) for j in range(nvars)]
    code = CODES[a.dtype]
    es = a.dtype.itemsize
    with open(prefix + ".fvi", "wb") as f:
        f.write(struct.pack("<HH6I5I", code, 0, nobs * nvars, nobs, nvars, es, es * 8, namelength,
                            0, 0, 0, 0, 0))
        for nm in list(obs_names) + list(var_names):
            b = str(nm).encode()[: namelength - 1]
            f.write(b + b"\0" * (namelength - len(b)))
    np.asfortranarray(a).T.tofile(prefix + ".fvd")


def read(prefix):
    """(data (nobs, nvars) float64, obs_names, var_names, header dict)."""
    with open(prefix + ".fvi", "rb") as f:
        h = f.read(48)
        code = struct.unpack("<H", h[:2])[0]
        nel, nobs, nvars, bpr, bits, namelen = struct.unpack("<6I", h[4:28])
        raw = f.read((nobs + nvars) * namelen)
    names = [raw[i * namelen:(i + 1) * namelen].split(b"\0", 1)[0].decode() for i in range(nobs + nvars)]
    data = np.fromfile(prefix + ".fvd", dtype=TYPES[code]).reshape(nvars, nobs).T.astype(np.float64)
    hdr = dict(type=code, nelements=nel, nobs=nobs, nvars=nvars, namelength=namelen)
    return data, names[:nobs], names[nobs:], hdr
