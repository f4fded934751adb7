"""Writes tests/golden/interactions_subset.npz from the reference's shipped
example data (examples/interactions/*.fvi/.fvd, read as DatABEL files).

Kept: all 4000 rows; XL (4 covariates incl. intercept); the first 100 XR
columns and the first 50 Y columns -- the sizes the reference's own check
program scans (tests/test.cpp:199-200: limit_t = 50, limit_m = 100); every
INT* interaction file; also copies the dosage factor files
examples/dosages_{1,2}.txt.  Run once in a container that has the reference
mounted; the tests only read the .npz.
"""
import os
import struct
import sys

import numpy as np

SRC = "/root/reference/examples/interactions"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "interactions_subset.npz")


def read_fvi(path):
    with open(path, "rb") as f:
        h = f.read(48)
        typ = struct.unpack("<H", h[:2])[0]
        _, nobs, nvars, _, _, namelen = struct.unpack("<6I", h[4:28])
        names = f.read((nobs + nvars) * namelen)
    def nm(i):
        raw = names[i * namelen:(i + 1) * namelen]
        return raw.split(b"\0", 1)[0].decode()
    return typ, nobs, nvars, namelen, [nm(i) for i in range(nobs)], [nm(nobs + j) for j in range(nvars)]


def read(name, ncols=None):
    typ, nobs, nvars, namelen, obs, var = read_fvi(os.path.join(SRC, name + ".fvi"))
    assert typ == 5 and namelen == 32, (name, typ, namelen)
    data = np.fromfile(os.path.join(SRC, name + ".fvd"), dtype="<f4").reshape(nvars, nobs).T
    k = nvars if ncols is None else ncols
    return np.ascontiguousarray(data[:, :k]), obs, var[:k]


def main():
    if not os.path.isdir(SRC):
        sys.exit("reference examples not found at " + SRC)
    out = {}
    XL, obs, xl_names = read("XL")
    XR, obs_r, xr_names = read("XR", 100)
    Y, obs_y, y_names = read("Y", 50)
    assert obs == obs_r == obs_y
    out.update(XL=XL, XR=XR, Y=Y, obs_names=np.array(obs), xl_names=np.array(xl_names),
               xr_names=np.array(xr_names), y_names=np.array(y_names))
    for f in sorted(os.listdir(SRC)):
        if f.startswith("INT") and f.endswith(".fvd"):
            nm = f[:-4]
            E, obs_i, e_names = read(nm)
            assert obs_i == obs
            out["E_" + nm] = E
            out["Enames_" + nm] = np.array(e_names)
    np.savez_compressed(OUT, **out)
    print("wrote", OUT, os.path.getsize(OUT), "bytes")
    # the check program's dosage factor files (examples/dosages_{1,2}.txt),
    # data read by its dosage runs, kept byte for byte
    for k in (1, 2):
        src = os.path.join(os.path.dirname(SRC), "dosages_%d.txt" % k)
        dst = os.path.join(os.path.dirname(OUT), "dosages_%d.txt" % k)
        with open(src, "rb") as fi, open(dst, "wb") as fo:
            fo.write(fi.read())


if __name__ == "__main__":
    main()
