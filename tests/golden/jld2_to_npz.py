"""Convert the reference's JLD2 regression fixtures into NumPy ``.npz`` files.

The four files ``test/version_update_test_verification/v1_*_MISMIP_100yr_output_
8kmres_maxiter1_timesteppt1.jld2`` of WAVI.jl are HDF5 files that hold six
Float64 datasets each (h, u, v, viscosity, grounded_fraction, bed_speed).  There
is no h5py in this image, so this script reads the small subset of HDF5 these
files use: superblock version 2, version-2 object headers, link messages in the
root group, and compact or contiguous data layout.  It only reads data; nothing
in the files is executed.

Julia stores arrays column-major; the datasets' HDF5 dimensions are therefore
the Julia shape reversed.  The arrays are written here in Julia's (nx, ny)
shape with Fortran order preserved (``x`` index fastest).

Usage: python tests/golden/jld2_to_npz.py /root/reference/test/version_update_test_verification
"""
import os
import struct
import sys

import numpy as np


def _u(b, off, n):
    return int.from_bytes(b[off:off + n], "little")


class _H5:
    def __init__(self, data):
        self.b = data
        sb = data.find(b"\x89HDF\r\n\x1a\n")
        if sb < 0 or data[sb + 8] not in (2, 3):
            raise ValueError("only HDF5 superblock v2/v3 is supported")
        self.so = data[sb + 9]
        self.sl = data[sb + 10]
        off = sb + 12
        self.base = _u(data, off, self.so)
        off += self.so * 2  # base, extension
        off += self.so      # end of file
        self.root = _u(data, off, self.so)

    def messages(self, addr):
        b = self.b
        p = self.base + addr
        if b[p:p + 4] != b"OHDR":
            raise ValueError("expected OHDR at %d" % p)
        flags = b[p + 5]
        q = p + 6
        if flags & 0x20:
            q += 16
        if flags & 0x10:
            q += 4
        csz = 1 << (flags & 3)
        chunk = _u(b, q, csz)
        q += csz
        end = q + chunk
        out = []
        while q + 4 <= end:
            mtype = b[q]
            msize = _u(b, q + 1, 2)
            mflags = b[q + 3]
            q += 4
            if flags & 0x04:
                q += 2
            out.append((mtype, b[q:q + msize]))
            q += msize
        return out

    def links(self):
        res = {}
        for mtype, m in self.messages(self.root):
            if mtype != 6:
                continue
            flags = m[1]
            q = 2
            ltype = 0
            if flags & 0x08:
                ltype = m[q]
                q += 1
            if flags & 0x04:
                q += 8
            if flags & 0x10:
                q += 1
            lsz = 1 << (flags & 3)
            nlen = _u(m, q, lsz)
            q += lsz
            name = m[q:q + nlen].decode()
            q += nlen
            if ltype == 0:
                res[name] = _u(m, q, self.so)
        return res

    def dataset(self, addr):
        dims = None
        dtype = None
        raw = None
        for mtype, m in self.messages(addr):
            if mtype == 1:  # dataspace
                ver, rank, fl = m[0], m[1], m[2]
                q = 4 if ver == 2 else 8
                dims = [_u(m, q + 8 * k, self.sl) for k in range(rank)]
            elif mtype == 3:  # datatype
                cls = m[0] & 0x0F
                size = _u(m, 4, 4)
                if cls != 1 or size != 8:
                    raise ValueError("only Float64 datasets are supported")
                dtype = np.dtype("<f8")
            elif mtype == 8:  # data layout
                ver = m[0]
                cls = m[1]
                if ver not in (3, 4):
                    raise ValueError("layout version %d unsupported" % ver)
                if cls == 0:
                    n = _u(m, 2, 2)
                    raw = m[4:4 + n]
                elif cls == 1:
                    a = _u(m, 2, self.so)
                    n = _u(m, 2 + self.so, self.sl)
                    raw = self.b[self.base + a:self.base + a + n]
                else:
                    raise ValueError("chunked layout unsupported")
        arr = np.frombuffer(raw, dtype=dtype).copy()
        # HDF5 dims are row-major of the reversed Julia shape.
        jshape = tuple(reversed(dims))
        return arr.reshape(jshape, order="F")


def convert(path):
    h = _H5(open(path, "rb").read())
    return {name: h.dataset(a) for name, a in h.links().items()}


def main(src):
    here = os.path.dirname(os.path.abspath(__file__))
    for fn in sorted(os.listdir(src)):
        if not fn.endswith(".jld2"):
            continue
        tag = fn.split("_MISMIP")[0]  # e.g. v1_6_2
        arrays = convert(os.path.join(src, fn))
        out = os.path.join(here, "mismip_plus_%s.npz" % tag)
        np.savez(out, **arrays)
        print(out, {k: v.shape for k, v in arrays.items()})


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else
         "/root/reference/test/version_update_test_verification")
