"""Raw float32 model and data files (the Madagascar-style binaries the reference's demos load,
demos/test_fwi_vel2d.py:11, through pyseistr's binread) and their inverse.

A file holds n1*n2*n3 little-endian float32 samples with the first axis fastest (Fortran order),
which is also the layout every pywave_amd entry point takes (z fastest).  Host-side helpers
only: the engine itself takes arrays.
"""
import numpy as np

__all__ = ["binread", "binwrite"]


def binread(fname, n1, n2=1, n3=1):
    """Read n1*n2*n3 float32 samples, first axis fastest; returns (n1,), (n1, n2) or (n1, n2, n3)."""
    n1, n2, n3 = int(n1), int(n2), int(n3)
    data = np.fromfile(fname, dtype="<f4")
    if data.size != n1 * n2 * n3:
        raise ValueError("%s holds %d float32 samples, expected %d x %d x %d = %d"
                         % (fname, data.size, n1, n2, n3, n1 * n2 * n3))
    shape = (n1,) if n2 == 1 and n3 == 1 else ((n1, n2) if n3 == 1 else (n1, n2, n3))
    return data.reshape(shape, order="F")


def binwrite(fname, data):
    """Write an array as float32 samples, first axis fastest (the inverse of binread)."""
    np.asarray(data, dtype="<f4").ravel(order="F").tofile(fname)
