"""PETSc binary Vec files, the format of the reference's g%d.bin,
g%d%d.bin and x%d%d-fft.bin (bgy3d_vec_save / bgy3d_vec_read,
bgy3d-vec.c:99-138: VecView to a binary viewer).

A file is the big-endian int32 VEC_FILE_CLASSID 1211214, the big-endian
int32 number of scalars, then the scalars as big-endian doubles in the
DMDA's natural order (x fastest).  The reference's "complex" Vecs are real
Vecs with two degrees of freedom per point, so x%d%d-fft.bin holds
[N2][N1][N0/2+1][2] doubles.
"""
import numpy as np

VEC_FILE_CLASSID = 1211214


def write_vec(path, array):
    a = np.ascontiguousarray(array, dtype=np.float64).ravel()
    with open(path, "wb") as f:
        f.write(np.array([VEC_FILE_CLASSID, a.size], dtype=">i4").tobytes())
        f.write(a.astype(">f8").tobytes())


def read_vec(path, shape=None):
    with open(path, "rb") as f:
        head = np.frombuffer(f.read(8), dtype=">i4")
        if head[0] != VEC_FILE_CLASSID:
            raise ValueError(f"{path}: not a PETSc Vec file (class id {head[0]})")
        n = int(head[1])
        data = np.frombuffer(f.read(8 * n), dtype=">f8")
    if data.size != n:
        raise ValueError(f"{path}: truncated, {data.size} of {n} scalars")
    data = data.astype(np.float64)
    return data.reshape(shape) if shape is not None else data
