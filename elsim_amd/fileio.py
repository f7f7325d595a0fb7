"""Fortran unformatted sequential records, byte-compatible with the
reference's READ(20)/WRITE(20) of whole arrays (e.g.
SBMES_3D_MPI_CH_Gr_CONJG_VD_T05_01_BQ_hx3.f90:151, 914; LevelSet_3D.f90:365;
AccDiff_3D.f90:166): [int32 nbytes][payload][int32 nbytes], array payload in
column-major order; and the output file names of OutPut3D_A (SBM:901-911).
"""
import struct

import numpy as np


def write_record(path, a, dtype=None):
    """Write one array as one record (Fortran order)."""
    a = np.asarray(a)
    if dtype is not None:
        a = a.astype(dtype)
    payload = np.asfortranarray(a).ravel(order="F").tobytes()
    n = len(payload)
    if n >= 2 ** 31:
        raise ValueError("record larger than 2 GiB needs subrecords")
    with open(path, "wb") as f:
        f.write(struct.pack("<i", n))
        f.write(payload)
        f.write(struct.pack("<i", n))


def read_record(path, shape, dtype):
    """Read the first record of a file as a Fortran-ordered array."""
    with open(path, "rb") as f:
        (n,) = struct.unpack("<i", f.read(4))
        payload = f.read(n)
        (m,) = struct.unpack("<i", f.read(4))
    if n != m:
        raise ValueError(f"corrupt record markers {n} != {m}")
    a = np.frombuffer(payload, dtype=dtype)
    if a.size != int(np.prod(shape)):
        raise ValueError(f"record holds {a.size} values, expected {shape}")
    return a.reshape(shape, order="F")


def output_name(prefix, rank, frame):
    """OutPut3D_A file name: FF + I4(1000+rank) + I4(1000+frame) + '.dat'."""
    return f"{prefix}{1000 + rank:4d}{1000 + frame:4d}.dat"
