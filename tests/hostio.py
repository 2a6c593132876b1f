"""Helpers that write and read the host layer's files independently of it
(DCD with numpy, raw signal files), for the host tests."""
import os
import struct
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SX_RUN = os.path.join(ROOT, "sassena_amd", "bin", "sx_run")


def write_dcd(path, frames, unit_cell=False):
    """CHARMM DCD: 84-byte header record, title record, NATOM record, then per
    frame an optional 48-byte unit-cell record and x, y, z float records."""
    NF, NA, _ = frames.shape

    def rec(f, payload):
        f.write(struct.pack("<i", len(payload)))
        f.write(payload)
        f.write(struct.pack("<i", len(payload)))

    with open(path, "wb") as f:
        icntrl = [0] * 20
        icntrl[0] = NF
        icntrl[2] = 1
        icntrl[10] = 1 if unit_cell else 0
        icntrl[19] = 24
        rec(f, b"CORD" + struct.pack("<20i", *icntrl))
        rec(f, struct.pack("<i", 1) + b"test".ljust(80))
        rec(f, struct.pack("<i", NA))
        for t in range(NF):
            if unit_cell:
                rec(f, struct.pack("<6d", 30.0, 0.0, 30.0, 0.0, 0.0, 30.0))
            for c in range(3):
                rec(f, frames[t, :, c].astype("<f4").tobytes())


def write_raw_trajectory(path, frames):
    NF, NA, _ = frames.shape
    with open(path, "wb") as f:
        f.write(b"SXRAW1" + struct.pack("<QQ", NA, NF))
        f.write(np.ascontiguousarray(frames, dtype="<f4").tobytes())


def read_raw_trajectory(path):
    with open(path, "rb") as f:
        assert f.read(6) == b"SXRAW1"
        NA, NF = struct.unpack("<QQ", f.read(16))
        return np.frombuffer(f.read(), dtype="<f4").reshape(NF, NA, 3)


def read_raw_signal(path):
    """-> (q [nq,3], fqt [nq,NF] complex, fq [nq], fq2 [nq])"""
    with open(path, "rb") as f:
        assert f.read(6) == b"SXSIG1"
        nq, NF = struct.unpack("<QQ", f.read(16))
        rec = np.frombuffer(f.read(), dtype="<f8").reshape(nq, 3 + 2 * NF + 4)
    q = rec[:, :3]
    fqt = rec[:, 3:3 + 2 * NF:2] + 1j * rec[:, 4:4 + 2 * NF:2]
    fq = rec[:, 3 + 2 * NF] + 1j * rec[:, 4 + 2 * NF]
    fq2 = rec[:, 5 + 2 * NF] + 1j * rec[:, 6 + 2 * NF]
    return q, fqt, fq, fq2


def sx_run(*args, cwd=None, check=True):
    res = subprocess.run([SX_RUN, *args], capture_output=True, text=True, cwd=cwd, timeout=600)
    if check and res.returncode != 0:
        raise RuntimeError(f"sx_run {' '.join(args)} failed ({res.returncode}): {res.stderr}")
    return res
