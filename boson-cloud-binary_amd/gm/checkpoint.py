"""checkpoint.bin of the reference (src/CrankNicolsonModes_restart.cpp:139-233),
read and written in plain Python (numpy), independently of the C++ writer.

Layout (little endian, no padding):
    char[8]   magic "GMCKPT01"
    int32[3]  step_id, Nr, n_modes
    float64[13] t, theta, phase, R, M1, M2, q, R1, R2, Omega, P, dt, p_bin
    uint64    size = Nr * n_modes
    complex128[size] psi (mode-major)
    rolling phase average (src/RollingPhase.cpp:169-184):
        float64 window_phase; float64[3] previous (time, phase, value);
        uint8 has_previous; uint8 ready; float64 initial_phase, integral,
        duration, average; uint64 n; n x float64[7] segments (time0, time1,
        phase0, phase1, value0, value1, integral)
"""
import os
import struct

import numpy as np

MAGIC = b"GMCKPT01"
SCALARS = ("t", "theta", "phase", "R", "M1", "M2", "q", "R1", "R2", "Omega", "P", "dt", "p_bin")


def read_checkpoint(path):
    with open(path, "rb") as f:
        b = f.read()
    if b[:8] != MAGIC:
        raise ValueError("not a checkpoint (magic)")
    step, Nr, K = struct.unpack_from("<3i", b, 8)
    vals = struct.unpack_from("<13d", b, 20)
    (size,) = struct.unpack_from("<Q", b, 124)
    if size != Nr * K:
        raise ValueError("state size does not match Nr * n_modes")
    o = 132
    psi = np.frombuffer(b, dtype="<c16", count=size, offset=o).copy()
    o += 16 * size
    rolling = {}
    rolling["window_phase"], = struct.unpack_from("<d", b, o)
    rolling["previous"] = struct.unpack_from("<3d", b, o + 8)
    rolling["has_previous"], rolling["ready"] = struct.unpack_from("<2B", b, o + 32)
    (rolling["initial_phase"], rolling["integral"], rolling["duration"],
     rolling["average"]) = struct.unpack_from("<4d", b, o + 34)
    (n,) = struct.unpack_from("<Q", b, o + 66)
    segs = np.frombuffer(b, dtype="<f8", count=7 * n, offset=o + 74).reshape(n, 7).copy()
    rolling["segments"] = segs
    if o + 74 + 56 * n != len(b):
        raise ValueError("trailing or missing bytes")
    return {"step_id": step, "Nr": Nr, "n_modes": K, "scalars": dict(zip(SCALARS, vals)), "psi": psi,
            "rolling": rolling}


def write_checkpoint(path, ck):
    """Write a dict as returned by read_checkpoint (to path.tmp, then rename)."""
    r = ck["rolling"]
    psi = np.ascontiguousarray(ck["psi"], dtype="<c16")
    parts = [MAGIC, struct.pack("<3i", ck["step_id"], ck["Nr"], ck["n_modes"]),
             struct.pack("<13d", *[ck["scalars"][k] for k in SCALARS]), struct.pack("<Q", psi.size), psi.tobytes(),
             struct.pack("<d", r["window_phase"]), struct.pack("<3d", *r["previous"]),
             struct.pack("<2B", int(r["has_previous"]), int(r["ready"])),
             struct.pack("<4d", r["initial_phase"], r["integral"], r["duration"], r["average"]),
             struct.pack("<Q", len(r["segments"])), np.ascontiguousarray(r["segments"], dtype="<f8").tobytes()]
    tmp = path + ".tmp"
    with open(tmp, "wb") as f:
        f.write(b"".join(parts))
    os.replace(tmp, path)
