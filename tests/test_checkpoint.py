"""checkpoint.bin layout (src/CrankNicolsonModes_restart.cpp:139-233): the
Python reader/writer against byte offsets fixed by the reference's writer."""
import struct

import numpy as np

from gm import checkpoint


def sample(Nr=5, K=3, nseg=2):
    rng = np.random.default_rng(1)
    psi = rng.standard_normal(Nr * K) + 1j * rng.standard_normal(Nr * K)
    return {"step_id": 7, "Nr": Nr, "n_modes": K,
            "scalars": {k: float(i + 0.5) for i, k in enumerate(checkpoint.SCALARS)}, "psi": psi,
            "rolling": {"window_phase": 12.566, "previous": (1.0, 2.0, 3.0), "has_previous": 1, "ready": 0,
                        "initial_phase": 0.25, "integral": 4.5, "duration": 6.0, "average": 0.75,
                        "segments": rng.standard_normal((nseg, 7))}}


def test_byte_layout_matches_the_reference_writer(tmp_path):
    ck = sample()
    p = tmp_path / "checkpoint.bin"
    checkpoint.write_checkpoint(str(p), ck)
    b = p.read_bytes()
    # restart.cpp:153-163: magic, int32 header, 13 doubles, uint64 size, psi, average
    assert b[:8] == b"GMCKPT01"
    assert struct.unpack_from("<3i", b, 8) == (7, 5, 3)
    assert struct.unpack_from("<13d", b, 20) == tuple(i + 0.5 for i in range(13))
    assert struct.unpack_from("<Q", b, 124) == (15,)
    assert np.array_equal(np.frombuffer(b, "<c16", 15, 132), ck["psi"])
    o = 132 + 16 * 15
    # RollingPhase.cpp:169-184: doubles and two one-byte bools, no padding
    assert struct.unpack_from("<d3d2B4dQ", b, o)[:4] == (12.566, 1.0, 2.0, 3.0)
    assert struct.unpack_from("<2B", b, o + 32) == (1, 0)
    assert struct.unpack_from("<Q", b, o + 66) == (2,)
    assert len(b) == o + 74 + 2 * 56
    assert not (tmp_path / "checkpoint.bin.tmp").exists()


def test_round_trip(tmp_path):
    ck = sample(Nr=33, K=6, nseg=5)
    p = str(tmp_path / "c.bin")
    checkpoint.write_checkpoint(p, ck)
    back = checkpoint.read_checkpoint(p)
    assert back["step_id"] == 7 and back["Nr"] == 33 and back["n_modes"] == 6
    assert back["scalars"] == ck["scalars"]
    assert np.array_equal(back["psi"], ck["psi"])
    assert np.array_equal(back["rolling"]["segments"], ck["rolling"]["segments"])
    assert back["rolling"]["previous"] == (1.0, 2.0, 3.0)
