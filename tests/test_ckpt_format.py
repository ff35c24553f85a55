"""The numpy restatement of the checkpoint format against the byte layout
spelled out from src/io.jl:21-37 (CPU)."""
import struct

import numpy as np

import ckpt_format as K


def test_header_layout_and_round_trip(tmp_path):
    Q = np.random.default_rng(1).standard_normal((5, 4, 3, 6))
    p = str(tmp_path / "x.r0000.ckpt")
    K.write(p, Q, 0.125, 42)
    raw = open(p, "rb").read()
    # UInt64 magic, then Int64 n_cons, n_global[3], n_local[3], coords[3], Float64 t, Int64 step
    fields = struct.unpack("<Q10qdq", raw[:104])
    assert fields[0] == 0x434c4553434b5054 and raw[:8] == bytes.fromhex("54504b4353454c43")
    assert fields[1:11] == (6, 5, 4, 3, 5, 4, 3, 0, 0, 0) and fields[11] == 0.125 and fields[12] == 42
    # Q[o+1:o+nx, ..., :] written column-major: x fastest, components last
    first = struct.unpack("<3d", raw[104:128])
    assert first == (Q[0, 0, 0, 0], Q[1, 0, 0, 0], Q[2, 0, 0, 0])
    assert len(raw) == 104 + Q.size * 8
    hdr, Q2 = K.read(p)
    assert hdr["t"] == 0.125 and hdr["step"] == 42 and np.array_equal(Q, Q2)
    assert K.name("run", 3) == "run.r0003.ckpt"
