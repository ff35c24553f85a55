"""The checkpoint byte format of src/io.jl:6-75 restated in numpy (a reader
and a writer of one rank's file), used to check the C ABI's files."""
import numpy as np

MAGIC = 0x434c4553434b5054   # "CLES CKPT"


def name(prefix, rank=0):
    """io.jl:8-9 _ckpt_name."""
    return f"{prefix}.r{rank:04d}.ckpt"


def write(path, Q_interior, t, step, n_global=None, coords=(0, 0, 0)):
    """Q_interior: (nx, ny, nz, n_cons) float64."""
    Q_interior = np.asarray(Q_interior, dtype="<f8")
    nx, ny, nz, nc = Q_interior.shape
    n_global = n_global or (nx, ny, nz)
    with open(path, "wb") as f:
        f.write(np.array([MAGIC], dtype="<u8").tobytes())
        f.write(np.array([nc, *n_global, nx, ny, nz, *coords], dtype="<i8").tobytes())
        f.write(np.array([t], dtype="<f8").tobytes())
        f.write(np.array([step], dtype="<i8").tobytes())
        f.write(np.asfortranarray(Q_interior).tobytes(order="F"))


def read(path):
    """(header dict, Q interior (nx, ny, nz, n_cons))."""
    raw = open(path, "rb").read()
    magic = np.frombuffer(raw[:8], dtype="<u8")[0]
    ints = np.frombuffer(raw[8:88], dtype="<i8")
    t = np.frombuffer(raw[88:96], dtype="<f8")[0]
    step = np.frombuffer(raw[96:104], dtype="<i8")[0]
    nc, ng, nl, co = int(ints[0]), tuple(int(x) for x in ints[1:4]), tuple(int(x) for x in ints[4:7]), \
        tuple(int(x) for x in ints[7:10])
    data = np.frombuffer(raw[104:], dtype="<f8")
    assert data.size == nl[0] * nl[1] * nl[2] * nc
    Q = data.reshape(nl + (nc,), order="F")
    return {"magic": int(magic), "n_cons": nc, "n_global": ng, "n_local": nl, "coords": co, "t": float(t),
            "step": int(step)}, Q
