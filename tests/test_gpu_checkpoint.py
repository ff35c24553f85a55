"""Checkpoint / restart of the device state in the reference's byte format
(src/io.jl:6-75): a file written by the GPU solver reads back through the
numpy restatement of the format; a file written by that restatement loads into
the GPU solver; a restart from a checkpoint continues bitwise as the
uninterrupted run; mismatched files are refused with the reference's messages."""
import numpy as np
import pytest

import cases
import ckpt_format as K

pytestmark = pytest.mark.gpu
TWO_PI = 2 * np.pi


@pytest.fixture(scope="module")
def ctx():
    import compactles_hip as ch
    return ch.Context(0)


def _tgv(ctx, n=(24, 20, 16)):
    L = (TWO_PI,) * 3
    X, Y, Z = cases.grid(n, L)
    q = cases.conserved(cases.tgv_prim(X, Y, Z))
    return cases.gpu_solver(ctx, n, L, q, art=True)


def _interior(s):
    Q = s.allocate_state()
    s.download(Q)
    return np.stack([Q[(slice(s.halo[0], s.halo[0] + s.n[0]), slice(s.halo[1], s.halo[1] + s.n[1]),
                       slice(s.halo[2], s.halo[2] + s.n[2]), c)] for c in range(s.n_cons)], axis=-1)


def test_gpu_file_reads_with_reference_format(ctx, tmp_path):
    s = _tgv(ctx)
    s.run(1e9, 5)
    prefix = str(tmp_path / "tgv")
    s.save_checkpoint(prefix)
    hdr, Q = K.read(K.name(prefix))
    assert hdr["magic"] == K.MAGIC and hdr["n_cons"] == 5 and hdr["n_global"] == (24, 20, 16)
    assert hdr["n_local"] == (24, 20, 16) and hdr["coords"] == (0, 0, 0)
    assert hdr["t"] == s.t and hdr["step"] == 5
    assert np.array_equal(Q, _interior(s))


def test_reference_format_file_loads(ctx, tmp_path):
    s = _tgv(ctx)
    Q = _interior(s) * (1 + 1e-3 * np.random.default_rng(2).standard_normal((24, 20, 16, 5)))
    prefix = str(tmp_path / "ext")
    K.write(K.name(prefix), Q, 0.75, 17)
    s.load_checkpoint(prefix)
    assert s.t == 0.75 and s.step_count == 17
    assert np.array_equal(_interior(s), Q)


def test_restart_matches_oracle_restart(ctx, tmp_path):
    """A restart builds a fresh solver, as load_checkpoint! into a new Solver
    does in the reference: its lagged artificial coefficients start at zero,
    so it is compared with the oracle restarted from the same file, not with
    the uninterrupted run."""
    a = _tgv(ctx)
    a.run(1e9, 10)
    prefix = str(tmp_path / "mid")
    a.save_checkpoint(prefix)
    hdr, Q = K.read(K.name(prefix))
    b = _tgv(ctx)
    b.load_checkpoint(prefix)
    d_gpu = b.run(1e9, 30, record=20)
    n, L = (24, 20, 16), (TWO_PI,) * 3
    o = cases.oracle_solver(n, L, [Q[..., c] for c in range(5)], art=True)
    o.set_time(hdr["t"], hdr["step"])
    d_or = o.run(1e9, 30, record=20)
    assert b.step_count == o.time[1] == 30
    assert np.abs(d_gpu - d_or).max() <= 1e-12 * d_or.max()
    got = _interior(b)
    for c in range(5):
        ref = o.interior(o.Q[..., c])
        scale = np.abs(ref).max() if c != 3 else max(np.abs(o.interior(o.Q[..., k])).max() for k in (1, 2, 3))
        assert np.abs(got[..., c] - ref).max() <= 1e-10 * scale, c


@pytest.mark.parametrize("what,msg", [("magic", "not a CompactLES checkpoint"), ("ncons", "conserved layout mismatch"),
                                      ("grid", "global grid mismatch"), ("coords", "rank layout mismatch")])
def test_mismatch_is_refused(ctx, tmp_path, what, msg):
    import compactles_hip as ch
    s = _tgv(ctx)
    Q = _interior(s)
    prefix = str(tmp_path / what)
    if what == "ncons":
        K.write(K.name(prefix), np.concatenate([Q, Q[..., :1]], axis=-1), 0.0, 0)
    elif what == "grid":
        K.write(K.name(prefix), Q, 0.0, 0, n_global=(24, 20, 17))
    elif what == "coords":
        K.write(K.name(prefix), Q, 0.0, 0, coords=(1, 0, 0))
    else:
        K.write(K.name(prefix), Q, 0.0, 0)
        raw = bytearray(open(K.name(prefix), "rb").read())
        raw[0] ^= 0xFF
        open(K.name(prefix), "wb").write(bytes(raw))
    with pytest.raises(ch.ClesError) as e:
        s.load_checkpoint(prefix)
    assert msg in str(e.value)
