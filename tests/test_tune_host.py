"""Line-tile table host logic (no GPU is touched): the thread-count rule for
untuned shapes, table save / load round trips, keys that separate grids,
precisions and boundary kinds, and malformed tables refused."""
import numpy as np
import pytest



def _spec(ch, n, precision="fp64", periodic=True):
    bc = ch.PeriodicBC() if periodic else ch.SlipWallBC()
    prob = ch.Problem(domain=((0.0, 2 * np.pi),) * 3, bcs=((bc, bc),) * 3,
                      ic=lambda X, Y, Z: ch.Prim(u=(0 * X, 0 * X, 0 * X), p=1.0, rho=1.0))
    return ch.SolverSpec(prob, ch.Numerics(n_global=n, precision=precision))


@pytest.fixture(autouse=True)
def _empty_table():
    import compactles_hip as ch
    ch.tune.clear()
    yield
    ch.tune.clear()


def _rule(dim, n_line, lines):
    """The documented rule restated: the largest chunk (x <= 16, y/z <= 32,
    >= 8) whose launch has at least 131072 threads."""
    c = 16 if dim == 0 else 32
    while c >= 8:
        if lines * -(-n_line // c) >= 131072:
            return c
        c //= 2
    return 8


@pytest.mark.parametrize("n", [(32, 32, 32), (64, 64, 64), (128, 128, 128), (256, 256, 256), (1024, 512, 1),
                               (256, 12, 10), (48, 200, 1)])
def test_untuned_rule(n):
    import compactles_hip as ch
    chunks, tuned = ch.tune.lookup(_spec(ch, n))
    N = n[0] * n[1] * n[2]
    want = tuple(_rule(d, n[d], N // n[d]) if n[d] > 1 else 0 for d in range(3))
    assert chunks == want and tuned == (False, False, False)


def test_rule_keeps_the_large_grid_tiles():
    # 256^3 keeps the 16-point x chunks of the specialised 256-point x kernel and 32-point y/z chunks
    import compactles_hip as ch
    assert ch.tune.lookup(_spec(ch, (256, 256, 256)))[0] == (16, 32, 32)
    assert ch.tune.lookup(_spec(ch, (32, 32, 32)))[0] == (8, 8, 8)


def test_save_load_round_trip(tmp_path):
    import compactles_hip as ch
    a, b = _spec(ch, (64, 64, 64)), _spec(ch, (64, 64, 64), precision="fp32")
    c = _spec(ch, (64, 64, 64), periodic=False)
    path = tmp_path / "tiles.txt"
    path.write_text("# header\n1 64 64 64 0 1 1 2 4 4\n2 64 64 64 0 1 1 2 4 16\n3 64 64 64 0 1 1 2 4 8\n")
    ch.tune.load(path)
    assert ch.tune.lookup(a) == ((4, 16, 8), (True, True, True))
    # other precision, other boundary kind: not the same key
    assert ch.tune.lookup(b)[1] == (False, False, False)
    assert ch.tune.lookup(c)[1] == (False, False, False)
    out = tmp_path / "saved.txt"
    ch.tune.save(out)
    ch.tune.clear()
    assert ch.tune.lookup(a)[1] == (False, False, False)
    ch.tune.load(out)
    assert ch.tune.lookup(a) == ((4, 16, 8), (True, True, True))
    rows = [l for l in out.read_text().splitlines() if not l.startswith("#")]
    assert sorted(rows) == sorted(["1 64 64 64 0 1 1 2 4 4", "2 64 64 64 0 1 1 2 4 16", "3 64 64 64 0 1 1 2 4 8"])


def test_malformed_table_refused(tmp_path):
    import compactles_hip as ch
    p = tmp_path / "bad.txt"
    p.write_text("1 64 64 64 0 1 1 2 4\n")
    with pytest.raises(ch.ClesError, match="malformed"):
        ch.tune.load(p)
    p.write_text("4 64 64 64 0 1 1 2 4 8\n")
    with pytest.raises(ch.ClesError, match="malformed"):
        ch.tune.load(p)
    with pytest.raises(ch.ClesError, match="cannot open"):
        ch.tune.load(tmp_path / "missing.txt")
