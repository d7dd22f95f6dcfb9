"""Host side of the case set-ups (no GPU; the geometry builder is host C in the library).
The C1 and three-level layouts have the sizes, relaxation times and node-update counts the
case drivers quote, and the seeded 256-case sweep of uneven two-level cylinder channels has
the configuration it is quoted on. Re ~ U[40, 200]; D in {10, 16, 24,
32, 40} finest cells (ratio 2: D/2 coarse); blockage D/H ~ U[0.1, 0.3] (H rounded to whole
coarse cells); channel length 20 D; refined block 4 D x 2 D with the cylinder one diameter
behind its upstream edge, on the channel's centre line."""
import numpy as np

import mbl


def test_c4_sweep_configuration():
    specs = mbl.c4_specs(256, seed=2024)
    assert len(specs) == 256
    fine_d = np.array([4.0 * s.R for s in specs])  # D = 2R coarse cells = 4R finest
    assert set(np.round(fine_d).astype(int)) == {10, 16, 24, 32, 40}
    re = np.array([s.Re for s in specs])
    assert re.min() >= 40.0 and re.max() <= 200.0 and re.max() - re.min() > 140.0
    for s in specs:
        dc = int(round(2.0 * s.R))
        h = s.LY - 1
        assert 0.1 - 0.01 <= dc / h <= 0.3 + 0.01
        assert s.LX - 1 == 20 * dc
        r = s.region
        assert (r.xend - r.xstart, r.yend - r.ystart) == (4 * dc, 2 * dc)
        assert s.CX - r.xstart == dc
        assert s.CY == (s.LY - 1) // 2 == (r.ystart + r.yend) // 2
        assert s.U == 0.1


def test_c4_sweep_is_seeded():
    a, b, c = mbl.c4_specs(32, seed=5), mbl.c4_specs(32, seed=5), mbl.c4_specs(32, seed=6)
    key = lambda specs: [(s.LX, s.LY, s.R, s.Re) for s in specs]
    assert key(a) == key(b) and key(a) != key(c)


def _blocks(lib, geom):
    _, areas, _, _, _ = geom.init(lib)
    return [areas[i] for i in range(geom.g.number_area)]


def test_c1_geometry_configuration(lib32):
    """C1: 300 x 80-cell channel, one ratio-2 block of 80 x 40 coarse cells (161 x 81 fine
    nodes), D = 10 coarse, Re = 20 at Umax = 0.1: nu = 0.05, tau 0.65 coarse and 0.8 fine;
    48,510 node updates per coarse step (the MLUPS denominator of the case driver)."""
    g = mbl.c1_geometry(lib32)
    assert (g.g.LEVEL, g.g.LX, g.g.LY) == (2, 301, 81)
    blocks = _blocks(lib32, g)
    fine = [a for a in blocks if a.m == 2]
    assert len(fine) == 1 and (fine[0].nx, fine[0].ny) == (161, 81)
    assert all(abs(a.tau - (0.65 if a.m == 1 else 0.8)) < 1e-12 for a in blocks)
    assert sum(a.nx * a.ny * a.m for a in blocks) == 48510


def test_c2b_geometry_configuration(lib32):
    """Three-level case: 2048 x 512-cell channel, two nested ratio-2 blocks, D = 16 coarse =
    64 finest, Re = 100: nu = 0.016 coarse, doubling per level (tau 0.548, 0.596, 0.692);
    1,759,140 node updates per coarse step."""
    g = mbl.c2b_geometry(lib32)
    assert (g.g.LEVEL, g.g.LX, g.g.LY) == (3, 2049, 513)
    blocks = _blocks(lib32, g)
    assert sorted({a.m for a in blocks}) == [1, 2, 4]
    finest = [a for a in blocks if a.m == 4]
    assert len(finest) == 1 and (finest[0].nx, finest[0].ny) == (513, 257)
    for a in blocks:
        assert abs(a.tau - (3.0 * 0.016 * a.m + 0.5)) < 1e-12
    assert sum(a.nx * a.ny * a.m for a in blocks) == 1759140
