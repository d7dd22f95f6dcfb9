"""Host side of the batched sweep (no GPU): the seeded 256-case sweep of uneven two-level
cylinder channels has the configuration it is quoted on. Re ~ U[40, 200]; D in {10, 16, 24,
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
