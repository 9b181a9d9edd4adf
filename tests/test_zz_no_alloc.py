"""sw2d_read_noalloc (the read that allocates no device memory: its
conversion passes through the library's fixed staging array) against
sw2d_read, through the C ABI: two calls in a row, the second on a smaller
grid, for every field it serves, bitwise; the refused fields and sizes
report SW2D_E_INVALID."""
import numpy as np
import pytest

from conftest import random_case

pytestmark = pytest.mark.gpu
import sw2d  # noqa: E402

FIELDS = (sw2d.F_H, sw2d.F_U, sw2d.F_V, sw2d.F_Z)


def say(m):
    print(m, flush=True)


@pytest.mark.parametrize("precision", [sw2d.FP32, sw2d.FP64])
def test_read_noalloc_equals_read(precision):
    # first a grid whose fields take several passes through the 16 MiB
    # staging array (2,097,152 doubles: 1600 x 1700 = 2.72 M cells), then a
    # smaller one: the array keeps the first call's contents, which must not
    # show in the second
    for (R, Cc), n_iter, seed in (((1600, 1700), 40, 3), ((70, 90), 25, 4)):
        say("case: %dx%d precision %d, %d iterations" % (R, Cc, precision, n_iter))
        z, h = random_case(seed, R, Cc, holes=0.03)
        s = sw2d.Sim(z, h, 2.0, dt=0.05, manning=0.03, cfl=0.45, edge_open=sw2d.EDGE_SOUTH,
                     precision=precision)
        s.step(n_iter)
        for f in FIELDS:
            a = s.read(f)
            b = s.read_noalloc(f)
            assert a.shape == b.shape
            assert np.array_equal(a, b), (R, f)
        # the same call again, after more iterations
        s.step(3)
        for f in FIELDS:
            assert np.array_equal(s.read(f), s.read_noalloc(f)), (R, f)
        for f in (sw2d.F_UU, sw2d.F_VV):
            with pytest.raises(sw2d.SW2DError) as e:
                s.read_noalloc(f)
            assert e.value.code == sw2d.E_INVALID
        s.close()


def test_read_noalloc_refuses_rows_wider_than_the_staging_array():
    say("case: one row of 2,097,152 cells (V has 2,097,153 values)")
    C = 2 * 1024 * 1024
    z = np.zeros((1, C))
    h = np.full((1, C), 0.1)
    s = sw2d.Sim(z, h, 1.0, dt=0.01, manning=0.03, precision=sw2d.FP32)
    assert np.array_equal(s.read(sw2d.F_H), s.read_noalloc(sw2d.F_H))   # C values: fits
    with pytest.raises(sw2d.SW2DError) as e:
        s.read_noalloc(sw2d.F_V)                                       # C + 1 values
    assert e.value.code == sw2d.E_INVALID
    s.close()
