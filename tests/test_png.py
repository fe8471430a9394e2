"""The PNG writer of minotert_amd.post (bench image output), CPU only: the file decodes with
Pillow to the same RGBA pixels, top row first."""
import numpy as np
import pytest

from minotert_amd.post import write_png


def test_png_round_trip(tmp_path):
    PIL = pytest.importorskip("PIL.Image")
    rng = np.random.default_rng(0)
    w, h = 37, 11
    px = rng.integers(0, 256, (h, w, 4), dtype=np.uint8)
    path = tmp_path / "x.png"
    write_png(str(path), px, w, h)
    with PIL.open(path) as im:
        assert im.size == (w, h) and im.mode == "RGBA"
        assert np.array_equal(np.asarray(im), px)
