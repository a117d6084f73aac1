"""Box generation from tags (tmk_make_boxes, host code; no device needed).

Chombo's BRMeshRefine is not available, so layouts are pinned by what a regrid needs and by
cases whose answer is fixed: every tag inside the proper nesting domain is covered; boxes are
disjoint, inside the finer domain, multiples of the block factor and at most max_grid_size;
every box keeps the coarse-fine fill's reach inside the coarse level; one rectangle of tags
gives exactly that rectangle (blocked and split), two separated ones give separate boxes.
"""
import numpy as np
import pytest

from tetemoko_amd import capi


def _check_layout(tags, boxes, r, bf, mg, coarse=None, nb=None):
    ny, nx = tags.shape
    fnx, fny = nx * r, ny * r
    cov = np.zeros((fny, fnx), dtype=np.int32)
    for (x0, y0, x1, y1) in boxes:
        assert 0 <= x0 <= x1 < fnx and 0 <= y0 <= y1 < fny
        assert x0 % bf == 0 and y0 % bf == 0
        assert (x1 + 1) % bf == 0 or x1 == fnx - 1
        assert (y1 + 1) % bf == 0 or y1 == fny - 1
        assert x1 - x0 + 1 <= mg and y1 - y0 + 1 <= mg
        cov[y0:y1 + 1, x0:x1 + 1] += 1
    assert cov.max() <= 1, "boxes overlap"
    if nb is None:
        nb = -(-4 // r) + 1
    covered_c = cov.reshape(ny, r, nx, r).max(axis=(1, 3)) > 0
    if coarse is None:
        pnd = np.ones((ny, nx), dtype=bool)
    else:
        cl = np.zeros((ny, nx), dtype=bool)
        for (x0, y0, x1, y1) in coarse:
            cl[y0:y1 + 1, x0:x1 + 1] = True
        pnd = np.zeros((ny, nx), dtype=bool)
        for j in range(ny):
            for i in range(nx):
                pnd[j, i] = cl[max(j - nb, 0):j + nb + 1, max(i - nb, 0):i + nb + 1].all()
        # every box lies in the nesting domain
        assert not (covered_c & ~pnd).any()
    cf = bf // r
    blk_pnd = np.ones(((ny + cf - 1) // cf, (nx + cf - 1) // cf), dtype=bool)
    for j in range(ny):
        for i in range(nx):
            if not pnd[j, i]:
                blk_pnd[j // cf, i // cf] = False
    need = tags.astype(bool) & np.repeat(np.repeat(blk_pnd, cf, axis=0), cf, axis=1)[:ny, :nx]
    assert not (need & ~covered_c).any(), "a nested tag is not covered"
    return covered_c


def test_one_rectangle_gives_that_rectangle():
    tags = np.zeros((32, 64), dtype=np.uint8)
    tags[4:12, 16:40] = 1
    boxes = capi.make_boxes(tags, ref_ratio=4, block_factor=8, max_grid_size=64)
    _check_layout(tags, boxes, 4, 8, 64)
    # blocks of 2 coarse cells, refined by 4: x 64..159, y 16..47, split into pieces of <= 64
    xs = sorted({(b[0], b[2]) for b in boxes})
    ys = sorted({(b[1], b[3]) for b in boxes})
    assert xs[0][0] == 64 and xs[-1][1] == 159 and ys == [(16, 47)]
    assert sum((b[2] - b[0] + 1) * (b[3] - b[1] + 1) for b in boxes) == 96 * 32


def test_two_separated_clusters():
    tags = np.zeros((40, 80), dtype=np.uint8)
    tags[2:6, 4:10] = 1
    tags[20:30, 50:70] = 1
    boxes = capi.make_boxes(tags, ref_ratio=2, block_factor=4, max_grid_size=1000, fill_ratio=0.8)
    covered = _check_layout(tags, boxes, 2, 4, 1000)
    # nothing between the clusters is refined
    assert not covered[10:18, :].any()
    assert covered.sum() < 0.2 * covered.size


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_random_tags_cover_nest_and_block(seed):
    rng = np.random.default_rng(seed)
    ny, nx = 48, 96
    tags = (rng.uniform(size=(ny, nx)) < 0.03).astype(np.uint8)
    yy, xx = np.mgrid[0:ny, 0:nx]
    tags |= ((xx - 40) ** 2 / 200.0 + (yy - 5) ** 2 / 30.0 < 1).astype(np.uint8)
    coarse = [(16, 0, 79, 23), (32, 24, 63, 39)]
    for r, bf, mg in ((2, 4, 32), (4, 8, 64)):
        boxes = capi.make_boxes(tags, coarse_boxes=coarse, ref_ratio=r, block_factor=bf, max_grid_size=mg)
        _check_layout(tags, boxes, r, bf, mg, coarse=coarse)


def test_fill_ratio_controls_refined_area():
    yy, xx = np.mgrid[0:64, 0:64]
    tags = ((xx - 32) ** 2 + (yy - 32) ** 2 < 400).astype(np.uint8)
    areas = []
    for fr in (0.5, 0.7, 0.9):
        boxes = capi.make_boxes(tags, ref_ratio=2, block_factor=2, max_grid_size=1000, fill_ratio=fr)
        _check_layout(tags, boxes, 2, 2, 1000)
        areas.append(sum((b[2] - b[0] + 1) * (b[3] - b[1] + 1) for b in boxes))
    # a higher fill ratio refines less area around the disc (never less than the disc itself)
    assert areas[0] >= areas[1] >= areas[2] >= 4 * tags.sum()


def test_bad_parameters_rejected():
    tags = np.ones((8, 8), dtype=np.uint8)
    for kw in ({"block_factor": 6, "ref_ratio": 4}, {"max_grid_size": 12, "block_factor": 8},
               {"fill_ratio": 0.0}):
        with pytest.raises(capi.TmkError):
            capi.make_boxes(tags, **kw)
