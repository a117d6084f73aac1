"""Boundary plot files (libtmk_h5.so; AMRLevelLinElast.cpp:1628-1756), read back with the HDF5
project's own h5dump tool, independently of the writer: the root and level headers, the fault
boxes in layout order, the data over each box grown by the output ghost, the offsets, the file
name, and the rejection of bad arguments."""
import os

import numpy as np
import pytest

from h5dump_util import H5DUMP, h5_attr, h5_data, h5_ints, h5_objects
from tetemoko_amd import bdry_plot as bp

pytestmark = pytest.mark.skipif(not os.path.exists(H5DUMP), reason="h5dump not available")


def _levels(rng, nb=11):
    # level 0: two fault boxes given out of layout order; level 1: no fault boxes;
    # level 2: three boxes, one apart from the others
    l0 = [(32, 0, 63, 0), (0, 0, 31, 0)]
    l2 = [(300, 0, 331, 0), (256, 0, 299, 0), (512, 0, 519, 0)]
    mk = lambda bl: [rng.standard_normal((nb, b[2] - b[0] + 9)) for b in bl]
    return [
        {"level": 0, "ref_ratio": 2, "dx": 25.0, "dt": 2e-3, "time": 0.75, "domain_nx": 64, "boxes": l0, "data": mk(l0)},
        {"level": 1, "ref_ratio": 2, "dx": 12.5, "dt": 1e-3, "time": 0.75, "domain_nx": 128, "boxes": [], "data": []},
        {"level": 2, "ref_ratio": 2, "dx": 6.25, "dt": 5e-4, "time": 0.75, "domain_nx": 1024, "boxes": l2, "data": mk(l2)},
    ]


def test_plot_name():
    assert bp.plot_name("run/bssa.", 120) == "run/bssa.boundary.000120.2d.hdf5"
    assert bp.plot_name("", 0) == "boundary.000000.2d.hdf5"


def test_headers_and_layout(tmp_path):
    rng = np.random.default_rng(1)
    levels = _levels(rng)
    names = bp.BDRY_NAMES["rs"]
    path = bp.write(str(tmp_path / bp.plot_name("x.", 40)), levels, 40, 0.75, names, max_level=3)
    objs = h5_objects(path)
    assert {"/Chombo_global", "/level_0", "/level_2", "/level_0/boxes", "/level_0/Processors",
            "/level_0/data:datatype=0", "/level_0/data:offsets=0", "/level_0/data_attributes"} <= objs
    assert "/level_1" not in objs  # no fault boxes, no group (AMRLevelLinElast.cpp:1721)
    assert h5_attr(path, "/max_level") == 3 and h5_attr(path, "/num_levels") == 3
    assert h5_attr(path, "/iteration") == 40 and h5_attr(path, "/time") == 0.75
    assert h5_attr(path, "/num_components") == 11
    assert [h5_attr(path, "/component_%d" % c) for c in range(11)] == names
    assert h5_attr(path, "/Chombo_global/SpaceDim") == 2
    for lv in (levels[0], levels[2]):
        g = "/level_%d" % lv["level"]
        assert h5_attr(path, g + "/ref_ratio") == lv["ref_ratio"]
        assert h5_attr(path, g + "/dx") == lv["dx"] and h5_attr(path, g + "/dt") == lv["dt"]
        assert h5_attr(path, g + "/time") == lv["time"]
        assert h5_attr(path, g + "/prob_domain") == [0, 0, lv["domain_nx"] - 1, 0]
        assert h5_attr(path, g + "/data_attributes/comps") == 11
        assert h5_attr(path, g + "/data_attributes/objectType") == "FArrayBox"
        assert h5_attr(path, g + "/data_attributes/ghost") == [4, 0]
        assert h5_attr(path, g + "/data_attributes/outputGhost") == [1, 0]


def test_data_boxes_offsets(tmp_path):
    rng = np.random.default_rng(2)
    levels = _levels(rng)
    path = bp.write(str(tmp_path / "f.hdf5"), levels, 7, 0.75, bp.BDRY_NAMES["rs"])
    for lv in (levels[0], levels[2]):
        g = "/level_%d" % lv["level"]
        order = sorted(range(len(lv["boxes"])), key=lambda k: lv["boxes"][k][0])
        boxes = np.array(h5_ints(path, g + "/boxes")).reshape(-1, 4)
        assert [tuple(b) for b in boxes] == [lv["boxes"][k] for k in order]
        assert (h5_data(path, g + "/Processors", "<i4", tmp_path) == 0).all()
        offs = h5_data(path, g + "/data:offsets=0", "<i8", tmp_path)
        data = h5_data(path, g + "/data:datatype=0", "<f8", tmp_path)
        want, pos = [], [0]
        for k in order:
            d = lv["data"][k]
            want.append(d[:, 3:-3].ravel())  # valid faces plus one x-ghost each side, per component
            pos.append(pos[-1] + want[-1].size)
        assert offs.tolist() == pos
        assert np.array_equal(data, np.concatenate(want))


def test_slip_weakening_names(tmp_path):
    rng = np.random.default_rng(3)
    b = [(0, 0, 15, 0)]
    lv = {"level": 0, "ref_ratio": 1, "dx": 50.0, "dt": 1e-3, "time": 0.0, "domain_nx": 16, "boxes": b,
          "data": [rng.standard_normal((9, 24))]}
    path = bp.write(str(tmp_path / "sw.hdf5"), [lv], 0, 0.0, bp.BDRY_NAMES["sw"])
    assert h5_attr(path, "/num_components") == 9
    assert h5_attr(path, "/component_8") == "rupture"
    assert np.array_equal(h5_data(path, "/level_0/data:datatype=0", "<f8", tmp_path), lv["data"][0][:, 3:-3].ravel())


def test_bad_arguments(tmp_path):
    rng = np.random.default_rng(4)
    good = _levels(rng)[0]
    names = bp.BDRY_NAMES["rs"]
    bad = [dict(good, boxes=[(32, 1, 63, 1), (0, 0, 31, 0)]),   # not on the fault row
           dict(good, boxes=[(32, 0, 64, 0), (0, 0, 31, 0)]),   # beyond the domain
           dict(good, level=-1)]
    for lv in bad:
        with pytest.raises(bp.H5Error):
            bp.write(str(tmp_path / "bad.hdf5"), [lv], 0, 0.0, names)
    with pytest.raises(bp.H5Error):  # data of the wrong shape
        bp.write(str(tmp_path / "bad.hdf5"), [dict(good, data=[good["data"][0][:, 1:], good["data"][1]])], 0, 0.0, names)
    with pytest.raises(bp.H5Error):  # unwritable path: the library's error comes back as an error code
        bp.write(str(tmp_path / "no" / "such" / "dir.hdf5"), [good], 0, 0.0, names)
