"""Array files (MetaImage / Dennerlein / raw): the reference's own test cases
(test/cpp/io/file_writer/meta_img_file_writer_tests.cpp,
test/cpp/io/file_reader/*_tests.cpp and array_file_reader_test_fix.*) restated
on this library's readers and writers.  The reference's fixture files are not
in its repository; its fixture is rows 2 x cols 3 (x slices 4), data 0, 1, 2,
... over the first file and continuing over the second, which is what these
tests write and read back."""
import os

import numpy as np
import pytest

import gct_amd as g


# ---- writer known answers (meta_img_file_writer_tests.cpp:89-140) ----------------

@pytest.mark.parametrize("dtype,name", [(np.uint8, "MET_UCHAR"), (np.int8, "MET_CHAR"), (np.uint16, "MET_USHORT"),
                                        (np.int16, "MET_SHORT"), (np.uint32, "MET_UINT"), (np.int32, "MET_INT"),
                                        (np.uint64, "MET_ULONG_LONG"), (np.int64, "MET_LONG_LONG"),
                                        (np.float32, "MET_FLOAT"), (np.float64, "MET_DOUBLE")])
def test_meta_type_names(dtype, name):
    assert g.meta_type_name(dtype) == name


def test_number_of_elements_is_24(tmp_path):
    f = tmp_path / "n.mha"
    g.write_meta(f, np.arange(24, dtype=np.float32), dim_size=(2, 3, 4))
    assert g.read_array(f).size == 24


def _fixture_3d(first=0):
    # rows 2, cols 3, slices 4 (array_file_reader_test_fix.cpp:14-20), iota data
    return np.arange(first, first + 24, dtype=np.float32)


@pytest.mark.parametrize("ext", [".mha", ".mhd"])
def test_meta_image_written_and_read_back(tmp_path, ext):
    data = _fixture_3d()
    f = tmp_path / f"temp{ext}"
    g.write_meta(f, data, dim_size=(3, 2, 4))
    md = g.read_metadata(f)
    assert (md.num_rows, md.num_cols, md.num_slices, md.type) == (2, 3, 4, g.elem_type(np.float32))
    assert np.array_equal(g.read_array(f), data)
    if ext == ".mhd":
        assert (tmp_path / "temp.raw").stat().st_size == 24 * 4
        assert "ElementDataFile = temp.raw" in f.read_text()


def test_meta_header_lines_match_the_reference_writer(tmp_path):
    f = tmp_path / "h.mha"
    g.write_meta(f, np.zeros(6, np.float32), dim_size=(3, 2), spacing=(0.5, 1.25), offset=(-1.0, 2.0))
    head = f.read_bytes().split(b"ElementDataFile = LOCAL\n")[0].decode()
    assert head.splitlines() == ["ObjectType = Image", "NDims = 2", "BinaryData = True", "DimSize = 3 2",
                                 "ElementSpacing = 0.500000 1.250000", "Offset = -1.000000 2.000000",
                                 "ElementType = MET_FLOAT", "BinaryDataByteOrderMSB = False"]
    assert f.stat().st_size == len(head) + len("ElementDataFile = LOCAL\n") + 24


def test_writer_errors(tmp_path):
    # invalid directory (meta_img_file_writer_tests.cpp:9-20)
    with pytest.raises(g.GctError) as e:
        g.write_meta(tmp_path / "invalid_dir" / "raw.mha", np.ones(1, np.float32), dim_size=(1, 1))
    assert e.value.status == 2
    # metadata / data mismatch (:22-31)
    with pytest.raises(g.GctError) as e:
        g.write_meta(tmp_path / "raw.mha", np.ones(1, np.float32), dim_size=(0, 0))
    assert e.value.status == 2


# ---- readers (mhd_file_reader_tests.cpp, array_file_reader_tests.cpp) -----------

@pytest.mark.parametrize("ext", [".mha", ".mhd"])
def test_series_of_meta_images(tmp_path, ext):
    g.write_meta(tmp_path / f"01{ext}", _fixture_3d(0), dim_size=(3, 2, 4))
    g.write_meta(tmp_path / f"02{ext}", _fixture_3d(24), dim_size=(3, 2, 4))
    # a file of another size and one of another extension are not part of the series
    g.write_meta(tmp_path / f"03{ext}", np.zeros(6, np.float32), dim_size=(3, 2))
    g.write_raw(tmp_path / "other.bin", _fixture_3d(100))
    assert np.array_equal(g.read_array(tmp_path / f"01{ext}"), _fixture_3d(0))
    assert np.array_equal(g.read_array(tmp_path / f"01{ext}", series=True), np.arange(48, dtype=np.float32))


def test_meta_reader_errors(tmp_path):
    with pytest.raises(g.GctError):
        g.read_array(tmp_path / "missing.mha")
    f = tmp_path / "bad.mha"
    f.write_text("NDims = 2\nDimSize = 3 2\nElementType = MET_HALF\nElementDataFile = LOCAL\n")
    with pytest.raises(g.GctError, match="not a valid ElementType"):
        g.read_metadata(f)
    f = tmp_path / "nodata.mhd"
    f.write_text("NDims = 2\nDimSize = 3 2\nElementType = MET_FLOAT\nElementDataFile = nodata.raw\n")
    with pytest.raises(g.GctError, match="not found"):
        g.read_array(f)
    (tmp_path / "nodata.raw").write_bytes(b"\0" * 20)          # 5 floats, 6 expected
    with pytest.raises(g.GctError, match="bytes requested"):
        g.read_array(f)


def test_meta_external_data_file_and_type_conversion(tmp_path):
    vals = np.array([-3, -1, 0, 2, 7, 300], np.int16)
    (tmp_path / "d.raw").write_bytes(vals.tobytes())
    (tmp_path / "d.mhd").write_text("ObjectType = Image\nNDims = 2\nDimSize = 3 2\nElementType = MET_SHORT\n"
                                    "ElementDataFile = d.raw\n")
    assert np.array_equal(g.read_array(tmp_path / "d.mhd", np.int16), vals)
    # static_cast conversions (array_file_reader.h:92-99)
    assert np.array_equal(g.read_array(tmp_path / "d.mhd", np.float32), vals.astype(np.float32))
    assert np.array_equal(g.read_array(tmp_path / "d.mhd", np.int32), vals.astype(np.int32))
    f = tmp_path / "t.mha"
    g.write_meta(f, np.array([0.9, 1.5, 2.99, 250.7], np.float64), dim_size=(2, 2))
    assert np.array_equal(g.read_array(f, np.uint8), np.array([0, 1, 2, 250], np.uint8))


@pytest.mark.parametrize("dtype", [np.uint8, np.uint16, np.float32, np.float64])
def test_den_files(tmp_path, dtype):
    # float_2d / uint8_2d fixtures: rows 2, cols 3, one slice; type from the element size
    data = np.arange(6).astype(dtype)
    f = tmp_path / "01.den"
    g.write_den(f, data, 2, 3, 1)
    raw = f.read_bytes()
    assert np.frombuffer(raw[:6], np.uint16).tolist() == [2, 3, 1]
    md = g.read_metadata(f)
    assert (md.num_rows, md.num_cols, md.num_slices, md.type) == (2, 3, 1, g.elem_type(dtype))
    assert np.array_equal(g.read_array(f, dtype), data)
    g.write_den(tmp_path / "02.den", np.arange(6, 12).astype(dtype), 2, 3, 1)
    assert np.array_equal(g.read_array(f, dtype, series=True), np.arange(12).astype(dtype))
    with pytest.raises(g.GctError):
        g.write_den(tmp_path / "x.den", data, 2, 2, 1)


def test_raw_files_with_caller_metadata(tmp_path):
    md = g.ArrayMetadata(2, 3, 4, g.elem_type(np.float32))
    g.write_raw(tmp_path / "01.bin", _fixture_3d(0))
    g.write_raw(tmp_path / "02.bin", _fixture_3d(24))
    got = g.read_metadata(tmp_path / "01.bin", raw_md=md)
    assert (got.num_rows, got.num_cols, got.num_slices) == (2, 3, 4)
    assert np.array_equal(g.read_array(tmp_path / "01.bin", raw_md=md), _fixture_3d(0))
    assert np.array_equal(g.read_array(tmp_path / "01.bin", series=True, raw_md=md), np.arange(48, dtype=np.float32))
    with pytest.raises(g.GctError):            # raw files need metadata
        g.read_array(tmp_path / "01.bin")
    bad = g.ArrayMetadata(2, 3, 5, g.elem_type(np.float32))
    with pytest.raises(g.GctError, match="bytes requested"):
        g.read_array(tmp_path / "01.bin", raw_md=bad)


# ---- projections and images (projection_reader.*, img_writer.*) -----------------

def test_projection_reader_shapes(tmp_path):
    p = np.arange(5 * 4 * 3, dtype=np.float32).reshape(5, 4, 3)       # 5 views of 4 rows x 3 cols
    f = tmp_path / "proj.mha"
    g.write_meta(f, p)
    assert g.read_projections(f, 2).shape == (5, 4, 3)
    assert np.array_equal(g.read_projections(f, 2), p)
    # 1-D detectors: columns per view, rows x slices views
    assert g.read_projections(f, 1).shape == (20, 3)


def test_image_writer_spacing_and_offset(tmp_path):
    img = g.Image(4, 3, 0.5, 0.25, 2, 2.0, mid=(1.0, -2.0, 3.0))
    vol = np.arange(24, dtype=np.float32).reshape(img.shape)
    f = tmp_path / "vol.mha"
    g.write_image(img, vol, f)
    head = f.read_bytes().split(b"ElementDataFile")[0].decode()
    assert "DimSize = 4 3 2" in head
    assert "ElementSpacing = 0.500000 0.250000 2.000000" in head
    # first voxel centre = mid - (N - 1) w / 2 (recon_img_geom.h:85-88)
    assert "Offset = 0.250000 -2.250000 2.000000" in head
    assert np.array_equal(g.read_array(f), vol.ravel())
    g.write_image(img, vol, tmp_path / "vol.den")
    md = g.read_metadata(tmp_path / "vol.den")
    assert (md.num_rows, md.num_cols, md.num_slices) == (3, 4, 2)
