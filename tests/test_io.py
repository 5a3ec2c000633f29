"""CPU tests of the file tasks `read` and `write` (src/ufo-read-task.c,
src/ufo-write-task.c) through the C task layer on host memory.

Files are made with numpy and Pillow (an independent TIFF codec) and read
back with Pillow, so a mistake in our TIFF reader and writer does not cancel
out.  The multipage expectations restate tests/test-multipage-readers.sh of
the reference (4 files x 15 images, image k holding the value k); the
counter-file names restate tests/test-file-write-regression.sh; foo.tif is
the reference's own test image (tests/foo.tif).
"""
import os

import numpy as np
import pytest
from PIL import Image

import ufo_tasks as U

HERE = os.path.dirname(os.path.abspath(__file__))

# tests/test-multipage-readers.sh: (image-start, image-step) -> first pixel of each output page
MULTIPAGE = [
    (5, 8, [5, 13, 21, 29, 37, 45, 53]),
    (17, 8, [17, 25, 33, 41, 49, 57]),
    (5, 30, [5, 35]),
    (17, 30, [17, 47]),
]


def _read(path, **props):
    t = U.Task("read")
    t.props.path = str(path)
    for k, v in props.items():
        setattr(t.props, k, v)
    t.setup()
    return t


def _write(filename, **props):
    t = U.Task("write")
    t.props.filename = str(filename)
    for k, v in props.items():
        setattr(t.props, k, v)
    t.setup()
    return t


def _pages(path):
    im = Image.open(path)
    out = []
    for k in range(getattr(im, "n_frames", 1)):
        im.seek(k)
        out.append(np.asarray(im).copy())
    return out


def _edf(path, frames, dtype_name, big_endian=False, size=None):
    """An ESRF data file: '{' key = value ; ... '}' padded to 1024 bytes, then the frames."""
    n, h, w = frames.shape
    keys = ["HeaderID = EH:000001:000000:000000", "Image = 1", "Dim_1 = %d" % w, "Dim_2 = %d" % h,
            "DataType = %s" % dtype_name, "ByteOrder = %s" % ("HighByteFirst" if big_endian else "LowByteFirst")]
    if size is not None:
        keys.append("Size = %d" % size)
    head = "{\n" + "".join(k + " ;\n" for k in keys)
    head = head + " " * (1024 - len(head) - 2) + "}\n"
    assert len(head) == 1024
    with open(path, "wb") as f:
        f.write(head.encode())
        f.write(frames.tobytes())


@pytest.fixture(scope="module")
def multipage(tmp_path_factory):
    """tests/make-input-multipage-readers: 4 files of 15 uint16 128x128 images, image k = k."""
    d = tmp_path_factory.mktemp("multipage")
    base = np.ones((15, 128, 128), dtype=np.uint16) * np.arange(15, dtype=np.uint16)[:, None, None]
    for i in range(4):
        cur = base + i * 15
        ims = [Image.fromarray(c) for c in cur]
        ims[0].save(d / ("multipage-image-%02d.tif" % i), save_all=True, append_images=ims[1:])
        cur.tofile(d / ("multipage-image-%02d.raw" % i))
        _edf(d / ("multipage-image-%02d.edf" % i), cur, "UnsignedShort")
    return d


@pytest.mark.parametrize("ext", ["tif", "raw", "edf"])
@pytest.mark.parametrize("start,step,expected", MULTIPAGE)
def test_multipage_readers(multipage, tmp_path, ext, start, step, expected):
    extra = dict(raw_width=128, raw_height=128, raw_bitdepth=16) if ext == "raw" else {}
    r = _read(multipage / ("multipage-image*." + ext), image_start=start, image_step=step, **extra)
    out = tmp_path / "multipage-out.tif"
    w = _write(out, tiff_bigtiff=False)
    for item in r.items():
        w.process_host(item)
    del w
    pages = _pages(out)
    assert [int(p[0, 0]) for p in pages] == expected
    assert all((p == p[0, 0]).all() for p in pages)


def test_reference_test_image():
    """tests/foo.tif of the reference: 16x16, 8-bit, every sample 1."""
    path = os.path.join(HERE, "golden", "foo.tif")
    items = list(_read(path).items())
    assert len(items) == 1
    np.testing.assert_array_equal(items[0], np.asarray(Image.open(path)).astype(np.float32))
    assert items[0].shape == (16, 16) and (items[0] == 1).all()


@pytest.mark.parametrize("dtype", ["uint8", "uint16", ">u2", "int16", "int32", "float32"])
def test_tiff_sample_types(tmp_path, dtype):
    rng = np.random.default_rng(3)
    info = np.iinfo(dtype) if np.dtype(dtype).kind in "iu" else None
    a = (rng.integers(info.min, info.max, (37, 53), endpoint=True).astype(dtype) if info is not None
         else rng.standard_normal((37, 53)).astype(dtype))
    Image.fromarray(a).save(tmp_path / "x.tif")
    got = list(_read(tmp_path / "x.tif").items())
    assert len(got) == 1
    np.testing.assert_array_equal(got[0], a.astype(np.float32))


def test_tiff_strips_and_roi(tmp_path):
    """Several strips per page (rows per strip 7) and the vertical ROI y / height / y-step."""
    a = np.arange(64 * 41, dtype=np.float32).reshape(64, 41)
    Image.fromarray(a).save(tmp_path / "s.tif", tiffinfo={278: 7})
    assert list(_read(tmp_path / "s.tif").items())[0].tolist() == a.tolist()
    for y, height, ystep in [(0, 0, 1), (5, 0, 1), (3, 20, 1), (10, 1, 20), (2, 50, 7), (63, 0, 3)]:
        r = _read(tmp_path / "s.tif", y=y, height=height, y_step=ystep)
        got = list(r.items())[0]
        stop = y + (height or 64 - y)
        np.testing.assert_array_equal(got, a[y:stop:ystep])


def test_roi_errors(tmp_path):
    Image.fromarray(np.zeros((10, 8), np.float32)).save(tmp_path / "z.tif")
    r = _read(tmp_path / "z.tif", y=10)
    with pytest.raises(U.TaskError) as e:
        r.get_requisition()
    assert e.value.code == 1 and "vertical ROI start 10 >= height 10" in str(e.value)
    r = _read(tmp_path / "z.tif", y=4, height=7)
    with pytest.raises(U.TaskError) as e:
        r.get_requisition()
    assert e.value.code == 1


def test_edf_types_byte_order_and_size(tmp_path):
    rng = np.random.default_rng(5)
    f = rng.standard_normal((3, 9, 11)).astype(np.float32)
    _edf(tmp_path / "a.edf", f.astype(">f4"), "FloatValue", big_endian=True)
    np.testing.assert_array_equal(np.stack(list(_read(tmp_path / "a.edf").items())), f)
    i = rng.integers(-2 ** 31, 2 ** 31 - 1, (2, 9, 11)).astype(np.int32)
    _edf(tmp_path / "b.edf", i, "SignedInteger")
    np.testing.assert_array_equal(np.stack(list(_read(tmp_path / "b.edf").items())), i.astype(np.float32))
    # Size limits the frames to the first two
    _edf(tmp_path / "c.edf", f, "Float", size=2 * 9 * 11 * 4)
    assert len(list(_read(tmp_path / "c.edf").items())) == 2


def test_edf_corrupt_header(tmp_path):
    with open(tmp_path / "bad.edf", "wb") as fh:
        fh.write(b"{ Dim_1 = 4 ; }\n" + bytes(100))
    r = _read(tmp_path / "bad.edf")
    with pytest.raises(U.TaskError) as e:
        r.get_requisition()
    assert "Corrupt EDF header" in str(e.value)


def test_raw_offsets_and_depths(tmp_path):
    rng = np.random.default_rng(7)
    frames = rng.integers(0, 255, (4, 6, 10)).astype(np.uint8)
    with open(tmp_path / "o.raw", "wb") as fh:
        for fr in frames:
            fh.write(b"\xaa" * 16 + fr.tobytes() + b"\xbb" * 8)
    r = _read(tmp_path / "o.raw", raw_width=10, raw_height=6, raw_bitdepth=8, raw_pre_offset=16, raw_post_offset=8)
    np.testing.assert_array_equal(np.stack(list(r.items())), frames.astype(np.float32))
    fl = rng.standard_normal((2, 6, 10)).astype(np.float32)
    fl.tofile(tmp_path / "f.raw")
    r = _read(tmp_path / "f.raw", raw_width=10, raw_height=6, raw_bitdepth=32)
    np.testing.assert_array_equal(np.stack(list(r.items())), fl)
    r = _read(tmp_path / "f.raw", raw_width=10, raw_height=6, raw_bitdepth=12)
    with pytest.raises(U.TaskError, match="bitdepth other than 8, 16 or 32"):
        r.get_requisition()


def test_file_selection(multipage):
    """start / step over sorted files, number, a single file ignores start."""
    def firsts(**kw):
        return [int(x[0, 0]) for x in _read(multipage / "multipage-image*.tif", **kw).items()]

    assert firsts(start=1, number=3) == [15, 16, 17]
    assert firsts(start=1, step=2, image_step=15) == [15, 45]
    assert firsts(step=3, image_step=5) == [0, 5, 10, 45, 50, 55]
    assert firsts(number=0) == []
    one = _read(multipage / "multipage-image-02.tif", start=3, image_start=14)
    assert [int(x[0, 0]) for x in one.items()] == [44]
    with pytest.raises(U.TaskError) as e:
        _read(multipage / "multipage-image*.tif", start=4)
    assert e.value.code == 0 and "start=4 skips too many files" in str(e.value)
    with pytest.raises(U.TaskError) as e:
        _read(multipage / "nothing*.tif")
    assert e.value.code == 0 and "does not match any files" in str(e.value)


def test_directory_path_and_type(multipage, tmp_path):
    """A directory means every readable file in it; `type` forces a reader for other names."""
    d = tmp_path / "dir"
    d.mkdir()
    for k in (2, 0, 1):
        Image.fromarray(np.full((3, 3), k, np.float32)).save(d / ("p%d.tif" % k))
    (d / "notes.txt").write_text("not an image")
    assert [int(x[0, 0]) for x in _read(d).items()] == [0, 1, 2]
    assert [int(x[0, 0]) for x in _read(str(d) + "/").items()] == [0, 1, 2]
    # image-start carried over files: edf-00 .. edf-03 hold 0..59, every 20th from 5
    assert [int(x[0, 0]) for x in _read(multipage / "*.edf", image_start=5, image_step=20).items()] == [5, 25, 45]
    a = np.arange(20, dtype=np.float32).reshape(4, 5)
    a.tofile(tmp_path / "frame.bin")
    got = list(_read(tmp_path / "frame.bin", type="raw", raw_width=5, raw_height=4, raw_bitdepth=32).items())
    np.testing.assert_array_equal(got[0], a)


def test_convert_false(tmp_path):
    Image.fromarray(np.ones((4, 4), np.uint16)).save(tmp_path / "u.tif")
    r = _read(tmp_path / "u.tif", convert=False)
    with pytest.raises(RuntimeError, match="convert=false"):
        list(r.items())
    Image.fromarray(np.ones((4, 4), np.float32)).save(tmp_path / "f.tif")
    assert len(list(_read(tmp_path / "f.tif", convert=False).items())) == 1


def test_write_counter_files(tmp_path):
    """tests/test-file-write-regression.sh: bytes-per-file=0 puts every image in its own %02i file."""
    w = _write(tmp_path / "foo-%02i.tif", bytes_per_file=0)
    for k in range(3):
        w.process_host(np.full((512, 512), k, np.float32))
    del w
    assert sorted(os.listdir(tmp_path)) == ["foo-00.tif", "foo-01.tif", "foo-02.tif"]
    for k in range(3):
        pages = _pages(tmp_path / ("foo-%02d.tif" % k))
        assert len(pages) == 1 and (pages[0] == k).all()


def test_write_bytes_per_file_and_counter(tmp_path):
    """Files split once the next image would pass bytes-per-file; counter-start / counter-step name them."""
    w = _write(tmp_path / "s-%i.raw", bytes_per_file=3 * 16 * 8 * 4, counter_start=10, counter_step=5)
    for k in range(7):
        w.process_host(np.full((8, 16), k, np.float32))
    del w
    names = sorted(os.listdir(tmp_path), key=lambda s: int(s[2:-4]))
    assert names == ["s-10.raw", "s-15.raw", "s-20.raw"]
    sizes = [os.path.getsize(tmp_path / n) // (16 * 8 * 4) for n in names]
    assert sizes == [3, 3, 1]
    # append: the counter starts after the files that exist
    w = _write(tmp_path / "s-%i.raw", append=True, bytes_per_file=0, counter_start=10)
    w.process_host(np.zeros((8, 16), np.float32))
    del w
    assert os.path.exists(tmp_path / "s-11.raw")


def test_write_explicit_bits32_accounting(tmp_path):
    """An explicit bits=32 counts 2 bytes per sample towards bytes-per-file (ufo-write-task.c:438-441)."""
    per = 16 * 8 * 4
    for explicit, expect in ((False, [2, 2]), (True, [4])):
        d = tmp_path / str(explicit)
        props = dict(bytes_per_file=2 * per)
        if explicit:
            props["bits"] = 32
        w = _write(d / "b-%i.raw", **props)
        for k in range(4):
            w.process_host(np.zeros((8, 16), np.float32))
        del w
        names = sorted(os.listdir(d))
        assert [os.path.getsize(d / n) // per for n in names] == expect


def test_write_setup_errors(tmp_path):
    with pytest.raises(U.TaskError, match="too many format specifiers"):
        _write(tmp_path / "a-%i-%i.tif")
    with pytest.raises(U.TaskError, match="valid file extension"):
        _write(tmp_path / "a.png")
    with pytest.raises(U.TaskError, match="integer format specifier"):
        _write(tmp_path / "a-%s.tif")
    # a missing directory is created
    w = _write(tmp_path / "new" / "dir" / "x.tif")
    w.process_host(np.zeros((2, 2), np.float32))
    del w
    assert os.path.exists(tmp_path / "new" / "dir" / "x.tif")


def test_write_bits_rejected_values_are_ignored():
    t = U.Task("write")
    t.props.bits = 16
    t.props.bits = 12  # the reference warns and keeps the previous value
    assert t.props.bits == 16


def test_write_8_16_bit_conversion(tmp_path):
    """Rescaling maps [min, max] onto [0, 255] / [0, 65535]; user bounds clip; without rescale
    values are truncated and wrap like the reference's x86 casts."""
    x = np.linspace(-1.0, 3.0, 40 * 30, dtype=np.float32).reshape(30, 40)
    for bits, top, dt in ((8, 255, np.uint8), (16, 65535, np.uint16)):
        w = _write(tmp_path / ("r%d.tif" % bits), bits=bits, tiff_bigtiff=False)
        w.process_host(x)
        w.process_host(np.full((30, 40), 7.0, np.float32))  # constant image: zero scale guard
        del w
        p = _pages(tmp_path / ("r%d.tif" % bits))
        assert p[0].dtype == dt and p[0].min() == 0 and p[0].max() == top
        assert (np.diff(p[0].astype(np.int64).ravel()) >= 0).all()
        assert (p[1] == 0).all()
        w = _write(tmp_path / ("c%d.tif" % bits), bits=bits, minimum=0.0, maximum=2.0, tiff_bigtiff=False)
        w.process_host(x)
        del w
        q = _pages(tmp_path / ("c%d.tif" % bits))[0].astype(np.int64)
        assert (q[x <= 0] == 0).all() and (q[x >= 2] == top).all()
        mid = (x > 0) & (x < 2)
        np.testing.assert_allclose(q[mid], np.floor(x[mid] / 2.0 * top), atol=1)
        w = _write(tmp_path / ("n%d.raw" % bits), bits=bits, rescale=False)
        w.process_host(np.array([[-1.0, 300.5, 70000.0, 5.9]], np.float32))
        del w
        got = np.fromfile(tmp_path / ("n%d.raw" % bits), dtype=dt)
        want = [top, 300 % (top + 1), 70000 % (top + 1), 5]
        assert got.tolist() == want


def test_bigtiff_round_trip(tmp_path):
    rng = np.random.default_rng(11)
    stack = rng.standard_normal((5, 33, 17)).astype(np.float32)
    w = _write(tmp_path / "big.tif")  # tiff-bigtiff defaults to true
    for s in stack:
        w.process_host(s)
    del w
    with open(tmp_path / "big.tif", "rb") as fh:
        assert fh.read(4) == b"II+\x00"
    np.testing.assert_array_equal(np.stack(_pages(tmp_path / "big.tif")), stack)
    np.testing.assert_array_equal(np.stack(list(_read(tmp_path / "big.tif").items())), stack)


def test_write_volume_planes(tmp_path):
    """A 3-D item is written plane by plane."""
    v = np.arange(3 * 4 * 5, dtype=np.float32).reshape(3, 4, 5)
    w = _write(tmp_path / "v.tif", tiff_bigtiff=False)
    w.process_host(v)
    del w
    np.testing.assert_array_equal(np.stack(_pages(tmp_path / "v.tif")), v)


def test_task_modes_and_properties():
    r, w = U.Task("read"), U.Task("write")
    assert r.mode & U.MODE_GENERATOR and r.mode & U.MODE_CPU
    assert w.mode & U.MODE_SINK
    assert r.L.ufomi_task_get_num_inputs(r.h) == 0 and w.L.ufomi_task_get_num_inputs(w.h) == 1
    assert r.props.path == "." and r.props.number == 4294967295 and r.props.convert is True
    assert r.props.type == "unspecified" and r.props.y_step == 1 and r.props.image_step == 1
    assert w.props.bytes_per_file == 137438953472 and w.props.bits == 32 and w.props.rescale is True
    assert w.props.tiff_bigtiff is True and w.props.filename == "" and w.props.counter_step == 1
    with pytest.raises(Exception):
        r.props.step = 0
    with pytest.raises(Exception):
        r.props.type = "hdf5"


def _bigtiff(path, strip_count, strip_value, width=4, height=4):
    """A little-endian BigTIFF of one 8-bit page with a hand-set StripOffsets entry."""
    import struct
    entries = [(256, 3, 1, width), (257, 3, 1, height), (258, 3, 1, 8), (278, 3, 1, 1),
               (273, 16, strip_count, strip_value)]
    ifd = struct.pack("<Q", len(entries)) + b"".join(struct.pack("<HHQQ", *e) for e in entries) + struct.pack("<Q", 0)
    with open(path, "wb") as f:
        f.write(b"II" + struct.pack("<HHHQ", 43, 8, 0, 16) + ifd + bytes(width * height))


@pytest.mark.parametrize("count,value", [
    (1 << 61, 0x4141414141414141),  # count x 8 bytes wraps to 0: must not be taken as inline
    (1 << 61, 0xFFFFFFFFFFFFFFF0),  # offset + index x 8 overflows
    (4, 0x7FFFFFFFFFFFFFF9),        # offset past any file
])
def test_bigtiff_crafted_strip_offsets_rejected(tmp_path, count, value):
    _bigtiff(tmp_path / "evil.tif", count, value)
    r = _read(tmp_path / "evil.tif")
    with pytest.raises((U.TaskError, RuntimeError)) as e:  # setup-time or read-time refusal
        r.get_requisition()
        list(r.items())
    assert "TIFF" in str(e.value)


def test_bigtiff_hand_made_reads(tmp_path):
    # the same writer with valid strips (4 rows of 4 bytes after the 16 + 8 + 5 x 20 + 8 byte header and IFD)
    _bigtiff(tmp_path / "ok.tif", 4, 0)
    import struct
    data = bytearray(open(tmp_path / "ok.tif", "rb").read())
    pixels = 16 + 8 + 5 * 20 + 8
    strips = len(data)
    data += struct.pack("<4Q", *(pixels + 4 * i for i in range(4)))
    data[16 + 8 + 4 * 20 + 12: 16 + 8 + 4 * 20 + 20] = struct.pack("<Q", strips)
    data[pixels: pixels + 16] = bytes(range(16))
    open(tmp_path / "ok.tif", "wb").write(bytes(data))
    r = _read(tmp_path / "ok.tif")
    r.get_requisition()
    (img,) = list(r.items())
    np.testing.assert_array_equal(img, np.arange(16, dtype=np.float32).reshape(4, 4))
