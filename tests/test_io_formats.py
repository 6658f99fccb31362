"""On-disk formats of the reference (CImg load_tiff / save_tiff /
save_analyze / load_analyze / get_split, lds.txt) against independent
implementations:

- TIFF: files written by libtiff 4.7 (through Pillow, WRITE_LIBTIFF) are read
  here, and files written here are read back by libtiff (Pillow,
  READ_LIBTIFF).  CImg's own TIFF path needs libtiff headers, which this
  image lacks, so the tag set CImg::_save_tiff sets (CImg.h:51588-51644) is
  checked tag by tag.  Layouts libtiff does not produce through Pillow
  (big-endian, BigTIFF, tiles, planar, signed samples, 64-bit float) come
  from a byte-level writer in this file that follows the TIFF 6.0 and
  BigTIFF layouts.
- Analyze: CImg::save_analyze / load_analyze compiled from the reference's
  CImg.h into oracle/_ref, byte for byte and value for value.
- get_split: CImg::get_split compiled into oracle/_ref.
"""
import os
import struct

import numpy as np
import pytest

import _oracle as O
import spacetimetomography_amd as stt

PIL = pytest.importorskip("PIL")
from PIL import Image, TiffImagePlugin  # noqa: E402

needs_ref = pytest.mark.skipif(not O.HAVE_REF, reason="oracle/_ref not built (reference tree absent)")


def _rng(seed):
    return np.random.default_rng(seed)


# ------------------------------------------------------------ TIFF: ours --

@pytest.mark.parametrize("shape", [(1, 1, 5, 7), (1, 6, 7, 9), (3, 4, 5, 6), (1, 3, 5, 2100)])
def test_tiff_roundtrip(tmp_path, shape):
    a = _rng(1).standard_normal(shape).astype(np.float32)
    p = str(tmp_path / "v.tif")
    stt.save_tiff(p, a)
    info = stt.image_query(p)
    assert (info.spectrum, info.depth, info.height, info.width) == shape
    assert (info.bits, info.sample_format) == (32, 3)
    b = stt.load_image(p)
    assert b.shape == shape
    assert np.array_equal(b.view(np.uint32), a.view(np.uint32))


def _pil_pages(path, force_libtiff):
    """Pages decoded by libtiff (force_libtiff) and each page's tags as
    Pillow's IFD parser lists them (libtiff decoding drops some from that
    list, so the tags come from a second, plain pass)."""
    tags = []
    with Image.open(path) as im:
        for k in range(im.n_frames):
            im.seek(k)
            tags.append(dict(im.tag_v2))
    old = TiffImagePlugin.READ_LIBTIFF
    TiffImagePlugin.READ_LIBTIFF = force_libtiff
    try:
        pages = []
        with Image.open(path) as im:
            for k in range(im.n_frames):
                im.seek(k)
                im.load()
                pages.append(np.array(im))
        return pages, tags
    finally:
        TiffImagePlugin.READ_LIBTIFF = old


@pytest.mark.parametrize("w", [9, 512, 2100])
def test_tiff_written_here_is_read_by_libtiff_with_cimg_tags(tmp_path, w):
    a = _rng(2).standard_normal((3, 5, w)).astype(np.float32)
    p = str(tmp_path / "v.tif")
    stt.save_tiff(p, a)
    pages, tags = _pil_pages(p, force_libtiff=True)
    assert len(pages) == 3
    for z in range(3):
        assert pages[z].dtype == np.float32
        assert np.array_equal(pages[z], a[z])
    rps = max(8192 // (w * 4), 1)  # libtiff's default strip size for (uint32)-1
    t = tags[0]
    # the fields CImg::_save_tiff sets for a float plane (CImg.h:51595-51622)
    assert t[256] == w and t[257] == 5
    assert t[258] == (32,) and t[339] == (3,) and t[277] == 1
    assert t[259] == 1 and t[262] == 1 and t[284] == 1
    assert t[274] == 1 and t[266] == 1 and t[305] == "CImg"
    assert t[278] == rps
    nstrips = (5 + rps - 1) // rps
    assert len(t[273]) == nstrips and len(t[279]) == nstrips
    assert sum(t[279]) == 5 * w * 4


def test_tiff_rgb_photometric_for_three_channels(tmp_path):
    a = _rng(3).standard_normal((3, 2, 4, 5)).astype(np.float32)
    p = str(tmp_path / "c.tif")
    stt.save_tiff(p, a)
    raw = open(p, "rb").read()
    assert raw[:4] == b"II*\x00"
    assert np.array_equal(stt.load_image(p), a)
    # Pillow opens no float RGB TIFF, so list the first IFD's entries here.
    first = struct.unpack_from("<I", raw, 4)[0]
    n = struct.unpack_from("<H", raw, first)[0]
    tags = {}
    for k in range(n):
        tag, typ, cnt = struct.unpack_from("<HHI", raw, first + 2 + 12 * k)
        if typ == 3 and cnt == 1:
            tags[tag] = struct.unpack_from("<H", raw, first + 2 + 12 * k + 8)[0]
        else:
            tags[tag] = cnt
    # photometric 2 for 3 or 4 samples per pixel (CImg.h:51594), samples chunky
    assert tags[262] == 2 and tags[277] == 3 and tags[284] == 1
    assert tags[258] == 3 and tags[339] == 3  # one BitsPerSample / SampleFormat per sample
    assert list(tags) == sorted(tags)  # entries in ascending tag order


# ---------------------------------------------------- TIFF: libtiff wrote --

def _libtiff_write(path, arrays, **kw):
    old = TiffImagePlugin.WRITE_LIBTIFF
    TiffImagePlugin.WRITE_LIBTIFF = True
    try:
        ims = [Image.fromarray(a) for a in arrays]
        ims[0].save(path, save_all=True, append_images=ims[1:], **kw)
    finally:
        TiffImagePlugin.WRITE_LIBTIFF = old


@pytest.mark.parametrize("compression", ["raw", "tiff_lzw", "packbits"])
@pytest.mark.parametrize("dtype", [np.float32, np.uint8, np.uint16, np.int32])
def test_reads_libtiff_written_stacks(tmp_path, compression, dtype):
    r = _rng(4)
    if dtype == np.float32:
        pages = [r.standard_normal((37, 53)).astype(np.float32) for _ in range(4)]
    else:
        info = np.iinfo(dtype)
        lo, hi = max(info.min, -(1 << 30)), min(info.max, 1 << 30)
        pages = [r.integers(lo, hi, size=(37, 53), endpoint=True).astype(dtype) for _ in range(4)]
    p = str(tmp_path / "s.tif")
    _libtiff_write(p, pages, compression=compression)
    got = stt.load_image(p)
    assert got.shape == (1, 4, 37, 53)
    want = np.stack(pages).astype(np.float32)
    assert np.array_equal(got[0], want)


@pytest.mark.parametrize("dtype", [np.uint8, np.uint16, np.int32])
def test_reads_libtiff_lzw_with_horizontal_predictor(tmp_path, dtype):
    r = _rng(5)
    info = np.iinfo(dtype)
    lo, hi = max(info.min, -(1 << 30)), min(info.max, 1 << 30)
    a = r.integers(lo, hi, size=(64, 300), endpoint=True).astype(dtype)
    # smooth ramps compress, noise exercises wrap-around of the differences
    a[:32] = (np.arange(300)[None, :] * 3 + np.arange(32)[:, None]).astype(dtype)
    p = str(tmp_path / "pred.tif")
    _libtiff_write(p, [a], compression="tiff_lzw", tiffinfo={317: 2})
    _, tags = _pil_pages(p, force_libtiff=True)
    assert tags[0][317] == 2 and tags[0][259] == 5
    assert np.array_equal(stt.load_image(p)[0, 0], a.astype(np.float32))


def test_large_lzw_page_resets_its_code_table(tmp_path):
    # 12-bit codes fill up and the encoder emits Clear codes mid-strip
    a = _rng(6).integers(0, 255, size=(400, 400), endpoint=True).astype(np.uint8)
    p = str(tmp_path / "big_lzw.tif")
    _libtiff_write(p, [a], compression="tiff_lzw")
    assert np.array_equal(stt.load_image(p)[0, 0], a.astype(np.float32))


def test_rejects_unsupported_tiffs(tmp_path):
    a = np.arange(12, dtype=np.float32).reshape(3, 4)
    p = str(tmp_path / "deflate.tif")
    _libtiff_write(p, [a], compression="tiff_adobe_deflate")
    with pytest.raises(stt.SttError, match="compression 8"):
        stt.load_image(p)
    p = str(tmp_path / "pal.tif")
    Image.fromarray(np.arange(12, dtype=np.uint8).reshape(3, 4)).convert("P").save(p)
    with pytest.raises(stt.SttError, match="palette"):
        stt.load_image(p)
    p = str(tmp_path / "bilevel.tif")
    Image.new("1", (4, 3)).save(p)
    with pytest.raises(stt.SttError, match="1 bits"):
        stt.load_image(p)
    p = str(tmp_path / "sizes.tif")
    _libtiff_write(p, [a, np.zeros((4, 4), np.float32)])
    with pytest.raises(stt.SttError, match="different sizes"):
        stt.load_image(p)
    p = str(tmp_path / "not.tif")
    open(p, "wb").write(b"hello world, not a tiff")
    with pytest.raises(stt.SttError, match="not a TIFF"):
        stt.load_image(p)
    p = str(tmp_path / "trunc.tif")
    stt.save_tiff(p, np.ones((2, 64, 64), np.float32))
    raw = open(p, "rb").read()
    open(p, "wb").write(raw[:20000])
    with pytest.raises(stt.SttError):
        stt.load_image(p)
    with pytest.raises(stt.SttError, match="extension"):
        stt.load_image(str(tmp_path / "x.png"))
    with pytest.raises(stt.SttError, match="cannot open"):
        stt.load_image(str(tmp_path / "missing.tif"))


# ------------------------------------------- TIFF: byte-level fixtures --

_TYPES = {3: ("H", 2), 4: ("I", 4), 16: ("Q", 8)}


def _tiff_bytes(pages, bps, sf, big_endian=False, bigtiff=False, tile=None, planar=1, rps=None,
                loop=False):
    """A TIFF file: every IFD first, then the pixel data (the opposite of
    libtiff's order).  pages: list of (spp, H, W) arrays already in the
    on-disk sample type."""
    e = ">" if big_endian else "<"
    hdr_len = 16 if bigtiff else 8
    ent_len, cnt_fmt, nxt_fmt, off_t = (20, "Q", "Q", 16) if bigtiff else (12, "H", "I", 4)
    spp, H, W = pages[0].shape
    chunks_per_page = []
    for pg in pages:
        ch = []
        planes = [pg[s:s + 1] for s in range(spp)] if planar == 2 else [pg]
        for pl in planes:
            inter = np.ascontiguousarray(np.moveaxis(pl, 0, -1))  # (H, W, samples)
            if tile:
                tw, th = tile
                for ty in range(0, H, th):
                    for tx in range(0, W, tw):
                        t = np.zeros((th, tw, inter.shape[2]), inter.dtype)
                        blk = inter[ty:ty + th, tx:tx + tw]
                        t[:blk.shape[0], :blk.shape[1]] = blk
                        ch.append(t.astype(t.dtype.newbyteorder(e)).tobytes())
            else:
                r = rps or H
                for y in range(0, H, r):
                    ch.append(inter[y:y + r].astype(inter.dtype.newbyteorder(e)).tobytes())
        chunks_per_page.append(ch)
    n_ent = 12 if tile else 11
    ifd_len = (8 if bigtiff else 2) + n_ent * ent_len + (8 if bigtiff else 4)
    # arrays of offsets/counts after all IFDs, then data
    nch = len(chunks_per_page[0])
    arr_len = 2 * nch * (8 if bigtiff else 4)
    base = hdr_len + len(pages) * (ifd_len + arr_len)
    data_off = base
    out_ifds = []
    for k, ch in enumerate(chunks_per_page):
        offs, cnts = [], []
        for c in ch:
            offs.append(data_off)
            cnts.append(len(c))
            data_off += len(c)
        ifd_at = hdr_len + k * (ifd_len + arr_len)
        arr_at = ifd_at + ifd_len
        ents = [(256, 4, 1, W), (257, 4, 1, H), (258, 3, 1, bps), (259, 3, 1, 1), (262, 3, 1, 1),
                (277, 3, 1, spp), (284, 3, 1, planar), (339, 3, 1, sf)]
        if tile:
            ents += [(322, 4, 1, tile[0]), (323, 4, 1, tile[1])]
            ents += [(324, off_t, nch, None), (325, off_t, nch, None)]
        else:
            ents += [(273, off_t, nch, None), (278, 4, 1, rps or H), (279, off_t, nch, None)]
        ents = sorted(ents)
        assert len(ents) == n_ent
        b = struct.pack(e + cnt_fmt, n_ent)
        arrays = b""
        for tag, typ, cnt, val in ents:
            b += struct.pack(e + "HH", tag, typ) + struct.pack(e + ("Q" if bigtiff else "I"), cnt)
            cap = 8 if bigtiff else 4
            if val is None:
                vals = offs if tag in (273, 324) else cnts
                vfmt, vsz = _TYPES[typ]
                payload = struct.pack(e + vfmt * cnt, *vals)
                if len(payload) <= cap:
                    b += payload.ljust(cap, b"\0")
                else:
                    b += struct.pack(e + ("Q" if bigtiff else "I"), arr_at + len(arrays))
                    arrays += payload
            else:
                vfmt, vsz = _TYPES[typ]
                b += struct.pack(e + vfmt, val).ljust(cap, b"\0")
        nxt = 0 if k + 1 == len(pages) else hdr_len + (k + 1) * (ifd_len + arr_len)
        if loop and k + 1 == len(pages):
            nxt = hdr_len
        b += struct.pack(e + nxt_fmt, nxt)
        b += arrays.ljust(arr_len, b"\0")
        out_ifds.append(b)
    if bigtiff:
        head = (b"MM" if big_endian else b"II") + struct.pack(e + "HHHQ", 43, 8, 0, hdr_len)
    else:
        head = (b"MM" if big_endian else b"II") + struct.pack(e + "HI", 42, hdr_len)
    return head + b"".join(out_ifds) + b"".join(b"".join(ch) for ch in chunks_per_page)


def _write(tmp_path, name, data):
    p = str(tmp_path / name)
    open(p, "wb").write(data)
    return p


@pytest.mark.parametrize("big_endian", [False, True])
@pytest.mark.parametrize("bigtiff", [False, True])
@pytest.mark.parametrize("layout", ["strips", "tiles", "planar_strips", "planar_tiles"])
def test_reads_byte_level_layouts(tmp_path, big_endian, bigtiff, layout):
    r = _rng(7)
    spp = 1 if layout in ("strips", "tiles") else 3
    pages = [r.standard_normal((spp, 21, 35)).astype(np.float32) for _ in range(3)]
    tile = (16, 16) if "tiles" in layout else None
    planar = 2 if "planar" in layout else 1
    data = _tiff_bytes(pages, 32, 3, big_endian=big_endian, bigtiff=bigtiff, tile=tile, planar=planar, rps=4)
    p = _write(tmp_path, "f.tif", data)
    got = stt.load_image(p)
    want = np.stack(pages, axis=1)  # (c, z, y, x)
    assert got.shape == want.shape
    assert np.array_equal(got, want)


@pytest.mark.parametrize("kind", ["i8", "u8", "i16", "u16", "f16_as_i16", "i32", "u32", "f64"])
def test_sample_types_convert_as_cimg(tmp_path, kind):
    # CImg::_load_tiff (CImg.h:47732-47760): 8 and 16 bit are unsigned only
    # for SampleFormat 1, 32 bit is uint / int / float for formats 1 / 2 / else.
    r = _rng(8)
    spec = {"i8": (np.int8, 8, 2), "u8": (np.uint8, 8, 1), "i16": (np.int16, 16, 2), "u16": (np.uint16, 16, 1),
            "f16_as_i16": (np.int16, 16, 3), "i32": (np.int32, 32, 2), "u32": (np.uint32, 32, 1),
            "f64": (np.float64, 64, 3)}[kind]
    dt, bps, sf = spec
    if np.issubdtype(dt, np.integer):
        info = np.iinfo(dt)
        a = r.integers(info.min, info.max, size=(1, 9, 11), endpoint=True, dtype=np.int64).astype(dt)
    else:
        a = r.standard_normal((1, 9, 11)) * 1e3
    for be in (False, True):
        p = _write(tmp_path, f"{kind}{int(be)}.tif", _tiff_bytes([a], bps, sf, big_endian=be))
        got = stt.load_image(p)[0, 0]
        assert np.array_equal(got, a[0].astype(np.float32)), (kind, be)


def test_rejects_looping_ifd_chain(tmp_path):
    pages = [np.zeros((1, 4, 4), np.float32)] * 2
    p = _write(tmp_path, "loop.tif", _tiff_bytes(pages, 32, 3, loop=True))
    with pytest.raises(stt.SttError, match="loops"):
        stt.load_image(p)


# -------------------------------------------------------------- Analyze --

@needs_ref
@pytest.mark.parametrize("name", ["flow.hdr", "flow.img", "flow", "flow.nii"])
def test_analyze_bytes_match_cimg(tmp_path, name):
    a = _rng(9).standard_normal((3, 6, 7, 8)).astype(np.float32)  # (c, z, y, x): a flow field
    ours, theirs = tmp_path / "ours", tmp_path / "theirs"
    ours.mkdir()
    theirs.mkdir()
    vox = np.array([0.5, 1.25, 2.0], np.float32)
    stt.save_analyze(str(ours / name), a, voxel=vox)
    assert O.ref.r_save_analyze(str(theirs / name).encode(), O.f32(a), 8, 7, 6, 3, O.f32(vox), 16) == 0
    assert sorted(os.listdir(ours)) == sorted(os.listdir(theirs))
    for f in os.listdir(ours):
        assert open(ours / f, "rb").read() == open(theirs / f, "rb").read(), f


@needs_ref
@pytest.mark.parametrize("datatype", [2, 4, 8, 16, 64])
def test_analyze_reader_matches_cimg_load(tmp_path, datatype):
    r = _rng(10)
    a = (r.standard_normal((2, 5, 6, 7)) * 60 + 100).astype(np.float32)
    p = str(tmp_path / "img.hdr")
    assert O.ref.r_save_analyze(p.encode(), O.f32(a), 7, 6, 5, 2, None, datatype) == 0
    dims = np.zeros(4, np.int32)
    vox = np.zeros(3, np.float32)
    want = np.zeros(a.size, np.float32)
    assert O.ref.r_load_analyze(p.encode(), O.i32(dims), O.f32(vox), O.f32(want), want.size) == 0
    info = stt.image_query(p)
    assert [info.width, info.height, info.depth, info.spectrum] == list(dims)
    assert list(info.voxel) == list(vox)
    got = stt.load_image(p)
    assert np.array_equal(got.ravel(), want)


@needs_ref
def test_cimg_reads_our_analyze(tmp_path):
    a = _rng(11).standard_normal((3, 4, 5, 6)).astype(np.float32)
    for name in ("u.hdr", "u.nii"):
        p = str(tmp_path / name)
        stt.save_analyze(p, a)
        dims = np.zeros(4, np.int32)
        out = np.zeros(a.size, np.float32)
        assert O.ref.r_load_analyze(p.encode(), O.i32(dims), None, O.f32(out), out.size) == 0
        assert list(dims) == [6, 5, 4, 3]
        assert np.array_equal(out, a.ravel())


def test_analyze_byte_swapped_header_and_scale(tmp_path):
    # A big-endian Analyze file with a scale factor: CImg's reader detects the
    # order from the header size (>= 4096 after reading it little-endian),
    # multiplies by the scale factor (offset 112) and reads pixdim at 80.
    a = (_rng(12).integers(-3000, 3000, size=(2, 3, 4))).astype(np.int16)
    hd = bytearray(348)
    struct.pack_into(">i", hd, 0, 348)
    struct.pack_into(">8h", hd, 40, 3, 4, 3, 2, 1, 0, 0, 0)
    struct.pack_into(">hh", hd, 70, 4, 16)
    struct.pack_into(">4f", hd, 76, 0.0, 0.5, 0.75, 2.5)
    struct.pack_into(">f", hd, 112, 0.25)
    open(tmp_path / "be.hdr", "wb").write(bytes(hd))
    open(tmp_path / "be.img", "wb").write(a.astype(">i2").tobytes())
    info = stt.image_query(str(tmp_path / "be.hdr"))
    assert (info.width, info.height, info.depth, info.spectrum) == (4, 3, 2, 1)
    assert list(info.voxel) == [0.5, 0.75, 2.5]
    got = stt.load_image(str(tmp_path / "be.img"))  # the .img name pairs with the .hdr
    assert np.array_equal(got[0], a.astype(np.float32) * np.float32(0.25))


def test_oversized_headers_are_rejected(tmp_path):
    hd = bytearray(348)
    struct.pack_into("<i", hd, 0, 348)
    struct.pack_into("<5h", hd, 40, 4, -1, -1, -1, -1)  # 65535^4 samples
    struct.pack_into("<h", hd, 70, 64)
    open(tmp_path / "huge.hdr", "wb").write(bytes(hd))
    open(tmp_path / "huge.img", "wb").write(bytes(8))
    with pytest.raises(stt.SttError, match="2\\^40"):
        stt.load_image(str(tmp_path / "huge.hdr"))
    pages = [np.zeros((1, 4, 4), np.float32)]
    data = bytearray(_tiff_bytes(pages, 32, 3))
    # width and length entries (tags 256, 257) to 2^24 - 1 and 2^24 - 1 pages worth: one page of 2^48 samples
    n = struct.unpack_from("<H", data, 8)[0]
    for k in range(n):
        tag = struct.unpack_from("<H", data, 10 + 12 * k)[0]
        if tag in (256, 257):
            struct.pack_into("<I", data, 10 + 12 * k + 8, (1 << 24) - 1)
    p = _write(tmp_path, "huge.tif", bytes(data))
    with pytest.raises(stt.SttError, match="2\\^40"):
        stt.load_image(p)


def test_analyze_rejects_bad_input(tmp_path):
    with pytest.raises(stt.SttError, match="16-bit"):
        stt.save_analyze(str(tmp_path / "big.hdr"), np.zeros((1, 1, 40000), np.float32))
    with pytest.raises(stt.SttError, match="extension"):
        stt.save_analyze(str(tmp_path / "x.dat"), np.zeros((2, 2), np.float32))
    hd = bytearray(348)
    struct.pack_into("<i", hd, 0, 348)
    struct.pack_into("<5h", hd, 40, 3, 2, 2, 2, 1)
    struct.pack_into("<h", hd, 70, 32)  # complex: not read by CImg either
    open(tmp_path / "c.hdr", "wb").write(bytes(hd))
    open(tmp_path / "c.img", "wb").write(bytes(64))
    with pytest.raises(stt.SttError, match="datatype 32"):
        stt.load_image(str(tmp_path / "c.hdr"))
    struct.pack_into("<h", hd, 70, 16)
    open(tmp_path / "c.hdr", "wb").write(bytes(hd))
    open(tmp_path / "c.img", "wb").write(bytes(8))
    with pytest.raises(stt.SttError, match="shorter"):
        stt.load_image(str(tmp_path / "c.hdr"))


# ---------------------------------------------------------------- split --

@needs_ref
def test_split_sizes_match_cimg():
    for size in list(range(1, 40)) + [450, 451, 960, 1000]:
        for nb in range(1, min(size, 20) + 1):
            want = np.zeros(nb, np.int32)
            assert O.ref.r_split_sizes(size, nb, O.i32(want)) == nb
            assert list(stt.split_sizes(size, nb)) == list(want), (size, nb)


def test_split_sizes_properties():
    for size, nb in [(450, 8), (10, 3), (7, 7), (1000, 16), (96, 8)]:
        s = stt.split_sizes(size, nb)
        assert s.sum() == size and len(s) == nb
        assert set(s.tolist()) <= {size // nb, -(-size // nb)}
    assert list(stt.split_sizes(96, 8)) == [12] * 8
    with pytest.raises(stt.SttError):
        stt.split_sizes(3, 4)


# ---------------------------------------------------------------- lds.txt --

def test_start_angles_frame_major(tmp_path):
    p = tmp_path / "lds.txt"
    p.write_text("0 36.5\t-12.25\n+7e1  1.5E-1\r\n.5 9. 100 \n 42 43")
    got = stt.read_start_angles(str(p), 3, 3)
    want = np.array([0, 36.5, -12.25, 70, 0.15, 0.5, 9, 100, 42], np.float32).reshape(3, 3)
    assert np.array_equal(got, want)
    assert got.dtype == np.float32


def test_start_angles_errors(tmp_path):
    p = tmp_path / "lds.txt"
    p.write_text("1 2 3")
    with pytest.raises(stt.SttError, match="3 numbers, 4 needed"):
        stt.read_start_angles(str(p), 2, 2)
    p.write_text("1 2x 3 4")
    with pytest.raises(stt.SttError, match="malformed"):
        stt.read_start_angles(str(p), 2, 2)
    p.write_text("1 inf 3 4")
    with pytest.raises(stt.SttError, match="malformed"):
        stt.read_start_angles(str(p), 2, 2)
    with pytest.raises(stt.SttError, match="cannot open"):
        stt.read_start_angles(str(tmp_path / "none.txt"), 1, 1)
