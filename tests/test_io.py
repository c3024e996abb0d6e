"""Request ingest and result output formats (host code of the library) -- CPU only.

Each format is checked against an independent implementation: Pillow's
libtiff-based TIFF codec reads what the library writes and writes what the
library reads (uncompressed and LZW); the 3-sample vector TIFF is decoded by a
TIFF 6.0 LZW decoder written here from the specification; the VTI file is
parsed here.  The odd-ification and degree rules follow include/tiff.h:54-59
and include/config.h:78-81.
"""
import io
import math
import os
import struct

import numpy as np
import pytest


def _pil_stack(path, stack, compression=None):
    from PIL import Image

    imgs = [Image.fromarray(np.ascontiguousarray(p, dtype=np.float32)) for p in stack]
    kw = {"compression": compression} if compression else {}
    imgs[0].save(path, format="TIFF", save_all=True, append_images=imgs[1:], **kw)


def _pil_read(path):
    from PIL import Image

    im = Image.open(path)
    out = []
    for i in range(getattr(im, "n_frames", 1)):
        im.seek(i)
        out.append(np.array(im, dtype=np.float32))
    return np.stack(out)


@pytest.mark.parametrize("compression", [None, "tiff_lzw"])
@pytest.mark.parametrize("hw", [(7, 7), (8, 8), (7, 8), (10, 9), (31, 33)])
def test_tiff_read_odd_ification_quirk(tmp_path, compression, hw):
    """tiff::read keeps rows h - (w even) and columns w - (h even): the checks are crossed."""
    from lamino_amd import ingest

    h, w = hw
    rng = np.random.default_rng(1)
    stack = rng.standard_normal((3, h, w)).astype(np.float32)
    p = str(tmp_path / "s.tif")
    _pil_stack(p, stack, compression)
    got = ingest.read_tiff(p)
    rows = h - (1 if w % 2 == 0 else 0)
    cols = w - (1 if h % 2 == 0 else 0)
    assert got.shape == (3, rows, cols)
    np.testing.assert_array_equal(got, stack[:, :rows, :cols])


def test_tiff_write_is_read_by_pillow_and_back(tmp_path):
    from lamino_amd import ingest

    rng = np.random.default_rng(2)
    vol = rng.standard_normal((4, 9, 11)).astype(np.float32)
    p = str(tmp_path / "v.tiff")
    ingest.write_tiff(p, vol)
    np.testing.assert_array_equal(_pil_read(p), vol)
    np.testing.assert_array_equal(ingest.read_tiff(p), vol)   # odd rows and columns: nothing dropped


def test_tiff_write3_names_and_content(tmp_path):
    from lamino_amd import ingest

    rng = np.random.default_rng(3)
    m = rng.standard_normal((3, 2, 5, 7)).astype(np.float32)
    paths = ingest.write_tiff3(str(tmp_path / "rec.tiff"), m)
    assert [os.path.basename(q) for q in paths] == ["rec0.tiff", "rec1.tiff", "rec2.tiff"]
    for i, q in enumerate(paths):
        np.testing.assert_array_equal(_pil_read(q), m[i])
    # a name without a .tif extension keeps its whole stem
    paths = ingest.write_tiff3(str(tmp_path / "out"), m)
    assert os.path.basename(paths[0]) == "out0.tiff" and os.path.exists(paths[2])


def _lzw_decode(data):
    """TIFF 6.0 section 13 LZW decoder (codes MSB first, 9-12 bits, early change)."""
    out = bytearray()
    table = [bytes([i]) for i in range(256)] + [b"", b""]
    width, pos, prev = 9, 0, None
    nbits = 8 * len(data)
    while pos + width <= nbits:
        code = 0
        for _ in range(width):
            code = (code << 1) | ((data[pos >> 3] >> (7 - (pos & 7))) & 1)
            pos += 1
        if code == 257:
            break
        if code == 256:
            table = table[:258]
            width, prev = 9, None
            continue
        if code < len(table):
            entry = table[code]
            if prev is not None:
                table.append(table[prev] + entry[:1])
        else:
            entry = table[prev] + table[prev][:1]
            table.append(entry)
        out += entry
        prev = code
        if len(table) + 1 >= (1 << width) and width < 12:
            width += 1
    return bytes(out)


def _tiff_pages(buf):
    """(tags dict, strip bytes) per page of a little-endian TIFF."""
    assert buf[:4] == b"II*\x00"
    off = struct.unpack_from("<I", buf, 4)[0]
    pages = []
    while off:
        n = struct.unpack_from("<H", buf, off)[0]
        tags = {}
        for q in range(n):
            tag, typ, cnt, val = struct.unpack_from("<HHII", buf, off + 2 + 12 * q)
            size = {2: 1, 3: 2, 4: 4}[typ] * cnt
            raw = buf[off + 2 + 12 * q + 8: off + 2 + 12 * q + 12] if size <= 4 else buf[val:val + size]
            if typ == 2:
                tags[tag] = raw[:cnt].rstrip(b"\x00").decode()
            else:
                tags[tag] = list(struct.unpack_from("<" + ("H" if typ == 3 else "I") * cnt, raw))
        strips = b"".join(buf[o:o + c] for o, c in zip(tags[273], tags[279]))
        pages.append((tags, strips))
        off = struct.unpack_from("<I", buf, off + 2 + 12 * n)[0]
    return pages


@pytest.mark.parametrize("shape", [(2, 5, 7), (3, 40, 61)])
def test_tiff_write_vectors_lzw_decodes_to_the_interleaved_components(tmp_path, shape):
    from lamino_amd import ingest

    rng = np.random.default_rng(4)
    m = rng.standard_normal((3,) + shape).astype(np.float32)
    m[1, 0] = 0.25            # long runs exercise the string table
    p = str(tmp_path / "vec.tiff")
    ingest.write_tiff_vectors(p, m)
    pages = _tiff_pages(open(p, "rb").read())
    assert len(pages) == shape[0]
    for i, (tags, strip) in enumerate(pages):
        assert tags[256] == [shape[2]] and tags[257] == [shape[1]]
        assert tags[258] == [32, 32, 32] and tags[339] == [3, 3, 3] and tags[277] == [3]
        assert tags[259] == [5] and tags[262] == [2]
        assert tags[270] == "3D vector field: x,y,z components; layout: [slice, row, col]"
        px = np.frombuffer(_lzw_decode(strip), dtype="<f4").reshape(shape[1], shape[2], 3)
        for c in range(3):
            np.testing.assert_array_equal(px[..., c], m[c, i])


def test_lzw_tables_fill_and_clear(tmp_path):
    """A page long enough to fill the 12-bit table several times (clear codes mid-strip)."""
    from lamino_amd import ingest

    rng = np.random.default_rng(5)
    m = rng.integers(0, 4, (3, 1, 300, 301)).astype(np.float32)
    p = str(tmp_path / "big.tiff")
    ingest.write_tiff_vectors(p, m)
    (tags, strip), = _tiff_pages(open(p, "rb").read())
    px = np.frombuffer(_lzw_decode(strip), dtype="<f4").reshape(300, 301, 3)
    for c in range(3):
        np.testing.assert_array_equal(px[..., c], m[c, 0])


def test_vti_layout_and_values(tmp_path):
    from lamino_amd import ingest

    rng = np.random.default_rng(6)
    m = rng.standard_normal((3, 3, 4, 5)).astype(np.float32)
    path = ingest.write_vti(str(tmp_path / "recon.tiff"), m)
    assert path.endswith("recon.vti") and os.path.exists(path)
    raw = open(path, "rb").read()
    head, rest = raw.split(b"   _", 1)
    text = head.decode()
    assert 'WholeExtent="0 4 0 3 0 2"' in text
    assert 'Origin="-2.5 -2 -1.5"' in text and 'Spacing="1 1 1"' in text
    assert '<DataArray type="Float64" Name="magnetization" NumberOfComponents="3" format="appended" offset="0"/>' in text
    n = 3 * 4 * 5
    (count,) = struct.unpack_from("<I", rest, 0)
    assert count == n * 3 * 8
    vals = np.frombuffer(rest[4:4 + count], dtype="<f8").reshape(n, 3)
    for c in range(3):
        np.testing.assert_array_equal(vals[:, c], m[c].ravel().astype(np.float64))
    assert rest[4 + count:] == b"\n  </AppendedData>\n</VTKFile>\n"


def test_read_angles_degree_rule(tmp_path):
    from lamino_amd import ingest

    deg = np.arange(0, 360, 2.5)
    p = tmp_path / "deg.txt"
    p.write_text("\n".join(f"{v}" for v in deg))
    got = ingest.read_angles(str(p))
    want = (deg.astype(np.float32).astype(np.float64) * math.pi / 180.0).astype(np.float32)
    np.testing.assert_array_equal(got, want.astype(np.float64))
    rad = np.linspace(0, 2 * math.pi, 16, endpoint=False)
    p2 = tmp_path / "rad.txt"
    p2.write_text(" ".join(f"{v:.9f}" for v in rad))
    np.testing.assert_allclose(ingest.read_angles(str(p2)), rad, rtol=0, atol=1e-6)
    # values are parsed as float: "6.283185307" rounds up past 2 pi, so the file counts as degrees
    p5 = tmp_path / "twopi.txt"
    p5.write_text("0 3.14159 6.283185307")
    assert float(np.float32(6.283185307)) > 2 * math.pi
    assert ingest.read_angles(str(p5))[2] < 0.2
    # only the largest value decides: negative degrees are kept as they are (config.h:78-81)
    p3 = tmp_path / "neg.txt"
    p3.write_text("-350 -180 -10 0")
    np.testing.assert_array_equal(ingest.read_angles(str(p3)), [-350.0, -180.0, -10.0, 0.0])
    p4 = tmp_path / "empty.txt"
    p4.write_text("\n")
    from lamino_amd._lib import LaminoError

    with pytest.raises(LaminoError, match="No angles found"):
        ingest.read_angles(str(p4))


def _cfg(tmp_path, body):
    p = tmp_path / "c.toml"
    p.write_text(body)
    return str(p)


def test_recon_params_defaults_and_regularizers(tmp_path):
    from lamino_amd import ingest

    cfg = ingest._toml(_cfg(tmp_path, """
[recon_params]
max_outer_iters = 7
recon_dims = [16, 64, 63]
tol = 0
"""))
    method, rd, p = ingest.parse_recon_params(cfg)
    assert method == 1 and rd == (15, 63, 63)
    assert p.max_iters == 50          # the template's max_outer_iters is not read
    assert p.tol == 0.0 and p.xtol == float(np.float32(1e-5))
    cfg = ingest._toml(_cfg(tmp_path, """
[recon_params]
max_iters = 30
recon_dims = [32, 128, 128]
[recon_params.regularizer]
method = "split_bregman"
[recon_params.regularizer.split_bregman]
lambda = 0.2
mu = 5
inner_iters = 4
"""))
    method, rd, p = ingest.parse_recon_params(cfg)
    assert (method, rd, p.max_iters, p.inner_iters) == (2, (31, 127, 127), 30, 4)
    assert p.lam == float(np.float32(0.2)) and p.mu == 5.0
    cfg = ingest._toml(_cfg(tmp_path, """
[recon_params]
recon_dims = [51, 256, 256]
[recon_params.regularizer]
method = "qGGMRF"
[recon_params.regularizer.qGGMRF]
sigma = 500.0
"""))
    method, rd, p = ingest.parse_recon_params(cfg)
    assert (method, rd, p.sigma, p.p, p.inner_iters) == (3, (51, 255, 255), 500.0, float(np.float32(1.2)), 3)


@pytest.mark.parametrize("body,msg", [
    ("", "Missing \\[recon_params\\]"),
    ("[recon_params]\nrecon_dims = [1, 2]\n", "must be an array of three integers"),
    ("[recon_params]\nrecon_dims = [1, 0, 3]\n", "recon_dims\\[1\\]' must be a positive integer"),
    ("[recon_params]\nrecon_dims = [3,3,3]\n[recon_params.regularizer]\nmethod = \"tv\"\n", "either 'qGGMRF'"),
    ("[recon_params]\nrecon_dims = [3,3,3]\n[recon_params.regularizer]\nmethod = \"qGGMRF\"\n",
     "Missing \\[recon_params.regularizer.qGGMRF\\]"),
])
def test_recon_params_errors(tmp_path, body, msg):
    from lamino_amd import ingest

    with pytest.raises(ingest.ConfigError, match=msg):
        ingest.parse_recon_params(ingest._toml(_cfg(tmp_path, body)))


def test_output_params_and_input_errors(tmp_path):
    from lamino_amd import ingest

    o = ingest.parse_output(ingest._toml(_cfg(tmp_path, "[output]\nformats = [\"vti\", \"tiff\", \"vti\"]\n")))
    assert o.filepath == "./recon.tiff" and o.formats == ["vti", "tiff"]
    assert ingest.parse_output(ingest._toml(_cfg(tmp_path, "[output]\nfilename = \"a/b.tiff\"\n"))).formats == \
        ["tiff", "vti"]
    with pytest.raises(ingest.ConfigError, match="invalid format 'png'"):
        ingest.parse_output(ingest._toml(_cfg(tmp_path, "[output]\nformats = [\"png\"]\n")))
    with pytest.raises(ingest.ConfigError, match="Missing \\[output\\]"):
        ingest.parse_output({})
    with pytest.raises(ingest.ConfigError, match="Missing \\[\\[input\\]\\]"):
        ingest.parse_input_datasets({})
    with pytest.raises(ingest.ConfigError, match="must have 'filename', 'angles', and 'gamma'"):
        ingest.parse_input_datasets({"input": [{"filename": "x"}]})
    with pytest.raises(ingest.ConfigError, match="Projection file does not exist"):
        ingest.parse_input_datasets({"input": [{"filename": str(tmp_path / "no.tif"), "angles": "a", "gamma": 0}]})
    with pytest.raises(ingest.ConfigError, match="TOML file does not exist"):
        ingest.load_config(str(tmp_path / "missing.toml"))
