"""The asset path's host side: the PNG decoder against PIL and against
images written with every scanline filter, the OBJ / MTL loader and
MakeArray on a hand-written model."""
import io
import struct
import zlib

import numpy as np
import pytest

from rsr_amd import assets

PIL = pytest.importorskip("PIL.Image")


def _pil_png(arr, mode):
    buf = io.BytesIO()
    PIL.fromarray(arr, mode).save(buf, format="PNG")
    return buf.getvalue()


@pytest.mark.parametrize("mode", ["RGBA", "RGB", "L", "LA", "P"])
def test_decode_png_matches_pil(mode):
    rng = np.random.default_rng(len(mode))
    h, w = 37, 53
    if mode == "P":
        img = PIL.fromarray(rng.integers(0, 255, (h, w, 3), dtype=np.uint8), "RGB").quantize(64)
        buf = io.BytesIO()
        img.save(buf, format="PNG")
        data = buf.getvalue()
    else:
        shape = {"RGBA": (h, w, 4), "RGB": (h, w, 3), "L": (h, w), "LA": (h, w, 2)}[mode]
        # smooth content so that the encoder picks several filters
        base = np.add.outer(np.arange(h), np.arange(w)) * 3
        arr = (base[..., None] + rng.integers(0, 8, shape if len(shape) == 3 else shape + (1,))).astype(np.uint8)
        arr = arr.reshape(shape)
        data = _pil_png(arr, mode)
    got = assets.decode_png(data)
    want = np.asarray(PIL.open(io.BytesIO(data)).convert("RGBA"))
    assert np.array_equal(got, want)


def _png_with_filters(img):
    """An RGBA8 PNG whose rows cycle through filter types 0-4."""
    h, w, _ = img.shape
    bpp, stride = 4, w * 4
    raw, prev = bytearray(), np.zeros(stride, np.int32)
    for y in range(h):
        ft = y % 5
        cur = img[y].reshape(-1).astype(np.int32)
        out = np.zeros(stride, np.int32)
        for i in range(stride):
            a = cur[i - bpp] if i >= bpp else 0
            b = prev[i]
            c = prev[i - bpp] if i >= bpp else 0
            if ft == 0:
                p = 0
            elif ft == 1:
                p = a
            elif ft == 2:
                p = b
            elif ft == 3:
                p = (a + b) >> 1
            else:
                pa, pb, pc = abs(b - c), abs(a - c), abs(a + b - 2 * c)
                p = a if pa <= pb and pa <= pc else (b if pb <= pc else c)
            out[i] = (cur[i] - p) & 255
        raw += bytes([ft]) + bytes(out.astype(np.uint8))
        prev = cur

    def chunk(kind, body):
        return struct.pack(">I", len(body)) + kind + body + struct.pack(">I", zlib.crc32(kind + body) & 0xFFFFFFFF)
    return b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, 6, 0, 0, 0)) + \
        chunk(b"IDAT", zlib.compress(bytes(raw))) + chunk(b"IEND", b"")


def test_decode_png_every_filter():
    img = np.random.default_rng(7).integers(0, 256, (11, 9, 4), dtype=np.uint8)
    assert np.array_equal(assets.decode_png(_png_with_filters(img)), img)


def test_obj_loader_and_make_array(tmp_path):
    (tmp_path / "m.mtl").write_text("newmtl red\nKd 0.8 0.1 0.1\nKa 0.1 0 0\nnewmtl blue # comment\nKd 0.1 0.2 0.9\n")
    (tmp_path / "m.obj").write_text(
        "mtllib m.mtl\n"
        "v 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nv 2 0 0\nv 2 1 0\n"
        "vt 0 0\nvt 1 0\nvt 1 1\n"
        "vn 0 0 1\n"
        "usemtl red\n"
        "f 1/1/1 2/2/1 3/3/1 4/1/1   # a quad: two triangles\n"
        "usemtl blue\n"
        "f 2//1 5//1 6//1 3//1\n")
    m = assets.load_obj(str(tmp_path / "m.obj"))
    assert len(m.faces) == 4 and [f[3] for f in m.faces] == [0, 0, 1, 1]
    assert m.faces[1][0] == [0, 2, 3]                    # fan (0, i + 1, i + 2)
    assert m.faces[2][2] == [0, 0, 0]                    # "p//n": texture index 1 - 1
    for n in m.face_normal:
        assert np.allclose(n, [0, 0, 1])
    assert np.allclose(m.smooth_normal[0], [0, 0, 1])
    (pos, kd), idx = assets.make_array(m, "PD")
    # vertices shared by both faces of one material merge; the two
    # materials keep vertices 2 and 3 apart
    assert len(idx) == 12 and idx.max() == 7
    assert len(pos) == 12   # 8 vertices and 4 of padding (VertexArray pad adds 1-4)
    assert np.allclose(kd[idx[0]], [0.8, 0.1, 0.1]) and np.allclose(kd[idx[-1]], [0.1, 0.2, 0.9])
    assert np.all(pos[8:] == 0)
