"""CPU tests: the oracle against the reference's own known answers, hardware
and closed forms, and the HIP library's host surface (no GPU calls).

The oracle (oracle/librgl_oracle.so) is the parity checker for the GPU
tests, so every building block it uses is pinned here by something other
than a restatement of its own formula:
  * rglv_triangle.t.cxx:208-235 -- the 8x8 top-left fill-rule mask;
  * rglv_triangle.t.cxx:291-377 -- 2560^2 uv interpolation within 0.01;
  * rmlm_mat4.t.cxx:21-73       -- mat4 storage order and transforms;
  * the host CPU's own rcpps    -- the captured 2048-entry table;
  * ryg-srgb.h:183-219          -- the reference's sRGB encoder compiled
                                   from its header, all float bit patterns;
  * closed forms and invariants (watertight shared edges, tile-size and
    thread-count invariance of whole frames).
"""
from __future__ import annotations

import ctypes
import math
import os
import re
import struct
import subprocess

import numpy as np
import pytest

from oracle_backend import ORACLE_DIR, OracleGPU, lib as olib, render as oracle_render

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
REFERENCE = "/root/reference"


def f2u(x: float) -> int:
    return struct.unpack("<I", struct.pack("<f", x))[0]


def u2f(u: int) -> float:
    return struct.unpack("<f", struct.pack("<I", u & 0xFFFFFFFF))[0]


def read_inc_table(path: str) -> np.ndarray:
    body = open(path).read()
    body = body[body.index("{") + 1:body.rindex("}")]
    return np.array([int(t, 0) for t in re.findall(r"0x[0-9a-fA-F]+|\d+", body)], dtype=np.uint64)


# ---------------------------------------------------------------------------
# rcpps (rmlv_mvec4.hxx:630-650 oneover; hardware reciprocal estimate)
# ---------------------------------------------------------------------------
def test_rcpps_tables_agree():
    fixture = np.fromfile(os.path.join(GOLDEN, "rcpps_table.bin"), dtype="<u4").astype(np.uint64)
    assert fixture.size == 2048
    # fixture: rcpps(1 + i/2048) bit patterns, all in [0.5, 1); the .inc
    # tables keep their mantissas
    assert np.all(fixture >> 23 == 126)
    mant = fixture & 0x7FFFFF
    assert np.array_equal(read_inc_table(os.path.join(ORACLE_DIR, "rcpps_table.inc")), mant)
    assert np.array_equal(read_inc_table(os.path.join(ROOT, "rsr_amd", "csrc", "rcpps_table.inc")), mant)


def _host_is_intel() -> bool:
    try:
        return "GenuineIntel" in open("/proc/cpuinfo").read()
    except OSError:
        return False


@pytest.mark.skipif(not _host_is_intel(), reason="the table is the Intel rcpps; other vendors' estimates differ")
def test_rcpps_model_matches_host_hardware():
    exe = os.path.join(ORACLE_DIR, "gen_rcpps_table")
    if not os.path.exists(exe):
        subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)
    # every 97th float bit pattern (44M inputs) through the model vs _mm_rcp_ps
    r = subprocess.run([exe, "check", os.path.join(GOLDEN, "rcpps_table.bin"), "97"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr


def test_rcpps_closed_forms():
    L = olib()
    assert f2u(L.rglo_rcpps(1.0)) == 0x3F7FF000          # Intel's rcpps(1) = 1 - 2^-12
    rng = np.random.default_rng(7)
    for x in rng.uniform(1.0, 2.0, 2000).astype(np.float32):
        r = L.rglo_rcpps(float(x))
        assert abs(r * float(x) - 1.0) <= 1.5 * 2.0 ** -12      # the documented |rel err| bound
        for k in (-20, -3, 5, 40):                            # exponent shift is exact
            assert L.rglo_rcpps(float(x) * 2.0 ** k) == r * 2.0 ** -k
        assert L.rglo_rcpps(-float(x)) == -r
    assert L.rglo_rcpps(0.0) == float("inf")
    assert L.rglo_rcpps(-0.0) == float("-inf")
    assert L.rglo_rcpps(u2f(0x00000001)) == float("inf")     # denormal input
    assert L.rglo_rcpps(float("inf")) == 0.0
    assert np.isnan(L.rglo_rcpps(float("nan")))


# ---------------------------------------------------------------------------
# rsqrtps (rmlv_soa.hxx:243-248 normalize; hardware reciprocal square root
# estimate)
# ---------------------------------------------------------------------------
def test_rsqrtps_tables_agree():
    fixture = np.fromfile(os.path.join(GOLDEN, "rsqrtps_table.bin"), dtype="<u4").astype(np.uint64)
    assert fixture.size == 4096
    # rsqrtps of [1, 4) lies in (0.5, 1): every entry has biased exponent 126
    assert np.all(fixture >> 23 == 126)
    for path in (os.path.join(ORACLE_DIR, "rsqrtps_table.inc"), os.path.join(ROOT, "rsr_amd", "csrc", "rsqrtps_table.inc")):
        assert np.array_equal(read_inc_table(path), fixture), path


@pytest.mark.skipif(not _host_is_intel(), reason="the table is the Intel rsqrtps; other vendors' estimates differ")
def test_rsqrtps_model_matches_host_hardware():
    exe = os.path.join(ORACLE_DIR, "gen_rsqrtps_table")
    if not os.path.exists(exe):
        subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)
    # every 97th float bit pattern through the model vs _mm_rsqrt_ps
    r = subprocess.run([exe, "check", os.path.join(GOLDEN, "rsqrtps_table.bin"), "97"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr


def test_rsqrtps_closed_forms():
    L = olib()
    rng = np.random.default_rng(9)
    for x in rng.uniform(1.0, 4.0, 3000).astype(np.float32):
        r = L.rglo_rsqrtps(float(x))
        assert abs(r * math.sqrt(float(x)) - 1.0) <= 1.5 * 2.0 ** -12   # the documented |rel err| bound
        for k in (-30, -2, 3, 25):                                      # x * 4^k -> r * 2^-k exactly
            assert L.rglo_rsqrtps(float(x) * 4.0 ** k) == r * 2.0 ** -k
    assert L.rglo_rsqrtps(1.0) == L.rglo_rsqrtps(4.0) * 2.0
    assert L.rglo_rsqrtps(0.0) == float("inf")
    assert L.rglo_rsqrtps(-0.0) == float("-inf")
    assert L.rglo_rsqrtps(u2f(0x00000001)) == float("inf")          # denormal input
    assert L.rglo_rsqrtps(float("inf")) == 0.0
    assert np.isnan(L.rglo_rsqrtps(-1.0)) and np.isnan(L.rglo_rsqrtps(float("-inf")))
    assert np.isnan(L.rglo_rsqrtps(float("nan")))


def test_oneover_newton_step():
    L = olib()
    rng = np.random.default_rng(8)
    for x in rng.uniform(0.01, 1000.0, 3000).astype(np.float32):
        y = L.rglo_oneover(float(x))
        assert abs(y * float(x) - 1.0) < 2.0 ** -20       # one Newton step squares the estimate error
    assert L.rglo_oneover(2.0) == pytest.approx(0.5, rel=2 ** -21)


def test_cvtt_truncates_and_saturates_to_int_min():
    L = olib()
    cases = {0.0: 0, -0.5: 0, 0.99: 0, 1.5: 1, -1.5: -1, 2147483520.0: 2147483520,
             2147483648.0: -2147483648, -2147483648.0: -2147483648, 1e20: -2147483648,
             float("inf"): -2147483648, float("-inf"): -2147483648, float("nan"): -2147483648}
    for x, want in cases.items():
        assert L.rglo_cvtt(x) == want, x


# ---------------------------------------------------------------------------
# sRGB resolve (ryg-srgb.h:183-219 via rglr_canvas_util.hxx:33-38)
# ---------------------------------------------------------------------------
def test_srgb_tables_agree():
    assert np.array_equal(read_inc_table(os.path.join(ORACLE_DIR, "srgb8_table.inc")),
                          read_inc_table(os.path.join(ROOT, "rsr_amd", "csrc", "srgb8_table.inc")))


def test_srgb_encode_closed_forms():
    L = olib()
    assert L.rglo_to_srgb8(0.25) == 137 and L.rglo_to_srgb8(0.5) == 188
    assert L.rglo_to_srgb8(0.0) == 0 and L.rglo_to_srgb8(-3.0) == 0 and L.rglo_to_srgb8(float("nan")) == 0
    assert L.rglo_to_srgb8(1.0) == 255 and L.rglo_to_srgb8(7.0) == 255
    xs = np.linspace(0.0, 1.0, 20001, dtype=np.float32)
    got = np.array([L.rglo_to_srgb8(float(x)) for x in xs])
    assert np.all(np.diff(got) >= 0)                              # monotone
    x64 = xs.astype(np.float64)
    exact = np.where(x64 <= 0.0031308, 12.92 * x64, 1.055 * x64 ** (1 / 2.4) - 0.055) * 255.0
    assert np.max(np.abs(got - exact)) < 0.6                      # within 0.6 of the exact curve


@pytest.mark.skipif(not os.path.isdir(REFERENCE), reason="reference sources are not present")
def test_srgb_encode_matches_reference_header():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR, "ref"], check=True)
    # every 13th group of four float bit patterns (330M inputs; step 1 covers all 2^32)
    r = subprocess.run([os.path.join(ORACLE_DIR, "_ref", "ref_srgb_check"), "13"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout


def test_srgb8_to_float_table_matches_reference_fixture():
    from rsr_amd.scenes import srgb8_to_float_table
    want = np.array([int(w, 16) for w in open(os.path.join(GOLDEN, "srgb8_to_float.txt")).read().split()],
                    dtype=np.uint32)
    assert np.array_equal(srgb8_to_float_table().view(np.uint32), want)


# ---------------------------------------------------------------------------
# mat4 (rmlm_mat4.t.cxx:21-73)
# ---------------------------------------------------------------------------
def _xform4(m16, v):
    out = np.zeros(4, np.float32)
    m = np.ascontiguousarray(m16, np.float32)
    x = np.ascontiguousarray(v, np.float32)
    olib().rglo_xform4(m.ctypes.data, x.ctypes.data, out.ctypes.data)
    return out


def test_mat4_storage_is_column_major():
    from rsr_amd.scenes import mat4_rows
    m = mat4_rows(np.arange(16, dtype=np.float32).reshape(4, 4))
    assert list(m) == [0, 4, 8, 12, 1, 5, 9, 13, 2, 6, 10, 14, 3, 7, 11, 15]


def test_mat4_known_answers():
    from rsr_amd.scenes import mat4_rows
    ident = mat4_rows(np.eye(4, dtype=np.float32))
    assert list(_xform4(ident, [3, 4, 5, 6])) == [3, 4, 5, 6]
    scale = mat4_rows(np.diag([1, 2, 3, 1]).astype(np.float32))      # mat4::scale drops w
    assert list(_xform4(scale, [10, 10, 10, 1])) == [10, 20, 30, 1]
    tr = np.eye(4, dtype=np.float32); tr[2, 3] = -10
    assert list(_xform4(mat4_rows(tr), [3, 4, 5, 1])) == [3, 4, -5, 1]


def test_mat4_product_order():
    from rsr_amd.scenes import mat4_rows
    rng = np.random.default_rng(3)
    a = rng.integers(-8, 9, (4, 4)).astype(np.float32)
    b = rng.integers(-8, 9, (4, 4)).astype(np.float32)
    out = np.zeros(16, np.float32)
    ma, mb = mat4_rows(a), mat4_rows(b)          # keep the arrays alive across the call
    olib().rglo_mat4_mul(ma.ctypes.data, mb.ctypes.data, out.ctypes.data)
    assert np.array_equal(out, mat4_rows(a @ b))                      # small integers: exact
    v = rng.integers(-8, 9, 4).astype(np.float32)
    assert np.array_equal(_xform4(mat4_rows(a), v), a @ v)


# ---------------------------------------------------------------------------
# scalar TriangleRasterizer (rglv_triangle.hxx:79-167) via the reference's
# own test harness shape
# ---------------------------------------------------------------------------
def _scalar_triangle(w, h, pts, uvs=((0, 0), (0, 0), (0, 0))):
    out = np.zeros((h, w, 6), np.float32)
    p = np.ascontiguousarray(pts, np.float32)
    u = np.ascontiguousarray(uvs, np.float32)
    assert olib().rglo_scalar_triangle(w, h, p.ctypes.data, u.ctypes.data, out.ctypes.data) == 0
    return out


def test_fill_rule_reference_8x8():
    # rglv_triangle.t.cxx:208-235 ("dx top-left fill rules #1"): mask of BS.x > 0
    expected = ["........",
                ".XX.....",
                ".XXXX...",
                "..X.....",
                "........",
                "........",
                "........",
                "........"]
    out = _scalar_triangle(8, 8, [[2, 4, 0, 1], [6, 2, 0, 1], [1, 1, 0, 1]])
    got = ["".join("X" if out[y, x, 1] > 0 else "." for x in range(8)) for y in range(8)]
    assert got == expected


def test_uv_interpolation_reference_2560():
    # rglv_triangle.t.cxx:291-377: two triangles covering 2560^2 with uv = the
    # target's (x, y) in lower-left coordinates; every pixel within 0.01
    n = 2560
    ul, ur, ll, lr = (0, 0, 0, 1), (n, 0, 0, 1), (0, n, 0, 1), (n, n, 0, 1)
    a = _scalar_triangle(n, n, [ul, ll, ur], [(0, n), (0, 0), (n, n)])
    b = _scalar_triangle(n, n, [ur, ll, lr], [(n, n), (0, 0), (n, 0)])
    cov = a[..., 0] + b[..., 0]
    assert np.all(cov == 1.0)                                          # every pixel exactly once
    u = a[..., 4] + b[..., 4]
    v = a[..., 5] + b[..., 5]
    ex = np.arange(n, dtype=np.float32)[None, :] + 0.5
    ey = (n - np.arange(n, dtype=np.float32)[:, None]) - 0.5
    assert np.max(np.abs(u - ex)) < 0.01 and np.max(np.abs(v - ey)) < 0.01


def _area2(a, b, c):
    return (b[0] - a[0]) * (c[1] - a[1]) - (c[0] - a[0]) * (b[1] - a[1])


def test_shared_edges_are_watertight():
    # the top-left rule covers each pixel of a split convex quad exactly once,
    # and the union is the quad's interior sampled at pixel centres (brute
    # force); the rasteriser draws one winding (negative area, y down, as in
    # the 8x8 known answer) and nothing of the other
    rng = np.random.default_rng(11)
    w = h = 24
    ys, xs = np.mgrid[0:h, 0:w]
    cx, cy = xs + 0.5, ys + 0.5

    def inside(a, b, c):
        def e(p0, p1):
            return (p1[0] - p0[0]) * (cy - p0[1]) - (p1[1] - p0[1]) * (cx - p0[0])
        e1, e2, e3 = e(a, b), e(b, c), e(c, a)
        pos = (e1 > 0) & (e2 > 0) & (e3 > 0)
        neg = (e1 < 0) & (e2 < 0) & (e3 < 0)
        tie = (e1 == 0) | (e2 == 0) | (e3 == 0)
        return pos | neg, tie

    done = 0
    while done < 41:
        ang = np.sort(rng.uniform(0, 2 * np.pi, 4))
        ctr = rng.uniform(6, 18, 2)
        rad = rng.uniform(3, 11, 2)
        q = np.round(16 * np.stack([ctr[0] + rad[0] * np.cos(ang), ctr[1] + rad[1] * np.sin(ang)], 1)) / 16
        if done == 0:   # a square whose diagonal runs through pixel centres
            q = np.array([[4.0, 4.0], [20.0, 4.0], [20.0, 20.0], [4.0, 20.0]])
        if _area2(q[0], q[1], q[2]) > 0:
            q = q[::-1].copy()
        if not all(_area2(q[i], q[(i + 1) % 4], q[(i + 2) % 4]) < 0 for i in range(4)):
            continue                                                   # keep strictly convex quads
        done += 1
        p = [[x, y, 0, 1] for x, y in q]
        t1 = _scalar_triangle(w, h, [p[0], p[1], p[2]])[..., 0]
        t2 = _scalar_triangle(w, h, [p[0], p[2], p[3]])[..., 0]
        assert not _scalar_triangle(w, h, [p[2], p[1], p[0]])[..., 0].any()
        assert np.all(t1 + t2 <= 1.0)
        i1, k1 = inside(q[0], q[1], q[2])
        i2, k2 = inside(q[0], q[2], q[3])
        sure = ~(k1 | k2)
        assert np.array_equal((t1 > 0)[sure], i1[sure])
        assert np.array_equal((t2 > 0)[sure], i2[sure])
        # pixel centres exactly on the shared diagonal (strictly between its
        # ends) are inside the quad: exactly one of the two triangles takes them
        dx, dy = q[2][0] - q[0][0], q[2][1] - q[0][1]
        t = ((cx - q[0][0]) * dx + (cy - q[0][1]) * dy) / (dx * dx + dy * dy)
        on_diag = (dx * (cy - q[0][1]) - dy * (cx - q[0][0]) == 0) & (t > 0) & (t < 1)
        assert np.all((t1 + t2)[on_diag] == 1.0)


# ---------------------------------------------------------------------------
# whole frames: invariants the reference's design guarantees
# ---------------------------------------------------------------------------
def test_oracle_frame_tile_size_and_thread_invariance():
    from rsr_amd import scenes
    base = scenes.scene_sphere_ground()
    c0, d0, p0, s0 = oracle_render(base, threads=1)
    for tb, th in (((4, 4), 3), ((16, 16), 8), ((2, 6), 5)):
        sc = scenes.scene_sphere_ground()
        sc.tile_blocks = tb
        c, d, p, s = oracle_render(sc, threads=th)
        assert np.array_equal(c, c0) and np.array_equal(d.view(np.uint32), d0.view(np.uint32))
        assert np.array_equal(p, p0)
        assert s["fragments"] == s0["fragments"]
    cov = d0 < 1.0
    assert np.array_equal(cov, p0 != 0xFFFFFFFF)          # coverage = depth < clear = a primitive id
    assert 0.2 < cov.mean() < 0.9


def test_oracle_frame_clipped_triangles_present():
    from rsr_amd import scenes
    sc = scenes.scene_clip_floor(160, 96, 8)
    c, d, p, s = oracle_render(sc, threads=4)
    assert s["clipped"] > 0 and s["clipped_tile_pairs"] > 0
    assert (d < 1.0).any()


# ---------------------------------------------------------------------------
# the HIP library's host surface (loads on a machine without a GPU)
# ---------------------------------------------------------------------------
def _header_functions():
    text = open(os.path.join(ROOT, "include", "rgl_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    return sorted(set(re.findall(r"\b(rgl_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from rsr_amd import rgl
    L = rgl.lib()
    names = _header_functions()
    assert len(names) >= 40
    for n in names:
        assert hasattr(L, n), n
    assert sorted(rgl.EXPORTS) == names


def test_mipmap_chain_matches_host_stack():
    from rsr_amd import rgl, scenes
    rng = np.random.default_rng(5)
    for dim in (4, 16, 256):
        rgba8 = rng.integers(0, 256, (dim, dim, 4), dtype=np.uint8)
        want = scenes.make_mipmap_stack(rgba8)
        buf = np.zeros(dim * dim * 8, np.float32)
        buf[:dim * dim * 4] = want[:dim * dim * 4]
        rgl.make_mipmap(buf, dim)
        assert np.array_equal(buf.view(np.uint32), want.view(np.uint32))
    # constant texture: every level is the constant
    buf = np.full(64 * 64 * 8, 0.0, np.float32)
    buf[:64 * 64 * 4] = 0.375
    rgl.make_mipmap(buf, 64)
    lv = buf.reshape(128, 64, 4)
    assert np.all(lv[64:65, :32] == 0.375) and np.all(lv[64 + 32 + 16 + 8 + 4 + 2, 0] == 0.375)


def test_mipmap_rejects_bad_sizes():
    from rsr_amd import rgl
    buf = np.zeros(12 * 12 * 8, np.float32)
    with pytest.raises(rgl.RglError):
        rgl.make_mipmap(buf, 12)
