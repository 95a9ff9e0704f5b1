"""The asset path's host side: PNG decoding (as picopng's decodePNG feeds
rglr::LoadPNG, rglr_texture_load.cxx:39-66) and Wavefront OBJ / MTL
loading into the mesh the viewer's $mesh node draws (rglv_obj.cxx,
rglv_mesh.cxx:66-87 normals, rglv_mesh_util.cxx:24-135 MakeArray).

The texel work of a loaded image (sRGB decode, power-of-two resample, mip
chain) runs on the device: rgl.texture_from_rgba8.
"""
from __future__ import annotations

import os
import struct
import zlib
from dataclasses import dataclass, field

import numpy as np

# ---------------------------------------------------------------------------
# PNG
# ---------------------------------------------------------------------------
_PNG_SIG = b"\x89PNG\r\n\x1a\n"


def _unfilter(raw: bytes, w: int, h: int, bpp: int) -> np.ndarray:
    """PNG scanline filters 0-4 (None, Sub, Up, Average, Paeth)."""
    stride = w * bpp
    out = np.zeros((h, stride), np.uint8)
    prev = np.zeros(stride, np.int32)
    pos = 0
    for y in range(h):
        ft = raw[pos]
        line = np.frombuffer(raw, np.uint8, stride, pos + 1).astype(np.int32)
        pos += 1 + stride
        cur = np.zeros(stride, np.int32)
        if ft == 0:
            cur = line
        elif ft == 2:
            cur = (line + prev) & 255
        elif ft in (1, 3, 4):
            for i in range(stride):
                a = cur[i - bpp] if i >= bpp else 0
                b = prev[i]
                if ft == 1:
                    p = a
                elif ft == 3:
                    p = (a + b) >> 1
                else:
                    c = prev[i - bpp] if i >= bpp else 0
                    pa, pb, pc = abs(b - c), abs(a - c), abs(a + b - 2 * c)
                    p = a if pa <= pb and pa <= pc else (b if pb <= pc else c)
                cur[i] = (line[i] + p) & 255
        else:
            raise ValueError(f"PNG: bad filter type {ft}")
        out[y] = cur
        prev = cur
    return out


def decode_png(data: bytes) -> np.ndarray:
    """An 8-bit, non-interlaced PNG (greyscale, RGB, palette, grey + alpha,
    RGBA) to H x W x 4 RGBA8, alpha 255 where the image has none."""
    if data[:8] != _PNG_SIG:
        raise ValueError("not a PNG")
    pos, idat, plte, trns, hdr = 8, [], None, None, None
    while pos < len(data):
        n, kind = struct.unpack(">I4s", data[pos:pos + 8])
        body = data[pos + 8:pos + 8 + n]
        pos += 12 + n
        if kind == b"IHDR":
            hdr = struct.unpack(">IIBBBBB", body)
        elif kind == b"PLTE":
            plte = np.frombuffer(body, np.uint8).reshape(-1, 3)
        elif kind == b"tRNS":
            trns = np.frombuffer(body, np.uint8)
        elif kind == b"IDAT":
            idat.append(body)
        elif kind == b"IEND":
            break
    if hdr is None:
        raise ValueError("PNG: no IHDR")
    w, h, depth, ctype, _, _, interlace = hdr
    if depth != 8 or interlace != 0:
        raise ValueError("PNG: only 8-bit non-interlaced images")
    chans = {0: 1, 2: 3, 3: 1, 4: 2, 6: 4}[ctype]
    px = _unfilter(zlib.decompress(b"".join(idat)), w, h, chans).reshape(h, w, chans)
    out = np.full((h, w, 4), 255, np.uint8)
    if ctype == 0:
        out[..., :3] = px
    elif ctype == 2:
        out[..., :3] = px
    elif ctype == 3:
        out[..., :3] = plte[px[..., 0]]
        if trns is not None:
            alpha = np.full(256, 255, np.uint8)
            alpha[:len(trns)] = trns
            out[..., 3] = alpha[px[..., 0]]
    elif ctype == 4:
        out[..., :3] = px[..., :1]
        out[..., 3] = px[..., 1]
    else:
        out[:] = px
    return out


def load_png(path: str) -> np.ndarray:
    with open(path, "rb") as f:
        return decode_png(f.read())


# ---------------------------------------------------------------------------
# OBJ / MTL
# ---------------------------------------------------------------------------
@dataclass
class Material:
    name: str = "__none__"
    ka: tuple = (0.0, 0.0, 0.0)
    kd: tuple = (0.0, 0.0, 0.0)
    ks: tuple = (0.0, 0.0, 0.0)
    specpow: float = 1.0
    d: float = 1.0
    imagename: str = ""


@dataclass
class Mesh:
    """rglv::Mesh: positions, normals, texture coordinates, triangles
    (position, normal and texture indices and a material), face normals
    and smooth vertex normals."""
    name: str = ""
    position: list = field(default_factory=list)
    normal: list = field(default_factory=list)
    texture: list = field(default_factory=list)
    faces: list = field(default_factory=list)     # (pos[3], nrm[3], tex[3], material)
    materials: list = field(default_factory=list)
    face_normal: list = field(default_factory=list)
    smooth_normal: list = field(default_factory=list)


def _floats(text, n):
    v = [np.float32(x) for x in text.split()[:n]]
    while len(v) < n:
        v.append(np.float32(0.0))
    return tuple(v)


def parse_mtl(text: str) -> list:
    """ParseMtllib (rglv_obj.cxx:130-178)."""
    out, m = [], Material()
    for line in text.splitlines():
        line = line.split("#", 1)[0].strip()
        if not line:
            continue
        cmd, _, rest = line.partition(" ")
        rest = rest.strip()
        if cmd == "newmtl":
            if m.name != "__none__":
                out.append(m)
                m = Material()
            m.name = rest
        elif cmd == "Ka":
            m.ka = _floats(rest, 3)
        elif cmd == "Kd":
            m.kd = _floats(rest, 3)
        elif cmd == "Ks":
            m.ks = _floats(rest, 3)
        elif cmd == "Ns":
            m.specpow = _floats(rest, 1)[0]
        elif cmd == "d":
            m.d = _floats(rest, 1)[0]
        elif cmd == "map_Kd":
            m.imagename = rest
    if m.name != "__none__":
        out.append(m)
    return out


def _vnormalize(v):
    v = np.asarray(v, np.float32)
    ln = np.float32(np.sqrt(np.float32(v[0] * v[0] + v[1] * v[1] + v[2] * v[2])))
    return v / ln


def load_obj(path: str) -> Mesh:
    """LoadOBJ (rglv_obj.cxx:182-282): v / vn / vt / f (fans of polygons,
    1-based "p/t/n" or "p//n" indices, missing ones 1), usemtl, mtllib;
    then ComputeFaceAndVertexNormals (rglv_mesh.cxx:66-87)."""
    mesh = Mesh(name=os.path.basename(path))
    objdir = os.path.dirname(path)
    mat_idx = -1
    with open(path) as f:
        for line in f:
            line = line.split("#", 1)[0].strip()
            if not line:
                continue
            cmd, _, rest = line.partition(" ")
            rest = rest.strip()
            if cmd == "mtllib":
                with open(os.path.join(objdir, rest)) as mf:
                    mesh.materials = parse_mtl(mf.read())
            elif cmd == "usemtl":
                names = [m.name for m in mesh.materials]
                if rest not in names:
                    raise RuntimeError(f'usemtl "{rest}" not found in mtl')
                mat_idx = names.index(rest)
            elif cmd == "v":
                mesh.position.append(_floats(rest, 3))
            elif cmd == "vn":
                mesh.normal.append(_floats(rest, 3))
            elif cmd == "vt":
                mesh.texture.append(_floats(rest, 2))
            elif cmd == "f":
                idx = []
                for chunk in rest.split(" "):
                    if not chunk:
                        continue
                    parts = (chunk.split("/") + ["", ""])[:3]
                    vals = [int(p) if p else 1 for p in parts]
                    idx.append((vals[0] - 1, vals[1] - 1, vals[2] - 1))   # (vv, vt, vn)
                for i in range(len(idx) - 2):
                    tri = (idx[0], idx[i + 1], idx[i + 2])
                    mesh.faces.append(([t[0] for t in tri], [t[2] for t in tri], [t[1] for t in tri], mat_idx))
    pos = [np.asarray(p, np.float32) for p in mesh.position]
    smooth = [np.zeros(3, np.float32) for _ in pos]
    for fp, _, _, _ in mesh.faces:
        i0, i1, i2 = fp
        d = np.cross(pos[i1] - pos[i0], pos[i2] - pos[i0]).astype(np.float32)
        mesh.face_normal.append(_vnormalize(d))
        for i in fp:
            smooth[i] = smooth[i] + d
    mesh.smooth_normal = [_vnormalize(n) for n in smooth]
    return mesh


def make_array(mesh: Mesh, spec: str):
    """MakeArray (rglv_mesh_util.cxx:24-135): one vertex per distinct tuple
    of the attributes the spec letters name (P position, N normal, T uv,
    F face normal, A/D/S material ka/kd/ks, c smooth normal, b the smooth
    normal mixed 0.666 toward the face normal and normalised), triangle
    indices, then the reference's padding (always 1-4 zero vertices up to a
    multiple of 4).  Returns ([n x 3 float32 array per letter], uint16
    indices)."""
    cols = [[] for _ in spec]
    index, idx = {}, []
    for fi, (fp, fn, ft, mi) in enumerate(mesh.faces):
        mat = mesh.materials[mi] if 0 <= mi < len(mesh.materials) else Material()
        for i in range(3):
            d = []
            for ch in spec:
                if ch == "P":
                    v = mesh.position[fp[i]]
                elif ch == "N":
                    v = mesh.normal[fn[i]]
                elif ch == "T":
                    v = (*mesh.texture[ft[i]], np.float32(0.0))
                elif ch == "F":
                    v = mesh.face_normal[fi]
                elif ch == "A":
                    v = mat.ka
                elif ch == "D":
                    v = mat.kd
                elif ch == "S":
                    v = mat.ks
                elif ch == "c":
                    v = mesh.smooth_normal[fp[i]]
                elif ch == "b":
                    s, f = mesh.smooth_normal[fp[i]], mesh.face_normal[fi]
                    v = _vnormalize(s + (f - s) * np.float32(0.666))
                else:
                    raise ValueError(f"unknown spec {ch}")
                d.append(tuple(np.float32(x) for x in v))
            key = tuple(d)
            if key not in index:
                index[key] = len(index)
                for c, v in zip(cols, d):
                    c.append(v)
            k = index[key]
            if k >= 65536:
                raise ValueError("more than 65536 vertices")
            idx.append(k)
    need = 4 - len(cols[0]) % 4
    for c in cols:
        c.extend([(0.0, 0.0, 0.0)] * need)
    return [np.asarray(c, np.float32) for c in cols], np.asarray(idx, np.uint16)
