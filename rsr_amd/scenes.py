"""Seeded synthetic scenes for the rgl pipeline workloads.

Every array is float32 / uint16 numpy data that both the HIP pipeline and
any other backend receive unchanged, so the scene generator's own floating
point (trigonometry, normalisation) never affects parity.

Layouts follow the reference's conventions:
  * vertex attributes are SoA float arrays padded to a multiple of 4
    (rglv::Float3Array / VertexArray_F3F3F3::pad, rglv_vao.hxx:141-272);
  * attribute slots: 0-2 position, 3-5 normal, 9-10 uv, 15 per-instance
    mat4 array (viewer/shaders.hxx:83-112, 366-367);
  * mat4 is column-major, built from row-major arguments
    (rmlm_mat4.hxx:156-164);
  * textures are float RGBA with the mip chain stacked below the base
    level in an N x 2N buffer (rglr_texture.cxx:33-81).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

# GL constants (rglv_gl.hxx:20-62)
GL_TRIANGLES = 0
GL_CULL_FACE, GL_SCISSOR_TEST, GL_BLEND, GL_DEPTH_TEST = 1, 2, 3, 4
GL_FRONT, GL_BACK = 1, 2
GL_UNSIGNED_SHORT = 1
GL_NEAREST_MIPMAP_NEAREST, GL_LINEAR_MIPMAP_NEAREST = 0, 1
GL_COLOR_BUFFER_BIT, GL_DEPTH_BUFFER_BIT = 1, 2
GL_LESS, GL_LEQUAL, GL_EQUAL = 0, 1, 2
GL_DEPTH_ATTACHMENT, GL_COLOR_ATTACHMENT0 = 0, 2
RB_COLOR_DEPTH, RB_RGBF32, RB_RGBAF32, RB_F32 = 0, 1, 2, 3
RGL_HINT_READ4, RGL_HINT_DENSE = 1, 2

# program ids: 0 rglv::BaseProgram, 1 DefaultPost, 2 ExposurePost, 3 IQPost, 4 Amy, 5 Depth, 6 Many,
# 7 OBJ1, 8 OBJ2, 9 OBJ2S, 10 Envmap, 26 Text, 41 Pattern and 65 AlphaTexture
# are the reference's (viewer/shaders.hxx, shaders_envmap.hxx); 80, 81 and 82 are this project's shaders.
PGM_BASE = 0
PGM_DEFAULT_POST = 1
PGM_EXPOSURE_POST = 2
PGM_IQ_POST = 3
PGM_AMY = 4
PGM_DEPTH = 5
PGM_MANY = 6
PGM_OBJ1 = 7
PGM_OBJ2 = 8
PGM_OBJ2S = 9
PGM_ENVMAP = 10
PGM_TEXT = 26
PGM_PATTERN = 41
PGM_ALPHA_TEXTURE = 65
PGM_LAMBERT_TEX = 80
PGM_MANY_BLINN_PHONG = 81
PGM_GGX_ENV = 82

# dispatch keys (viewer/shaders.cxx:48-52 and :71)
KEY_RGBF32_F32_LESS = 0x6e2
KEY_COLOR_DEPTH_LESS = 0x62


def mat4_rows(*rows) -> np.ndarray:
    """mat4 from row-major arguments, stored column-major (rmlm_mat4.hxx:156-164)."""
    m = np.array(rows, dtype=np.float64).reshape(4, 4)
    return np.ascontiguousarray(m.T.reshape(16)).astype(np.float32)


def perspective2(fovy: float, aspect: float, znear: float, zfar: float) -> np.ndarray:
    """gluPerspective (rglv_math.cxx:152-167), float32 like the reference."""
    h = np.float32(math.tan((fovy / 2) / 180 * math.pi) * znear)
    w = np.float32(h * np.float32(aspect))
    l, r, b, t = -w, w, -h, h
    n, f = np.float32(znear), np.float32(zfar)
    two = np.float32(2)
    return mat4_rows(
        (two * n) / (r - l), 0, (r + l) / (r - l), 0,
        0, (two * n) / (t - b), (t + b) / (t - b), 0,
        0, 0, -((f + n) / (f - n)), -((two * f * n) / (f - n)),
        0, 0, -1, 0)


def look_at(eye, center, up) -> np.ndarray:
    """gluLookAt (rglv_math.cxx:129-146)."""
    eye, center, up = (np.asarray(v, np.float64) for v in (eye, center, up))
    fwd = center - eye
    fwd /= np.linalg.norm(fwd)
    side = np.cross(fwd, up)
    side /= np.linalg.norm(side)
    up2 = np.cross(side, fwd)
    rot = np.array([[*side, 0], [*up2, 0], [*(-fwd), 0], [0, 0, 0, 1]])
    tr = np.eye(4)
    tr[:3, 3] = -eye
    m = rot @ tr
    return np.ascontiguousarray(m.T.reshape(16)).astype(np.float32)


def pad4(a: np.ndarray, fill: float = 0.0) -> np.ndarray:
    n = len(a)
    p = (n + 3) & ~3
    out = np.full(p, fill, dtype=np.float32)
    out[:n] = a
    return out


# ---------------------------------------------------------------------------
# textures
# ---------------------------------------------------------------------------
def xorshift32_stream(seed: int, n: int) -> np.ndarray:
    x = seed & 0xFFFFFFFF
    if x == 0:
        x = 0x9E3779B9
    out = np.empty(n, dtype=np.uint32)
    for i in range(n):
        x ^= (x << 13) & 0xFFFFFFFF
        x ^= x >> 17
        x ^= (x << 5) & 0xFFFFFFFF
        out[i] = x
    return out


def srgb8_to_float_table() -> np.ndarray:
    """D3D10 sRGB8 -> linear table of ryg::srgb8_to_float (ryg-srgb.h:33-56),
    which rglr_texture_load.cxx:55-64 applies to texture texels: c = i * (1/255)
    in float, c / 12.92 below 0.04045, else the double-precision pow of the
    float base with the float exponent 2.4f, rounded to float."""
    c = np.arange(256, dtype=np.float32) * np.float32(1.0 / 255.0)
    base = ((c + np.float32(0.055)) / np.float32(1.055)).astype(np.float64)
    lin = np.where(c <= np.float32(0.04045), c / np.float32(12.92), base ** np.float64(np.float32(2.4)))
    return lin.astype(np.float32)


def value_noise_rgba8(dim: int, seed: int, lattice: int = 16) -> np.ndarray:
    """dim x dim RGBA8 value noise: a lattice x lattice grid of xorshift32
    values, smoothstep-bilinear upsampled with wrap-around."""
    vals = xorshift32_stream(seed, lattice * lattice * 3).reshape(lattice, lattice, 3)
    vals = (vals >> 24).astype(np.float64) / 255.0
    t = (np.arange(dim) + 0.5) * lattice / dim
    i0 = np.floor(t).astype(np.int64) % lattice
    i1 = (i0 + 1) % lattice
    f = t - np.floor(t)
    s = f * f * (3 - 2 * f)
    rows0 = vals[i0][:, i0] * (1 - s)[None, :, None] + vals[i0][:, i1] * s[None, :, None]
    rows1 = vals[i1][:, i0] * (1 - s)[None, :, None] + vals[i1][:, i1] * s[None, :, None]
    img = rows0 * (1 - s)[:, None, None] + rows1 * s[:, None, None]
    rgba = np.empty((dim, dim, 4), dtype=np.uint8)
    rgba[..., :3] = np.clip(np.round(img * 255.0), 0, 255).astype(np.uint8)
    rgba[..., 3] = 255
    return rgba


def make_mipmap_stack(rgba8: np.ndarray) -> np.ndarray:
    """Float RGBA texture with the box-filtered mip chain stacked below the
    base level: an N x 2N buffer of RGBA floats, row stride N
    (Texture::maybe_make_mipmap, rglr_texture.cxx:33-81)."""
    n = rgba8.shape[0]
    assert rgba8.shape[:2] == (n, n) and n & (n - 1) == 0
    table = srgb8_to_float_table()
    buf = np.zeros((2 * n, n, 4), dtype=np.float32)
    buf[:n] = table[rgba8]
    buf[:n, :, 3] = rgba8[..., 3].astype(np.float32) / np.float32(255.0)
    src_row, size, dst_row = 0, n, n
    while size > 1:
        lv = buf[src_row:src_row + size, :size]
        a = lv[0::2, 0::2]
        b = lv[0::2, 1::2]
        c = lv[1::2, 0::2]
        d = lv[1::2, 1::2]
        s = ((a + b) + c) + d
        half = size // 2
        buf[dst_row:dst_row + half, :half] = s / np.float32(4.0)
        src_row, dst_row, size = dst_row, dst_row + half, half
    return np.ascontiguousarray(buf.reshape(-1))


# ---------------------------------------------------------------------------
# meshes
# ---------------------------------------------------------------------------
def uv_sphere(slices: int, stacks: int, radius: float, center):
    """UV sphere with a seam column: (stacks+1)*(slices+1) vertices and
    2*slices*(stacks-1) triangles (pole bands emit one triangle per slice)."""
    cx, cy, cz = center
    pos, nrm, uv = [], [], []
    for i in range(stacks + 1):
        phi = math.pi * i / stacks
        for j in range(slices + 1):
            th = 2 * math.pi * j / slices
            nx, ny, nz = math.sin(phi) * math.cos(th), math.cos(phi), math.sin(phi) * math.sin(th)
            pos.append((cx + radius * nx, cy + radius * ny, cz + radius * nz))
            nrm.append((nx, ny, nz))
            uv.append((j / slices, 1.0 - i / stacks))
    idx = []
    row = slices + 1
    for i in range(stacks):
        for j in range(slices):
            a, b = i * row + j, i * row + j + 1
            c, d = (i + 1) * row + j, (i + 1) * row + j + 1
            if i != 0:
                idx += [a, b, c]  # ccw seen from outside
            if i != stacks - 1:
                idx += [b, d, c]
    return (np.array(pos, np.float32), np.array(nrm, np.float32),
            np.array(uv, np.float32), np.array(idx, np.uint16))


def icosphere(subdiv: int):
    """Subdivided icosahedron on the unit sphere: 10*4^s+2 vertices,
    20*4^s triangles, shared vertices."""
    t = (1.0 + math.sqrt(5.0)) / 2.0
    verts = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t),
             (0, -1, -t), (0, 1, -t), (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    verts = [np.array(v) / np.linalg.norm(v) for v in verts]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4),
             (11, 10, 2), (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8),
             (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(subdiv):
        cache = {}

        def mid(a, b):
            key = (min(a, b), max(a, b))
            if key not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]

        nf = []
        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    p = np.array(verts, np.float32)
    u = (np.arctan2(p[:, 2], p[:, 0]) / (2 * math.pi) + 0.5).astype(np.float32)
    v = (np.arcsin(np.clip(p[:, 1], -1, 1)) / math.pi + 0.5).astype(np.float32)
    return p, p.copy(), np.stack([u, v], 1), np.array(faces, np.uint16).reshape(-1)


def cube():
    """Unit cube (half-extent 1) with per-face normals and uvs: 24 vertices,
    12 triangles, counter-clockwise seen from outside."""
    pos, nrm, uv, idx = [], [], [], []
    for axis in range(3):
        for sign in (1.0, -1.0):
            n = np.zeros(3); n[axis] = sign
            u_ax = np.zeros(3); u_ax[(axis + 1) % 3] = 1.0
            v_ax = np.cross(n, u_ax)
            base = len(pos)
            for (a, b) in ((-1, -1), (1, -1), (1, 1), (-1, 1)):
                pos.append(n + a * u_ax + b * v_ax)
                nrm.append(n)
                uv.append(((a + 1) / 2, (b + 1) / 2))
            idx += [base, base + 1, base + 2, base, base + 2, base + 3]
    return (np.array(pos, np.float32), np.array(nrm, np.float32), np.array(uv, np.float32),
            np.array(idx, np.uint16))


def random_rotation(rng: np.random.RandomState) -> np.ndarray:
    z = rng.uniform(-1, 1)
    a = rng.uniform(0, 2 * math.pi)
    r = math.sqrt(1 - z * z)
    axis = np.array([r * math.cos(a), r * math.sin(a), z])
    ang = rng.uniform(0, 2 * math.pi)
    k = np.array([[0, -axis[2], axis[1]], [axis[2], 0, -axis[0]], [-axis[1], axis[0], 0]])
    return np.eye(3) + math.sin(ang) * k + (1 - math.cos(ang)) * (k @ k)


# ---------------------------------------------------------------------------
# scene = arrays + the GL call list that draws them
# ---------------------------------------------------------------------------
@dataclass
class Mesh:
    buffers: dict          # slot -> float32 array (padded)
    indices: np.ndarray    # uint16
    instances: int = 1     # >1 -> DrawElementsInstanced


@dataclass
class Scene:
    name: str
    width: int
    height: int
    tile_blocks: tuple = (8, 8)
    clear_color: tuple = (0.1, 0.1, 0.15)
    view: np.ndarray = None
    projection: np.ndarray = None
    program: int = PGM_LAMBERT_TEX
    uniforms: np.ndarray = None
    texture: np.ndarray = None
    texture_dim: int = 0
    texture_filter: int = GL_LINEAR_MIPMAP_NEAREST
    meshes: list = field(default_factory=list)
    cull: bool = True
    cull_face: int = GL_BACK
    depth_func: int = GL_LESS
    post_program: int = PGM_DEFAULT_POST
    fixed_guard_band: bool = False   # guard band factor raised to >= 1 (large targets)
    blend: bool = False              # GL_BLEND (BlendAlpha) for every draw: the in-order tile path

    @property
    def triangles(self) -> int:
        return sum(len(m.indices) // 3 * m.instances for m in self.meshes)

    @property
    def instances(self) -> int:
        return sum(m.instances for m in self.meshes)

    @property
    def transformed_vertices(self) -> int:
        """Vertices referenced by the draws, counted once per instance."""
        return sum((int(m.indices.max()) + 1) * m.instances for m in self.meshes)


def scene_sphere_ground(width: int = 640, height: int = 360, seed: int = 1) -> Scene:
    """UV sphere (32x16, r=1, centre (0,1,0)) above a 2-triangle ground
    quad (+-10 at y=0); 60 degree camera; directional Lambert; 256^2
    seeded value-noise texture with mips, bilinear sampling."""
    sp_pos, sp_nrm, sp_uv, sp_idx = uv_sphere(32, 16, 1.0, (0.0, 1.0, 0.0))
    g_pos = np.array([(-10, 0, -10), (10, 0, -10), (10, 0, 10), (-10, 0, 10)], np.float32)
    g_nrm = np.tile(np.array([[0, 1, 0]], np.float32), (4, 1))
    g_uv = np.array([(0, 0), (4, 0), (4, 4), (0, 4)], np.float32)
    g_idx = np.array([0, 2, 1, 0, 3, 2], np.uint16)

    def mesh(pos, nrm, uv, idx):
        b = {}
        for k in range(3):
            b[k] = pad4(pos[:, k])
            b[3 + k] = pad4(nrm[:, k])
        b[9], b[10] = pad4(uv[:, 0]), pad4(uv[:, 1])
        return Mesh(b, idx)

    L = np.array([-1.0, -2.0, -1.0])
    to_light = -L / np.linalg.norm(L)
    uni = np.zeros(32, np.float32)
    uni[0:3] = to_light
    uni[4], uni[5] = 0.2, 0.8
    tex = make_mipmap_stack(value_noise_rgba8(256, seed))
    return Scene(
        name=f"sphere_ground_{width}x{height}", width=width, height=height,
        view=look_at((0, 2, 6), (0, 0.75, 0), (0, 1, 0)),
        projection=perspective2(60.0, width / height, 0.1, 100.0),
        program=PGM_LAMBERT_TEX, uniforms=uni, texture=tex, texture_dim=256,
        meshes=[mesh(g_pos, g_nrm, g_uv, g_idx), mesh(sp_pos, sp_nrm, sp_uv, sp_idx)])


def scene_clip_floor(width: int = 640, height: int = 360, n: int = 32, seed: int = 9) -> Scene:
    """A tessellated n x n floor (2 n^2 triangles) spanning x in [-20, 20],
    z in [-40, 10] at y = -1 under a camera at the origin looking slightly
    down: the triangles near the camera cross the near plane and the guard
    band with parts on screen, so clipped fans cover many tiles."""
    xs = np.linspace(-20.0, 20.0, n + 1)
    zs = np.linspace(-40.0, 10.0, n + 1)
    X, Z = np.meshgrid(xs, zs)
    pos = np.stack([X.ravel(), np.full(X.size, -1.0), Z.ravel()], 1).astype(np.float32)
    nrm = np.tile(np.array([[0, 1, 0]], np.float32), (len(pos), 1))
    uv = np.stack([(X.ravel() + 20) / 10, (Z.ravel() + 40) / 10], 1).astype(np.float32)
    idx = []
    for j in range(n):
        for i in range(n):
            a = j * (n + 1) + i
            b, c, d = a + 1, a + n + 1, a + n + 2
            idx += [a, c, b, b, c, d]
    b_ = {}
    for k in range(3):
        b_[k] = pad4(pos[:, k])
        b_[3 + k] = pad4(nrm[:, k])
    b_[9], b_[10] = pad4(uv[:, 0]), pad4(uv[:, 1])
    uni = np.zeros(32, np.float32)
    L = np.array([0.3, 1.0, 0.2])
    uni[0:3] = L / np.linalg.norm(L)
    uni[4], uni[5] = 0.25, 0.75
    tex = make_mipmap_stack(value_noise_rgba8(256, seed))
    return Scene(
        name=f"clip_floor{n}_{width}x{height}", width=width, height=height,
        view=look_at((0, 0, 0), (0, -0.35, -1), (0, 1, 0)),
        projection=perspective2(60.0, width / height, 0.1, 100.0),
        program=PGM_LAMBERT_TEX, uniforms=uni, texture=tex, texture_dim=256,
        meshes=[Mesh(b_, np.array(idx, np.uint16))])


def instance_matrices(n: int, seed: int, center, extent: float, smin: float, smax: float) -> np.ndarray:
    rng = np.random.RandomState(seed)
    out = np.empty((n, 16), np.float32)
    for i in range(n):
        r = random_rotation(rng)
        s = rng.uniform(smin, smax)
        p = np.asarray(center) + rng.uniform(-extent / 2, extent / 2, 3)
        m = np.eye(4)
        m[:3, :3] = r * s
        m[:3, 3] = p
        out[i] = m.T.reshape(16)
    return out


def scene_icospheres(width: int = 1920, height: int = 1080, instances: int = 256, seed: int = 2,
                     tex_seed: int = 3, tex_dim: int = 1024, center=(0.0, 0.0, -25.0), extent: float = 40.0,
                     smin: float = 0.2, smax: float = 2.0, program: int = PGM_MANY_BLINN_PHONG) -> Scene:
    """instances x subdivision-3 icosphere (1280 triangles each) with seeded
    rotation, uniform scale 0.2-2.0 and position uniform in a 40-unit cube
    centred 25 units in front of a 60 degree camera; one instanced draw;
    Blinn-Phong with one point light; mipmapped value-noise texture."""
    pos, nrm, uv, idx = icosphere(3)
    b = {}
    for k in range(3):
        b[k] = pad4(pos[:, k])
        b[3 + k] = pad4(nrm[:, k])
    b[9], b[10] = pad4(uv[:, 0]), pad4(uv[:, 1])
    b[15] = instance_matrices(instances, seed, center, extent, smin, smax).reshape(-1)
    uni = np.zeros(32, np.float32)
    uni[0:3] = (0.0, 0.0, 0.0)       # eye
    uni[4:7] = (10.0, 15.0, 0.0)     # point light
    uni[8], uni[9], uni[10] = 0.15, 0.85, 0.6
    tex = make_mipmap_stack(value_noise_rgba8(tex_dim, tex_seed))
    return Scene(
        name=f"icospheres{instances}_{width}x{height}", width=width, height=height,
        view=look_at((0, 0, 0), (0, 0, -1), (0, 1, 0)),
        projection=perspective2(60.0, width / height, 0.1, 100.0),
        program=program, uniforms=uni, texture=tex, texture_dim=tex_dim,
        meshes=[Mesh(b, idx, instances)])


def _mesh_buffers(pos, nrm, uv):
    b = {}
    for k in range(3):
        b[k] = pad4(pos[:, k])
        b[3 + k] = pad4(nrm[:, k])
    b[9], b[10] = pad4(uv[:, 0]), pad4(uv[:, 1])
    return b


_SHARED_TEXTURES = {}


def shared_texture(dim: int, seed: int) -> np.ndarray:
    """One mipmapped value-noise texture per (dim, seed), shared by scenes."""
    key = (dim, seed)
    if key not in _SHARED_TEXTURES:
        _SHARED_TEXTURES[key] = make_mipmap_stack(value_noise_rgba8(dim, seed))
    return _SHARED_TEXTURES[key]


def scene_batch(n: int = 64, seed: int = 4, wmin: int = 128, wmax: int = 1024, hmin: int = 128, hmax: int = 768,
                tmin: int = 500, tmax: int = 50000, tex_dim: int = 256) -> list:
    """Batched workload: n independent scenes.  Framebuffers log-uniform from
    wmin x hmin to wmax x hmax (multiples of 8), triangle counts log-uniform
    in [tmin, tmax] made of seeded icosphere (1280 triangles) and cube (12
    triangles) instances in front of a 60 degree camera; Blinn-Phong with one
    point light; one shared mipmapped texture."""
    rng = np.random.RandomState(seed)
    ipos, inrm, iuv, iidx = icosphere(3)
    cpos, cnrm, cuv, cidx = cube()
    tex = shared_texture(tex_dim, seed)
    out = []
    for k in range(n):
        w = int(round(math.exp(rng.uniform(math.log(wmin), math.log(wmax))) / 8)) * 8
        h = int(round(math.exp(rng.uniform(math.log(hmin), math.log(hmax))) / 8)) * 8
        tris = int(math.exp(rng.uniform(math.log(tmin), math.log(tmax))))
        n_ico = int(tris * rng.uniform(0.3, 0.9)) // 1280
        n_cube = max(1, (tris - n_ico * 1280) // 12)
        meshes = []
        sseed = int(rng.randint(1, 1 << 30))
        depth = float(rng.uniform(8.0, 30.0))
        extent = depth * 0.9
        if n_ico > 0:
            b = _mesh_buffers(ipos, inrm, iuv)
            b[15] = instance_matrices(n_ico, sseed, (0.0, 0.0, -depth), extent, 0.2, 1.5).reshape(-1)
            meshes.append(Mesh(b, iidx, n_ico))
        b = _mesh_buffers(cpos, cnrm, cuv)
        b[15] = instance_matrices(n_cube, sseed + 1, (0.0, 0.0, -depth), extent, 0.1, 0.8).reshape(-1)
        meshes.append(Mesh(b, cidx, n_cube))
        uni = np.zeros(32, np.float32)
        uni[4:7] = (10.0, 15.0, 0.0)
        uni[8], uni[9], uni[10] = 0.15, 0.85, 0.6
        out.append(Scene(
            name=f"batch{seed}_{k}_{w}x{h}", width=w, height=h,
            view=look_at((0, 0, 0), (0, 0, -1), (0, 1, 0)),
            projection=perspective2(60.0, w / h, 0.1, 100.0),
            program=PGM_MANY_BLINN_PHONG, uniforms=uni, texture=tex, texture_dim=tex_dim, meshes=meshes))
    return out


ENV_DIM, ENV_TAPS, GGX_LIGHTS = 16, 8, 16


def env_cube(dim: int, seed: int) -> np.ndarray:
    """6 faces of dim x dim RGBA floats: a sky gradient (bright +y, dark
    -y) times smooth seeded colour noise, faces in +x -x +y -y +z -z order."""
    rng = np.random.RandomState(seed)
    out = np.zeros((6, dim, dim, 4), np.float32)
    t = (np.arange(dim) + 0.5) / dim * 2 - 1
    S, T = np.meshgrid(t, t)
    for f in range(6):
        # direction of each texel (inverse of the face selection in env_sample)
        if f == 0: d = np.stack([np.ones_like(S), -T, -S], -1)
        elif f == 1: d = np.stack([-np.ones_like(S), -T, S], -1)
        elif f == 2: d = np.stack([S, np.ones_like(S), T], -1)
        elif f == 3: d = np.stack([S, -np.ones_like(S), -T], -1)
        elif f == 4: d = np.stack([S, -T, np.ones_like(S)], -1)
        else: d = np.stack([-S, -T, -np.ones_like(S)], -1)
        d = d / np.linalg.norm(d, axis=-1, keepdims=True)
        sky = 0.35 + 0.65 * np.clip(d[..., 1] * 0.5 + 0.5, 0, 1)
        lat = rng.uniform(0.6, 1.0, (4, 4, 3))
        xi = np.clip(((S + 1) / 2 * 3), 0, 2.999)
        yi = np.clip(((T + 1) / 2 * 3), 0, 2.999)
        i0, j0 = xi.astype(int), yi.astype(int)
        fx, fy = xi - i0, yi - j0
        c = (lat[j0, i0] * ((1 - fx) * (1 - fy))[..., None] + lat[j0, i0 + 1] * (fx * (1 - fy))[..., None]
             + lat[j0 + 1, i0] * ((1 - fx) * fy)[..., None] + lat[j0 + 1, i0 + 1] * (fx * fy)[..., None])
        out[f, ..., :3] = c * sky[..., None]
        out[f, ..., 3] = 1.0
    return out.reshape(-1)


def heavy_tables(light_seed: int = 5, env_seed: int = 6):
    """The heavy shader's tables: environment cube (buffer 11), taps (12),
    lights (13), and its uniforms with the eye at the origin."""
    rng = np.random.RandomState(light_seed)
    lights = np.zeros((GGX_LIGHTS, 8), np.float32)
    # on a shell around the instance volume: directions uniform, 35-55 units
    # from its centre
    d = rng.normal(size=(GGX_LIGHTS, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    lights[:, 0:3] = np.array([0.0, 0.0, -25.0]) + d * rng.uniform(35.0, 55.0, (GGX_LIGHTS, 1))
    lights[:, 4:7] = rng.uniform(0.3, 1.0, (GGX_LIGHTS, 3)) * 400.0
    rng = np.random.RandomState(env_seed + 1000)
    taps = np.zeros((ENV_TAPS, 4), np.float32)
    taps[:, 0:3] = rng.normal(0.0, 0.6, (ENV_TAPS, 3))
    taps[:, 3] = 1.0 / ENV_TAPS
    uni = np.zeros(32, np.float32)
    uni[0:3] = (0.0, 0.0, 0.0)   # eye
    uni[3] = 0.03                # ambient
    uni[4] = 0.8                 # environment specular
    uni[5] = 0.25                # environment diffuse
    return env_cube(ENV_DIM, env_seed), taps.reshape(-1), lights.reshape(-1), uni


def scene_heavy(width: int = 1920, height: int = 1080, instances: int = 256, seed: int = 2,
                light_seed: int = 5, env_seed: int = 6, rough_seed: int = 7, tex_seed: int = 3,
                tex_dim: int = 1024) -> Scene:
    """The config-2 geometry under the heavy shader: 16 seeded point lights,
    GGX specular with per-instance roughness uniform in [0.05, 1], and an
    8-tap seeded environment lookup (program GGXEnv, id 82)."""
    sc = scene_icospheres(width, height, instances, seed, tex_seed=tex_seed, tex_dim=tex_dim, program=PGM_GGX_ENV)
    b = sc.meshes[0].buffers
    b[14] = np.random.RandomState(rough_seed).uniform(0.05, 1.0, instances).astype(np.float32)
    b[11], b[12], b[13], sc.uniforms = heavy_tables(light_seed, env_seed)
    sc.name = f"heavy_icospheres{instances}_{width}x{height}"
    return sc


STRESS_CENTER = (0.0, 0.0, -30.0)


def scene_stress(view: int = 0, width: int = 3840, height: int = 2160, instances: int = 6250, seed: int = 8,
                 cluster_radius: float = 6.0, smin: float = 0.06, smax: float = 0.46, tex_dim: int = 1024,
                 fixed_guard_band: bool = True) -> Scene:
    """Stress workload, one of four camera views: `instances` subdivision-2
    icospheres (320 triangles each; 6250 -> 2,000,000 triangles) with scale
    log-uniform in [smin, smax] (projected triangle areas about 0.25-16 px^2
    at 3840x2160 from 30 units) and positions uniform in a ball of
    `cluster_radius` around the centre, so that the covered region has a
    depth complexity of several layers; the heavy GGX shader of the learned
    shading workload.  The four views look at the centre from azimuths
    0/90/180/270 degrees (+15 degrees) at 30 units."""
    pos, nrm, uv, idx = icosphere(2)
    rng = np.random.RandomState(seed)
    # positions uniform in the ball, log-uniform scales, uniform random rotations
    d = rng.normal(size=(instances, 3))
    d *= (cluster_radius * rng.uniform(0, 1, (instances, 1)) ** (1.0 / 3.0)) / np.linalg.norm(d, axis=1, keepdims=True)
    scl = np.exp(rng.uniform(math.log(smin), math.log(smax), instances))
    q = rng.normal(size=(instances, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
    R = np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w),
                  2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w),
                  2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)], 1).reshape(instances, 3, 3)
    M = np.zeros((instances, 4, 4))
    M[:, :3, :3] = R * scl[:, None, None]
    M[:, :3, 3] = np.asarray(STRESS_CENTER) + d
    M[:, 3, 3] = 1.0
    mats = np.ascontiguousarray(M.transpose(0, 2, 1)).reshape(instances, 16).astype(np.float32)
    b = _mesh_buffers(pos, nrm, uv)
    b[15] = mats.reshape(-1)
    b[11], b[12], b[13], uni = heavy_tables()
    b[14] = np.random.RandomState(seed + 1).uniform(0.05, 1.0, instances).astype(np.float32)
    ang = math.radians(90.0 * view + 15.0)
    c = np.asarray(STRESS_CENTER)
    eye = c + 30.0 * np.array([math.sin(ang), 0.25, math.cos(ang)])
    uni[0:3] = eye
    return Scene(
        name=f"stress{instances}_view{view}_{width}x{height}", width=width, height=height,
        view=look_at(tuple(eye), tuple(c), (0, 1, 0)),
        projection=perspective2(60.0, width / height, 0.1, 200.0),
        program=PGM_GGX_ENV, uniforms=uni, texture=shared_texture(tex_dim, 3), texture_dim=tex_dim,
        meshes=[Mesh(b, idx, instances)], fixed_guard_band=fixed_guard_band)


def fragment_state_key(scene: Scene) -> int:
    """GLState::FragmentStateKey of the scene's draw state without the
    program id (rglv_gl.hxx:136-146): depth test on, blending as the scene
    sets it, depth and colour writes on, RB_RGBF32 colour, RB_F32 depth."""
    return ((1 << 1) | (scene.depth_func << 2) | (int(scene.blend) << 4) | (1 << 5) | (1 << 6) | (RB_RGBF32 << 7)
            | (RB_F32 << 9))


def render_scene(gpu, scene: Scene, color_out, depth_out, prim_out=None, buffers=None) -> None:
    """Issue the GL calls that draw *scene* into the given targets.

    *gpu* is any object with the rglv::GPU / rglv::GL method set (reset,
    install, gl_*); *buffers* maps id(array) -> the pointer the backend
    expects for that array (device pointer for the HIP pipeline)."""
    ptr = (lambda a: buffers[id(a)]) if buffers is not None else (lambda a: a)
    gpu.set_guard_band(scene.fixed_guard_band)
    gpu.reset(scene.width, scene.height, *scene.tile_blocks)
    gpu.install(scene.program, 0, 0)
    gpu.install(scene.program, fragment_state_key(scene), 1)
    gpu.install(scene.post_program, 0, 2)
    gpu.gl_renderbuffer_type(GL_COLOR_ATTACHMENT0, RB_RGBF32)
    gpu.gl_renderbuffer_type(GL_DEPTH_ATTACHMENT, RB_F32)
    gpu.gl_clear_color(*scene.clear_color)
    gpu.gl_clear_depth(1.0)
    gpu.gl_clear(GL_COLOR_BUFFER_BIT | GL_DEPTH_BUFFER_BIT)
    gpu.gl_use_program(scene.program)
    gpu.gl_use_uniforms(gpu.gl_alloc_uniform_buffer(scene.uniforms))
    gpu.gl_view_matrix(scene.view)
    gpu.gl_projection_matrix(scene.projection)
    if scene.cull:
        gpu.gl_enable(GL_CULL_FACE)
        gpu.gl_cull_face(scene.cull_face)
    gpu.gl_depth_func(scene.depth_func)
    if scene.blend:
        gpu.gl_enable(GL_BLEND)
    if scene.texture is not None:
        gpu.gl_bind_texture(0, ptr(scene.texture), scene.texture_dim, scene.texture_dim,
                            scene.texture_dim, scene.texture_filter)
    for m in scene.meshes:
        for slot in range(16):
            if slot in m.buffers:
                gpu.gl_use_buffer(slot, ptr(m.buffers[slot]), len(m.buffers[slot]))
            else:
                gpu.gl_use_buffer(slot, None, 0)
        if m.instances > 1:
            gpu.gl_draw_elements_instanced(GL_TRIANGLES, len(m.indices), GL_UNSIGNED_SHORT,
                                           ptr(m.indices), m.instances)
        else:
            gpu.gl_draw_elements(GL_TRIANGLES, len(m.indices), GL_UNSIGNED_SHORT, ptr(m.indices),
                                 RGL_HINT_DENSE | RGL_HINT_READ4)
    gpu.gl_use_program(scene.post_program)
    gpu.gl_store_color_tc(color_out, True)
    gpu.gl_store_depth(depth_out)
    if prim_out is not None:
        gpu.gl_store_primitive_id(prim_out)


def scene_arrays(scene: Scene) -> list:
    """Every array the scene's GL calls reference (to upload once)."""
    arrs = []
    if scene.texture is not None:
        arrs.append(scene.texture)
    for m in scene.meshes:
        arrs.extend(m.buffers.values())
        arrs.append(m.indices)
    return arrs


# ---------------------------------------------------------------------------
# scenes given as GL call lists (the form the viewer's scene-graph nodes
# issue: several draws with state changes between them, clears and stores
# between draws, blending, scissor, depth and colour write masks)
# ---------------------------------------------------------------------------
@dataclass
class CallScene:
    """A frame as a list of rglv::GL calls.  calls: (method, *args) with
    method the GL member name in snake case ("enable", "draw_elements", ...);
    an argument "@name" is the array arrays[name] (a device pointer on the
    GPU), ("out", k) is output buffer k.  outputs: the kinds of the store
    targets (output_array: "color", "depth", "prim", "rgba", "rgba_half",
    "quads")."""
    name: str
    width: int
    height: int
    tile_blocks: tuple = (8, 8)
    arrays: dict = field(default_factory=dict)
    calls: list = field(default_factory=list)
    outputs: list = field(default_factory=list)
    fixed_guard_band: bool = False


def output_array(kind: str, width: int, height: int) -> np.ndarray:
    """A store target of a CallScene output, filled with a pattern every
    store must overwrite: "color" u32 0x00RRGGBB, "depth" f32, "prim" u32
    (H x W); "rgba" (H x W x 4 f32, StoreColor(FloatingPointCanvas*,
    false)); "rgba_half" (ceil(H/2) x ceil(W/2) x 4, the downsampled
    store); "quads" (ceil(H/2) x ceil(W/2) x 16, StoreColor(QFloat4Canvas*))."""
    hh, hw = (height + 1) // 2, (width + 1) // 2
    if kind == "depth":
        return np.full((height, width), np.nan, np.float32)
    if kind == "rgba":
        return np.full((height, width, 4), np.nan, np.float32)
    if kind == "rgba_half":
        return np.full((hh, hw, 4), np.nan, np.float32)
    if kind == "quads":
        return np.full((hh, hw, 16), np.nan, np.float32)
    return np.full((height, width), 0x5A5A5A5A, np.uint32)


def fragment_state_key_of(st: dict) -> int:
    """GLState::FragmentStateKey without the program id (rglv_gl.hxx:136-146)."""
    key = int(st["scissor"])
    key |= int(st["depth_test"]) << 1
    if st["depth_test"]:
        key |= st["depth_func"] << 2
    key |= int(st["blend"]) << 4
    key |= int(st["depth_write"]) << 5
    key |= int(st["color_write"]) << 6
    key |= st["color_type"] << 7
    key |= st["depth_type"] << 9
    return key


def call_installs(sc: CallScene) -> list:
    """The Install entries the scene's draws and stores need: (program,
    state key, kind) with kind 0 bin, 1 draw, 2 blt -- what the viewer's
    shaders.cxx installs up front for the states it uses."""
    st = dict(scissor=False, depth_test=True, depth_func=GL_LESS, blend=False, depth_write=True, color_write=True,
              color_type=0, depth_type=0, program=0)
    out = []
    for c in sc.calls:
        m, a = c[0], c[1:]
        if m in ("enable", "disable"):
            name = {GL_SCISSOR_TEST: "scissor", GL_BLEND: "blend", GL_DEPTH_TEST: "depth_test"}.get(a[0])
            if name:
                st[name] = m == "enable"
        elif m == "depth_func":
            st["depth_func"] = a[0]
        elif m == "depth_write_mask":
            st["depth_write"] = bool(a[0])
        elif m == "color_write_mask":
            st["color_write"] = bool(a[0])
        elif m == "renderbuffer_type":
            st["color_type" if a[0] == GL_COLOR_ATTACHMENT0 else "depth_type"] = a[1]
        elif m == "use_program":
            st["program"] = a[0]
        elif m.startswith("draw_"):
            out += [(st["program"], 0, 0), (st["program"], fragment_state_key_of(st), 1)]
        elif m == "store_color_tc":
            out.append((st["program"], 0, 2))
    return sorted(set(out))


def render_calls(gpu, sc: CallScene, outs: list, ptr=lambda a: a) -> None:
    """Issue the scene's calls on *gpu* (any backend with the rglv::GPU /
    rglv::GL method names); outs[k] is the backend's handle of output k,
    ptr maps an array to the backend's handle of it."""
    gpu.set_guard_band(sc.fixed_guard_band)
    gpu.reset(sc.width, sc.height, *sc.tile_blocks)
    for pid, key, kind in call_installs(sc):
        gpu.install(pid, key, kind)

    def arg(a):
        if isinstance(a, str) and a.startswith("@"):
            return ptr(sc.arrays[a[1:]])
        if isinstance(a, tuple) and len(a) == 2 and a[0] == "out":
            return outs[a[1]]
        return a

    for c in sc.calls:
        m, a = c[0], [arg(x) for x in c[1:]]
        if m == "use_buffer":
            arr = sc.arrays[c[2][1:]] if isinstance(c[2], str) else None
            gpu.gl_use_buffer(a[0], a[1], len(arr) if arr is not None else 0)
        elif m == "alloc_uniforms":
            gpu.gl_use_uniforms(gpu.gl_alloc_uniform_buffer(sc.arrays[c[1][1:]]))
        else:
            getattr(gpu, "gl_" + m)(*a)


def mesh_calls(mesh_name: str, mesh: Mesh, instanced_arrays: bool = False) -> list:
    """use_buffer calls for a mesh's slots and its draw, naming its arrays
    "<mesh_name>.<slot>" and "<mesh_name>.idx" (see mesh_arrays)."""
    calls = [("use_buffer", s, f"@{mesh_name}.{s}" if s in mesh.buffers else None) for s in range(16)]
    if mesh.instances > 1:
        calls.append(("draw_elements_instanced", GL_TRIANGLES, len(mesh.indices), GL_UNSIGNED_SHORT,
                      f"@{mesh_name}.idx", mesh.instances))
    else:
        calls.append(("draw_elements", GL_TRIANGLES, len(mesh.indices), GL_UNSIGNED_SHORT, f"@{mesh_name}.idx",
                      RGL_HINT_DENSE | RGL_HINT_READ4))
    return calls


def mesh_arrays(mesh_name: str, mesh: Mesh) -> dict:
    out = {f"{mesh_name}.{s}": a for s, a in mesh.buffers.items()}
    out[f"{mesh_name}.idx"] = mesh.indices
    return out
