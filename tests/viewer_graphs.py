"""Scene graphs of the viewer's rendering nodes ($gpu, $layer, $mesh,
$many, $truecolor) for the caller tests: rsr_amd.nodes turns each into the
GL call lists the reference's nodes issue."""
from __future__ import annotations

import numpy as np

from rsr_amd import nodes, scenes


def _ground(n=8, size=12.0, seed=1):
    xs = np.linspace(-size, size, n + 1)
    X, Z = np.meshgrid(xs, xs)
    pos = np.stack([X.ravel(), np.zeros(X.size), Z.ravel()], 1)
    nrm = np.tile([0.0, 1.0, 0.0], (len(pos), 1))
    kd = np.random.default_rng(seed).uniform(0.08, 0.25, (len(pos), 3))
    idx = []
    for j in range(n):
        for i in range(n):
            a = j * (n + 1) + i
            idx += [a, a + n + 1, a + 1, a + 1, a + n + 1, a + n + 2]
    return pos, nrm, kd, idx


def graph(post=scenes.PGM_DEFAULT_POST, uf0=None, size=(640, 360), tile_blocks=(8, 8), shadows=True, many=True):
    gpos, gnrm, gkd, gidx = _ground()
    ground = nodes.MeshNode("ground", gpos, gnrm, gkd, gidx, make_shadows=True, take_shadows=shadows)
    spos, snrm, suv, sidx = scenes.uv_sphere(24, 12, 1.0, (0.0, 1.2, 0.0))
    skd = np.random.default_rng(2).uniform(0.1, 0.35, (len(spos), 3))
    sphere = nodes.MeshNode("sphere", spos, snrm, skd, sidx, make_shadows=True, take_shadows=shadows)
    gls = [ground, sphere]
    if many:
        ipos, inrm, iuv, iidx = scenes.icosphere(1)
        b = scenes._mesh_buffers(ipos, inrm, iuv)
        mats = scenes.instance_matrices(24, 7, (0.0, 1.0, -2.0), 7.0, 0.3, 0.7)
        gls.append(nodes.ManyNode("many", ipos, inrm, iuv, iidx, mats))
    cam = nodes.Camera(eye=(0.0, 3.5, 9.0), at=(0.0, 0.8, 0.0))
    lights = [nodes.SpotLight(pos=(18.0, 30.0, 22.0), at=(0.0, 0.0, 0.0), angle=40.0, size=256)] if shadows else []
    layer = nodes.LayerNode((0.15, 0.12, 0.2), cam, gls, lights)
    gpu = nodes.GpuNode(size, tile_blocks, [layer])
    return nodes.build_frames(gpu, nodes.TrueColorNode(post, uf0, True))
