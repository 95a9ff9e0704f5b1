"""GPU parity on scene variants that exercise the paths the two headline
scenes leave quiet: near-plane and guard-band clipping with visible
clipped fans, LEQUAL with culling off (back faces drawn and reordered),
front-face culling, 32-pixel tile grids (one of them large enough for
the three-launch scan), and the reference's own Amy and Many programs.  Every buffer is compared bit-exactly with the oracle."""
import dataclasses

import numpy as np
import pytest
import torch  # noqa: F401

from rsr_amd import scenes
from test_parity_gpu import _compare, _gpu_render

pytestmark = pytest.mark.gpu


def _check(s, min_clipped_pairs=0):
    from oracle_backend import render as oracle_render
    ref = oracle_render(s, threads=8)
    got = _gpu_render(s)
    _compare(s.name, ref, got)
    assert ref[3]["clipped_tile_pairs"] >= min_clipped_pairs
    assert got[3]["clipped_tile_pairs"] == ref[3]["clipped_tile_pairs"]


def test_clip_floor():
    # a floor under and behind the camera: near-plane and guard-band clipping
    # with visible fans
    _check(scenes.scene_clip_floor(), min_clipped_pairs=20)
    _check(scenes.scene_clip_floor(1920, 1080, 64), min_clipped_pairs=200)


def test_instances_around_camera():
    # instances around the camera: triangles crossing the near plane off screen
    s = scenes.scene_icospheres(640, 360, instances=96, seed=11, center=(0.0, 0.0, -4.0), extent=12.0,
                                smin=0.3, smax=1.5, tex_dim=256)
    _check(s)


def test_leq_no_cull():
    s = dataclasses.replace(scenes.scene_sphere_ground(), cull=False, depth_func=scenes.GL_LEQUAL,
                            name="sphere_ground_leq_nocull")
    _check(s)


def test_front_face_cull():
    s = dataclasses.replace(scenes.scene_sphere_ground(), cull_face=scenes.GL_FRONT, name="sphere_ground_cullfront")
    _check(s)


def test_tile_32():
    s = dataclasses.replace(scenes.scene_icospheres(960, 540, instances=96, seed=5, tex_dim=256),
                            tile_blocks=(4, 4), name="icospheres96_960x540_tile32")
    _check(s)


def test_tile_32_1080p_split_scan():
    # 2040 tiles in one frame: the scan takes its three-launch form (local
    # block scans, block bases, final offsets and work items)
    s = dataclasses.replace(scenes.scene_icospheres(1920, 1080, instances=128, seed=9, tex_dim=256),
                            tile_blocks=(4, 4), name="icospheres128_1920x1080_tile32")
    _check(s)


def test_reference_programs():
    s1 = dataclasses.replace(scenes.scene_sphere_ground(), program=scenes.PGM_AMY, name="sphere_ground_amy")
    _check(s1)
    s2 = scenes.scene_icospheres(640, 360, instances=48, seed=7, tex_dim=64, program=scenes.PGM_MANY)
    s2.uniforms[0] = 0.25
    _check(s2)


def test_heavy_shader_1080p():
    # the GGX + environment heavy shader on the 1080p icosphere geometry
    # (the learned shader's reference): colour bit-exact like the plain path
    _check(scenes.scene_heavy())


def test_heavy_shader_clipped():
    # heavy shader with near-plane clipping: per-instance roughness and the
    # environment taps carried through clipped fans
    s = scenes.scene_heavy(640, 360, instances=96, seed=11)
    s.meshes[0].buffers[15] = scenes.instance_matrices(96, 11, (0.0, 0.0, -4.0), 12.0, 0.3, 1.5).reshape(-1)
    _check(s)
