"""The viewer's node call sequences (rsr_amd.nodes: $gpu, $layer with a
spot light and its shadow-map frames, $mesh with OBJ2S, $many, $truecolor)
through the C ABI on the GPU: every frame's outputs equal the oracle's,
with each shadow map fed to the next frame as the reference feeds it."""
import numpy as np
import pytest
import torch  # noqa: F401

import viewer_graphs as V
from rsr_amd import nodes, scenes

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("post,uf0,tb", [(scenes.PGM_DEFAULT_POST, None, (8, 8)), (scenes.PGM_IQ_POST, None, (8, 8)),
                                          (scenes.PGM_EXPOSURE_POST, 0.7, (4, 4))])
def test_viewer_frames_match_oracle(post, uf0, tb):
    from oracle_backend import render_calls as oracle_calls
    from rsr_amd import rgl
    ref = nodes.run_frames(V.graph(post, uf0, tile_blocks=tb), lambda sc: oracle_calls(sc, threads=8)[0])

    def gpu_draw(sc):
        print(f"[{sc.name}]", flush=True)
        return rgl.render_calls(sc)[0]
    got = nodes.run_frames(V.graph(post, uf0, tile_blocks=tb), gpu_draw)
    for k, (r, g) in enumerate(zip(ref, got)):
        for j, (a, b) in enumerate(zip(r, g)):
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), (k, j)
