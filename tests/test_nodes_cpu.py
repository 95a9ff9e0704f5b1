"""The viewer's node call sequences (rsr_amd.nodes) on the oracle: the
frames they issue are well formed and draw what the nodes describe."""
import numpy as np

import viewer_graphs as V
from oracle_backend import render_calls
from rsr_amd import nodes, scenes


def _run(frames):
    return nodes.run_frames(frames, lambda sc: render_calls(sc, threads=4)[0])


def test_viewer_frame_order_and_outputs():
    fr = V.graph()
    # one light: a shadow frame per Render pass (gllayer.cxx:151-181 runs in
    # both), then the main frame
    assert [sc.name.split("_")[0] for sc in fr] == ["shadow0", "shadow1", "viewer"]
    main = fr[-1]
    first = [c[0] for c in main.calls[:5]]
    assert first == ["renderbuffer_type", "renderbuffer_type", "clear_color", "clear_depth", "clear"]
    assert main.calls[-3][0] == "store_color_tc"
    outs = _run(fr)
    smap = outs[1][0]
    assert (smap < 1.0).sum() > 100            # the sphere's back faces in the light's depth
    color, depth, prim = outs[-1]
    drawn = prim != 0xFFFFFFFF
    assert drawn.mean() > 0.5 and len(np.unique(color[drawn])) > 1000


def test_shadow_darkens_the_ground():
    # OBJ2SProgram with the stored map: ground pixels behind the sphere from
    # the light are attenuated to 0.111 of the lit value
    lit = _run(V.graph(many=False))[-1][0]
    r = (lit >> 16) & 255
    dark = (r < 40).sum()
    assert dark > 200
