"""Calls on different streams with their own workspaces are independent:
four streams (default priority) extract four different volumes, eight calls
each, launched interleaved so that their kernels overlap on the device, and
every output equals the oracle's.  A fifth stream integrates frames into a
volume meanwhile; its result equals the oracle's too."""
import numpy as np
import pytest
import torch

from marching_cube_amd import fusion, mesh, synth
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def test_concurrent_streams_with_own_workspaces(dev):
    print("[case] 4 extraction streams x 8 calls + 1 integration stream", flush=True)
    vols = [synth.sphere_volume(200, radius=0.31)[0], synth.sphere_volume(160, radius=0.27)[0],
            synth.noise_volume((40, 70, 90), seed=3),
            synth.noise_volume((24, 200, 33), seed=4, p_zero_w=0.05)]
    hs = [1.0 / 200, 1.0 / 160, 0.01, 0.02]
    refs = [O.extract(v, h, want=("vv",), nthreads=8) for v, h in zip(vols, hs)]
    streams = [torch.cuda.Stream(device=dev) for _ in vols]
    dvols = [torch.from_numpy(v.view(np.int32)).to(dev) for v in vols]
    wss = [mesh.Workspace(dev) for _ in vols]
    outs = []
    for ref in refs:
        outs.append([torch.full((ref["n"], 7), -1, dtype=torch.int32, device=dev)
                     for _ in range(8)])
    counts = [torch.zeros(8, dtype=torch.int64, device=dev) for _ in vols]
    # integration on its own stream at the same time
    intr = [525.0, 525.0, 319.5, 239.5]
    objs = synth.kinfu_scene(8, 1.0, seed=9)
    poses = synth.orbit_poses(3, 1.0, radius=1.6, seed=9)
    depth = fusion.synth_depth(intr, poses, objs, 320, 240, seed=9, device=dev)
    vox = fusion.voxel_new((96, 96, 96), 0.03, 1.0 / 96, device=dev)
    torch.cuda.synchronize()
    s_int = torch.cuda.Stream(device=dev)
    with torch.cuda.stream(s_int):
        fusion.integrate(vox, intr, poses, depth, trunc=0.03, max_w=64)
    for it in range(8):
        for k, s in enumerate(streams):
            with torch.cuda.stream(s):
                mesh.extract_into(dvols[k], hs[k], vv=outs[k][it], cap=refs[k]["n"],
                                  d_count=counts[k][it:it + 1], workspace=wss[k], sync=False)
    torch.cuda.synchronize()
    for k, ref in enumerate(refs):
        wss[k].check()
        want = ref["vv"].view(np.int32).reshape(-1, 7).tobytes()
        assert (counts[k].cpu().numpy() == ref["n"]).all(), k
        for it in range(8):
            assert outs[k][it].cpu().numpy().tobytes() == want, (k, it)
    ref_vol = O.volume_init((96, 96, 96))
    O.integrate(ref_vol, 1.0 / 96, 0.03, 64, intr, poses, depth.cpu().numpy())
    assert np.array_equal(vox.data.cpu().numpy().view(np.uint32), ref_vol)
