"""The serving loop on the GPU: requests for two resident scenes at mixed
resolutions, submitted together, with staggered arrivals and under a budget
that forces them to wait, give bitwise the same images as each request
rendered alone (same kernels on the same inputs; tolerance 0)."""
import numpy as np
import pytest
import torch

from nsvf_amd import scene as S
from nsvf_amd.render import Renderer
from nsvf_amd.serve import RenderServer, latency_summary

pytestmark = pytest.mark.gpu


def requests(n, seed, device):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        scene = ["a", "b"][int(rng.integers(2))]
        res = int(rng.choice([24, 40, 64]))
        radius = float(rng.uniform(2.0, 3.0))
        rs, rd = S.camera_rays(res, res, S.sphere_camera(radius, 1000 * seed + i))
        out.append((scene, torch.from_numpy(rs).to(device), torch.from_numpy(rd).to(device)))
    return out


@pytest.fixture(scope="module")
def server(device):
    srv = RenderServer(device, slots=3, precision="bf16", hierarchical=True)
    srv.load_scene("a", S.shell_scene(res=32))
    srv.load_scene("b", S.shell_scene(res=40, jitter_seed=1, emb_seed=5, mlp_seed=6))
    return srv


def alone(srv, scene, rs, rd):
    r = Renderer(srv.device, precision="bf16", hierarchical=True, compat=srv.compat)
    rgb, depth, missed = r.render(srv.scenes[scene], rs, rd)
    torch.cuda.synchronize()
    return rgb.clone(), depth.clone(), missed.clone()


def check_same(srv, reqs, items):
    for req, (scene, rs, rd) in zip(reqs, items):
        assert req.state == "done", (req.rid, req.state, req.error)
        ref = alone(srv, scene, rs, rd)
        for got, want in zip(req.result, ref):
            assert torch.equal(got, want), f"request {req.rid} differs from its render alone"


def test_mixed_batch_matches_alone(server):
    """Submitted together, requests of one scene run as batches of
    concatenated rays (mixed resolutions: the traversal over row strips);
    each request's pixels equal its render alone."""
    items = requests(9, 1, server.device)
    reqs = [server.submit(*it) for it in items]
    server.drain()
    torch.cuda.synchronize()
    check_same(server, reqs, items)
    assert max(r.batch for r in reqs) > 1
    scenes = sum(server._scene_bytes.values())
    assert server.budget.used == scenes + server.scheduler.held_bytes()
    server.release_idle()
    assert server.budget.used == scenes


def test_staggered_arrivals_match_alone(server):
    items = requests(8, 2, server.device)
    arrivals = [(0.002 * i, *it) for i, it in enumerate(items)]
    reqs = server.run_arrivals(arrivals)
    torch.cuda.synchronize()
    check_same(server, sorted(reqs, key=lambda r: r.rid), items)
    s = latency_summary(reqs)
    assert s["done"] == 8 and s["p99_ms"] >= s["p50_ms"] > 0


def test_tight_budget_queues_and_matches_alone(device):
    """Six identical 64x64 requests under a budget of the scenes plus 1.5
    times one request's workspace: several geometries fit at once, but only one
    request can shade at a time; every result still equals its render alone."""
    srv = RenderServer(device, slots=4, budget_bytes=1 << 40, precision="bf16")
    srv.load_scene("a", S.shell_scene(res=32))
    srv.load_scene("b", S.shell_scene(res=40, jitter_seed=1, emb_seed=5, mlp_seed=6))
    rs, rd = S.camera_rays(64, 64, S.sphere_camera(2.5, 50))
    items = [("a", torch.from_numpy(rs).to(device), torch.from_numpy(rd).to(device))
             for _ in range(6)]

    def need(scene, rs, rd):
        probe = Renderer(device, precision="bf16")
        g = probe.geometry_bytes(srv.scenes[scene], rs.shape[0])
        probe.begin(srv.scenes[scene], rs, rd)
        return g + probe.shading_bytes()

    needs = [need(*it) for it in items]
    assert len(set(needs)) == 1
    srv.budget.total = srv.budget.used + int(1.5 * needs[0])
    reqs = [srv.submit(*it) for it in items]
    shading = 0
    while not srv.scheduler.idle():
        srv.poll()
        assert srv.budget.used <= srv.budget.total
        shading = max(shading, sum(st.phase == "shading" for st in srv.scheduler.slots))
    torch.cuda.synchronize()
    assert shading == 1, shading
    check_same(srv, reqs, items)


@pytest.mark.parametrize("precision", ["bf16x3", "fp32"])
def test_segmented_batch_equals_views_alone(device, precision):
    """Four views of one scene rendered as one batch of concatenated rays
    with their segments: each view's image equals its render alone bit for
    bit (each segment's fine pass normalises over its own longest coarse
    row, as the view does alone)."""
    sc = S.shell_scene(res=32)
    from nsvf_amd.render import DeviceScene
    dsc = DeviceScene(sc, device)
    items = requests(4, 3, device)
    rs = torch.cat([it[1] for it in items])
    rd = torch.cat([it[2] for it in items])
    segs = [it[1].shape[0] for it in items]
    r = Renderer(device, precision=precision, hierarchical=True, compat=False)
    r.begin(dsc, rs, rd, segments=segs)
    batch = [t.clone() for t in r.finish()]
    torch.cuda.synchronize()
    a = 0
    for (scene, rs_i, rd_i), n in zip(items, segs):
        alone_r = Renderer(device, precision=precision, hierarchical=True, compat=False)
        ref = alone_r.render(dsc, rs_i, rd_i)
        torch.cuda.synchronize()
        for got, want in zip(batch, ref):
            assert torch.equal(got[a:a + n], want)
        a += n


def test_serving_sized_scenes_mix_matches_alone(device):
    """The serving configuration: four resident shells of about 20k, 100k,
    250k and 400k voxels at the headline precision, twelve requests at 256^2,
    400^2, 512^2 and 800^2 over all four scenes, half submitted together and
    half arriving while those run (same-scene requests share launches); every
    image equals its render alone bit for bit."""
    scenes = {f"shell{r}": S.shell_scene(res=r, emb_seed=e, mlp_seed=m)
              for r, e, m in [(46, 11, 21), (79, 12, 22), (107, 13, 23), (125, 14, 24)]}
    srv = RenderServer(device, slots=4, precision="bf16x3", hierarchical=True)
    for name, sc in scenes.items():
        srv.load_scene(name, sc)
    names = list(scenes)
    rng = np.random.default_rng(17)
    items = []
    for i in range(12):
        name = names[i % 4] if i < 8 else names[int(rng.integers(4))]
        res = [256, 400, 512, 800][int(rng.integers(4))]
        cam = S.sphere_camera(float(rng.uniform(2.0, 3.0)), 500 + i)
        rs, rd = S.camera_rays(res, res, cam)
        items.append((name, torch.from_numpy(rs).to(device), torch.from_numpy(rd).to(device), res))
    print("requests:", [(n, w) for n, _, _, w in items], flush=True)
    reqs = [srv.submit(n, rs, rd, width=w) for n, rs, rd, w in items[:6]]
    for _ in range(3):
        srv.poll()
    reqs += [srv.submit(n, rs, rd, width=w) for n, rs, rd, w in items[6:]]
    srv.drain()
    torch.cuda.synchronize()
    assert max(r.batch for r in reqs) > 1
    print("batches of several widths traced in pixel tiles:",
          sum(st.slot.tiled_batches for st in srv.scheduler.slots), flush=True)
    for req, (name, rs, rd, w) in zip(reqs, items):
        assert req.state == "done", (req.rid, req.state, req.error)
        r = Renderer(device, precision="bf16x3", hierarchical=True, compat=srv.compat)
        ref = r.render(srv.scenes[name], rs, rd, width=w)
        torch.cuda.synchronize()
        for got, want in zip(req.result, ref):
            assert torch.equal(got, want), f"request {req.rid} ({name}, {w}^2) differs from alone"
        del r
