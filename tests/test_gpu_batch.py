"""Batched views on MI355X: one persistent launch per pass must give each
view the image its own per-view ``encode`` gives (same trace code, same RNG
streams, same blend order)."""
import numpy as np
import pytest

from stargaze_amd import scenes
from stargaze_amd.pathtracer import INTEGRATOR_FUSED, Batch, Options, PathTracer

pytestmark = pytest.mark.gpu


def small_views(count=8, seed=11, halo_every=0):
    """Views without a landscape: batched passes over landscape views are
    switched off (test_batch_with_landscape_views)."""
    return scenes.batch_views(count, seed=seed, min_pixels=16 * 16, max_pixels=96 * 64, spp=(1, 3, 4, 16),
                              halo_every=halo_every)


def tracers(views, limit=None):
    out = []
    for v in views:
        o = Options(samples_per_frame=v["spp"], max_bounces=v["bounces"], milky_way=0.0)
        if limit is not None:
            o.sample_limit = limit(v)
        t = PathTracer((v["width"], v["height"]), options=o)
        if v["terrain"] is not None:
            t.upload_terrain(*v["terrain"])
        t.set_integrator(INTEGRATOR_FUSED)
        out.append(t)
    return out


def compare(views, a, b, label):
    worst = 0.0
    for i, (v, ta, tb) in enumerate(zip(views, a, b)):
        assert ta.samples() == tb.samples(), f"{label}: view {i} sample count"
        x, y = ta.read_accumulation(), tb.read_accumulation()
        bx, by = ta.read_backdrop(), tb.read_backdrop()
        assert np.all(np.isfinite(x)), f"{label}: view {i} non-finite"
        scale = max(1e-30, float(np.abs(y).max()))
        diff = max(float(np.abs(x - y).max()), float(np.abs(bx - by).max())) / scale
        worst = max(worst, diff)
        # Same code and RNG, built without multiply-add contraction: the
        # step functions round alike in both kernels, so the images are
        # bit-identical.
        assert diff == 0.0, f"{label}: view {i} ({v['kind']} {v['width']}x{v['height']} spp {v['spp']}): {diff:.3e}"
    print(f"{label}: {len(views)} views, worst relative difference {worst:.3e}", flush=True)


def close(batch, *groups):
    batch.close()
    for g in groups:
        for t in g:
            t.close()


@pytest.mark.parametrize("block", [64, 128, 256])
@pytest.mark.parametrize("use_graph", [1, 0])
def test_batch_matches_per_view(block, use_graph):
    print(f"batch parity block={block} graph={use_graph}", flush=True)
    views = small_views()
    frames = [v["frame"] for v in views]
    mine, ref = tracers(views), tracers(views)
    batch = Batch(mine)
    batch.set_launch(block=block, use_graph=use_graph, workgroups=64, blocks_per_pull=2)
    for _ in range(3):
        batch.encode(frames)
        for t, f in zip(ref, frames):
            t.encode(f)
    compare(views, mine, ref, f"block {block} graph {use_graph}")
    close(batch, mine, ref)


def test_batch_follows_frame_changes_and_limits():
    """A moved camera restarts one view's history and a sample limit shrinks
    another's last pass: the table is re-uploaded and the graph re-captured."""
    print("batch table changes", flush=True)
    views = small_views(6, seed=5)
    frames = [v["frame"] for v in views]
    limit = lambda v: 2 * v["spp"] + 1  # third pass takes one sample
    mine, ref = tracers(views, limit), tracers(views, limit)
    batch = Batch(mine)
    for p in range(5):
        if p == 2:
            g = frames[1].globals.copy()
            g["cam_forward"] = [0.05, 0.0, -1.0, g["cam_forward"][3]]
            frames[1] = scenes.Frame(g, frames[1].bodies, frames[1].ground)
        batch.encode(frames)
        for t, f in zip(ref, frames):
            t.encode(f)
    compare(views, mine, ref, "changes")
    close(batch, mine, ref)


def test_batch_view_times_and_orders():
    print("batch view times", flush=True)
    views = small_views(8, seed=3)
    frames = [v["frame"] for v in views]
    mine = tracers(views)
    batch = Batch(mine)
    for order in (0, 1, 2):
        batch.set_launch(order=order)
        batch.encode(frames)
        ms, span = batch.view_times()
        assert span > 0.0
        assert np.all(ms > 0.0) and np.all(ms <= span * 1.0001), (ms, span)
    for serial in (False, True):
        per_view = batch.encode_views(frames, serial=serial)
        assert np.all(per_view > 0.0)
    # One queue: completion times follow submission order.
    assert np.all(np.diff(per_view) >= -1e-3), per_view
    close(batch, mine)


def test_batch_autotune_then_parity():
    print("batch autotune", flush=True)
    views = small_views(6, seed=9)
    frames = [v["frame"] for v in views]
    mine, ref = tracers(views), tracers(views)
    batch = Batch(mine)
    best, ms = batch.autotune(frames, repeats=1)
    assert best["block"] in (64, 128, 256) and best["workgroups"] >= 1 and ms > 0.0
    assert batch.launch == best
    assert all(t.samples() == 0 for t in mine)
    for _ in range(2):
        batch.encode(frames)
        for t, f in zip(ref, frames):
            t.encode(f)
    compare(views, mine, ref, "after autotune")
    close(batch, mine, ref)


def test_batch_rejects_misuse():
    views = small_views(2, seed=1)
    mine = tracers(views)
    batch = Batch(mine)
    with pytest.raises(Exception):
        Batch([mine[0]])  # already a view of a batch
    with pytest.raises(Exception):
        batch.set_launch(block=96)
    with pytest.raises(ValueError):
        batch.encode([views[0]["frame"]])
    close(batch, mine)


@pytest.mark.skip(reason="batched passes of landscape views are switched off: their one GPU run after the "
                         "air-model change ended in a device memory fault that host sanitizer runs of the same "
                         "views did not reproduce; the case stays off the GPU until its cause is shown")
def test_batch_with_landscape_views():
    views = small_views(8, seed=11, halo_every=4)
    frames = [v["frame"] for v in views]
    mine, ref = tracers(views), tracers(views)
    batch = Batch(mine)
    for _ in range(3):
        batch.encode(frames)
        for t, f in zip(ref, frames):
            t.encode(f)
    compare(views, mine, ref, "landscape views")
    close(batch, mine, ref)


def test_batch_rejects_landscape_views():
    """The switched-off path fails loudly instead of rendering."""
    views = small_views(4, seed=11, halo_every=4)
    frames = [v["frame"] for v in views]
    mine = tracers(views)
    batch = Batch(mine)
    with pytest.raises(Exception, match="landscape"):
        batch.encode(frames)
    close(batch, mine)


def test_batched_serving_views_match_per_view_launches():
    """The batched-serving workload at full size (96 random systems, 128^2
    to 1080p, spp 4-64, air on a third of them): one batched pass against
    one sg_encode per view, bit for bit, with the autotuner's launch."""
    import parity

    print("batched serving workload: 96 views", flush=True)
    views = scenes.system_views(96)
    frames = [v["frame"] for v in views]
    mine, ref = tracers(views), tracers(views)
    batch = Batch(mine)
    batch.autotune(frames, repeats=1)
    batch.encode(frames)
    batch.synchronize()
    for t, f in zip(ref, frames):
        t.encode(f)
    for i, (ta, tb) in enumerate(zip(mine, ref)):
        assert ta.samples() == tb.samples() == views[i]["spp"], f"view {i} sample count"
        a, b = ta.read_accumulation(), tb.read_accumulation()
        assert np.all(np.isfinite(a)), f"view {i} non-finite"
        parity.assert_same_draws(a, b)
        parity.assert_same_draws(ta.read_backdrop(), tb.read_backdrop())
    close(batch, mine, ref)


def test_batched_serving_views_match_oracle_per_tile():
    """Views of the batched-serving generator shrunk to 32^2-80x45 pixels
    (same seed, same scene draws per view): each batched view against the
    FP32 oracle per 16x16 tile (3 sigma) and per pixel."""
    import oracle
    import parity

    views = scenes.system_views(12, min_pixels=32 * 32, max_pixels=80 * 45, spp=(8, 16))
    frames = [v["frame"] for v in views]
    mine = tracers(views)
    for t in mine:
        t.enable_moments(True)
    batch = Batch(mine)
    for _ in range(2):
        batch.encode(frames)
    batch.synchronize()
    zs, agree = [], []
    for i, (v, t) in enumerate(zip(views, mine)):
        print(f"oracle view {i}: {v['width']}x{v['height']} {v['spp']} spp {v['kind']}", flush=True)
        o = oracle.OracleTracer((v["width"], v["height"]),
                                Options(samples_per_frame=v["spp"], max_bounces=v["bounces"], milky_way=0.0), "f32")
        o.enable_moments(True)
        for _ in range(2):
            o.encode(v["frame"])
        a, am = t.read_accumulation(), t.read_moments()
        b, bm = o.read_accumulation(), o.read_moments()
        zs.append(parity.tile_z(a, am, b, bm).ravel())
        agree.append(parity.pixel_agreement(a, b))
    z = np.concatenate(zs)
    over, worst = parity.assert_tiles(z)
    print(f"batched views vs oracle: tiles over 3 sigma {over:.4f}, worst z {worst:.2f}, "
          f"pixels agreeing min {min(agree):.4f} mean {np.mean(agree):.4f}")
    assert min(agree) >= 0.95
    close(batch, mine)
