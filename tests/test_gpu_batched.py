"""The batched step's launch shapes against the default shape, through the
C ABI: persistent rasterizers, radix digit widths, graph replay and the
pair-overflow retry must not change what a step computes.

Bit-exact after the first step: records, pairs, tile ranges, the forward
image and LAST (the same arithmetic in the same order whatever the shape).
Float-atomic sums (GRAD2D) agree to 1e-5 of their largest value: the order
of the atomic additions is not fixed, so neither are the last bits of the
Adam update, and later steps (and the parameters after them) are compared
within 1e-4.
"""
import numpy as np
import pytest

from helpers import gs, make_camera, make_scene, smooth_target

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def device():
    name = gs.init(0)
    print(f"\n[gpu] {name}", flush=True)
    yield name
    gs.synchronize()


def case(msg):
    print(f"[case] {msg}", flush=True)


def train_params():
    return gs.TrainParams(lambda_dssim=0.2)


def step_state(ws, n, views):
    out = []
    for v in range(views):
        pairs = int(ws.download(gs.BUF_TOTALS, np.uint32, 1, view=v)[0])
        T = ws.tiles_x * ws.tiles_y
        out.append({
            "records": ws.download(gs.BUF_RECORDS, np.float32, n * 12, view=v),
            "pairs": ws.download(gs.BUF_PAIR_GAUSSIANS, np.uint32, pairs, view=v),
            "ranges": ws.download(gs.BUF_RANGES, np.uint32, (T + 1) * 2, view=v),
            "image": ws.download(gs.BUF_IMAGE, np.float32, view=v),
            "last": ws.download(gs.BUF_LAST, np.uint32, view=v),
            "grad2d": ws.download(gs.BUF_GRAD2D, np.float32, 9 * n, view=v).reshape(n, 9).T,
        })
    return out


def run_steps(n, deg, w, h, views, steps, seed, tuning, reserve=None):
    """`steps` batched optimizer steps from the same seeded start; returns the
    state after the first step, the parameters after the last, and the loss
    sums per step."""
    params, sh = make_scene(n, deg, seed)
    cams = [make_camera(w, h, 0.4 + 1.1 * v) for v in range(views)]
    targets = [gs.DeviceBuffer(smooth_target(w, h, seed + v)) for v in range(views)]
    try:
        with gs.Workspace(n, deg, w, h) as ws:
            ws.set_views(views)
            if reserve is not None:
                ws.reserve_pairs(reserve)
            ws.set_tuning(**tuning)
            ws.load(params, sh)
            losses, first = [], None
            for s in range(steps):
                frames = [gs.make_frame(c, n, deg, tp=train_params(), step=s + 1) for c in cams]
                _, ls = ws.train_views(frames, [t.ptr for t in targets])
                losses.append(ls)
                if first is None:
                    first = step_state(ws, n, views)
                # eager work of the program's own between (graph) replays
                ws.download(gs.BUF_TOTALS, np.uint32, 16)
            after = ws.store(n)[0]
            overflows = ws.overflows()
    finally:
        for t in targets:
            t.free()
    return first, after, np.asarray(losses), overflows


BASE = dict(raster_groups_per_cu=0, depth_digit_bits=8, tile_digit_bits=8, graph=0)


def compare(a, b):
    sa, pa, la, _ = a
    sb, pb, lb, _ = b
    for va, vb in zip(sa, sb):
        np.testing.assert_array_equal(va["records"].view(np.uint32), vb["records"].view(np.uint32))
        np.testing.assert_array_equal(va["pairs"], vb["pairs"])
        np.testing.assert_array_equal(va["ranges"], vb["ranges"])
        np.testing.assert_array_equal(va["last"], vb["last"])
        np.testing.assert_array_equal(va["image"].view(np.uint32), vb["image"].view(np.uint32))
        scale = np.abs(va["grad2d"]).max(axis=1, keepdims=True) + 1e-30
        assert (np.abs(va["grad2d"] - vb["grad2d"]) / scale).max() <= 1e-5
    np.testing.assert_allclose(la[0], lb[0], rtol=1e-6)  # first step: same render
    np.testing.assert_allclose(la, lb, rtol=1e-4)
    assert np.abs(pa - pb).max() <= 1e-4 * max(1.0, np.abs(pa).max())


SHAPE_CASES = [
    ("persistent-1", dict(BASE, raster_groups_per_cu=1)),
    ("persistent-4", dict(BASE, raster_groups_per_cu=4)),
    ("counting-tile-sort", dict(BASE, tile_digit_bits=0)),
    ("digits-6-6", dict(BASE, depth_digit_bits=6, tile_digit_bits=6)),
    ("digits-5-4", dict(BASE, depth_digit_bits=5, tile_digit_bits=4)),
    ("graph", dict(BASE, graph=1)),
    ("graph-persistent-digits", dict(raster_groups_per_cu=2, depth_digit_bits=6, tile_digit_bits=4, graph=1)),
]


@pytest.mark.parametrize("name,tuning", SHAPE_CASES, ids=[c[0] for c in SHAPE_CASES])
def test_launch_shapes_compute_the_same_step(name, tuning):
    case(f"launch shape {name}: 3 views, 3 steps")
    args = (1500, 3, 160, 120, 3, 3, 71)
    base = run_steps(*args, BASE)
    other = run_steps(*args, tuning)
    compare(base, other)


def test_depth16_order_with_small_digits():
    case("depth16 keys (2 passes of 8 bits, or 4 of 4; radix or counting tile sort) give the same pairs")
    n, deg, w, h = 1200, 1, 128, 96
    params, sh = make_scene(n, deg, 72)
    cam = make_camera(w, h, 1.0)
    out = []
    for bits, tile_bits in ((8, 8), (4, 4), (8, 0)):
        with gs.Workspace(n, deg, w, h) as ws:
            ws.set_tuning(depth_digit_bits=bits, tile_digit_bits=tile_bits)
            ws.load(params, sh)
            frame = gs.make_frame(cam, n, deg, flags=gs.FLAG_DEPTH16)
            ws.project(frame)
            pairs = ws.bin(frame)
            out.append((ws.download(gs.BUF_PAIR_GAUSSIANS, np.uint32, pairs),
                        ws.download(gs.BUF_DEPTH_ORDER, np.uint32, n)))
    for other in out[1:]:
        np.testing.assert_array_equal(out[0][0], other[0])
        np.testing.assert_array_equal(out[0][1], other[1])


def test_pair_overflow_retry_matches_a_step_with_room():
    """trainer.cpp:819-836 and test_trainer.cpp:832-853: a step whose pairs
    overflow the buffers leaves the model alone and runs again; the result
    is that of a step that had room from the start."""
    case("pair overflow: large splats, first step overflows the initial buffers")
    n, deg, w, h = 6000, 0, 512, 384
    params, sh = make_scene(n, deg, 73)
    params = params.copy()
    params[:, 4:7] += 1.5  # large splats: many tiles each
    cam = make_camera(w, h, 1.2)
    target = gs.DeviceBuffer(smooth_target(w, h, 73))
    results = []
    try:
        for reserve in (None, 4_000_000):
            with gs.Workspace(n, deg, w, h) as ws:
                if reserve:
                    ws.reserve_pairs(reserve)
                ws.load(params, None)
                frame = gs.make_frame(cam, n, deg, tp=train_params(), step=1)
                pairs, losses = ws.train_views([frame], [target.ptr])
                results.append((pairs[0], losses[0], ws.store(n)[0], ws.overflows(),
                                ws.download(gs.BUF_GRAD2D, np.float32, 9 * n).reshape(n, 9).T))
    finally:
        target.free()
    (p0, l0, a0, o0, g0), (p1, l1, a1, o1, g1) = results
    print(f"[info] pairs {p0}, overflow steps {o0} / {o1}", flush=True)
    assert p0 == p1 and p0 > 6000 * 4 * 5 // 4
    assert o0 >= 1 and o1 == 0
    np.testing.assert_allclose(l0, l1, rtol=1e-6)
    scale = np.abs(g1).max(axis=1, keepdims=True) + 1e-30
    assert (np.abs(g0 - g1) / scale).max() <= 1e-5
    assert np.abs(a0 - a1).max() <= 1e-5 * max(1.0, np.abs(a1).max())


def test_no_update_flag_leaves_the_model_alone():
    case("GS_HIP_FLAG_NO_UPDATE: parameters, moments, stats unchanged")
    n, deg, w, h = 800, 2, 96, 72
    params, sh = make_scene(n, deg, 74)
    cam = make_camera(w, h, 1.0)
    with gs.Workspace(n, deg, w, h) as ws:
        ws.load(params, sh)
        ws.set_target(smooth_target(w, h, 74))
        before = [ws.download(b, np.float32) for b in (gs.BUF_PARAMS, gs.BUF_M_PARAMS, gs.BUF_STATS, gs.BUF_SH)]
        frame = gs.make_frame(cam, n, deg, tp=train_params(), step=1, flags=gs.FLAG_NO_UPDATE)
        pairs, losses = ws.train_step(frame)
        after = [ws.download(b, np.float32) for b in (gs.BUF_PARAMS, gs.BUF_M_PARAMS, gs.BUF_STATS, gs.BUF_SH)]
        g2 = ws.download(gs.BUF_GRAD2D, np.float32, 9 * n)
    assert pairs > 0 and np.abs(g2).max() > 0
    for b, a in zip(before, after):
        np.testing.assert_array_equal(b.view(np.uint32), a.view(np.uint32))


def test_autotune_keeps_a_valid_shape_and_the_model():
    case("autotune on 2 views: a valid shape, the model untouched")
    n, deg, w, h = 3000, 3, 192, 144
    params, sh = make_scene(n, deg, 75)
    cams = [make_camera(w, h, 0.7 + v) for v in range(2)]
    with gs.Workspace(n, deg, w, h) as ws:
        ws.set_views(2)
        ws.load(params, sh)
        before = ws.download(gs.BUF_PARAMS, np.float32)
        frames = [gs.make_frame(c, n, deg, tp=train_params(), step=1) for c in cams]
        best, ms = ws.autotune(frames, None, repeats=2)
        after = ws.download(gs.BUF_PARAMS, np.float32)
        assert ws.tuning().as_dict() == best.as_dict()
    print(f"[info] best {best.as_dict()} at {ms:.3f} ms", flush=True)
    assert ms > 0
    assert best.raster_groups_per_cu in (0, 1, 2, 4, 8)
    assert 4 <= best.depth_digit_bits <= 8 and (best.tile_digit_bits == 0 or 4 <= best.tile_digit_bits <= 8)
    np.testing.assert_array_equal(before.view(np.uint32), after.view(np.uint32))


def test_step_targets_are_read_in_place():
    """A step's device targets are read where they lie: TARGET keeps what
    set_target put there, a step given the same photo through its pointer
    equals a step on TARGET, and a later step without targets reads TARGET."""
    case("targets read in place: TARGET untouched, pointer step == TARGET step")
    n, deg, w, h = 900, 1, 96, 72
    params, sh = make_scene(n, deg, 76)
    cam = make_camera(w, h, 1.2)
    photo, other = smooth_target(w, h, 76), smooth_target(w, h, 77)
    buf = gs.DeviceBuffer(photo)
    try:
        frame = gs.make_frame(cam, n, deg, tp=train_params(), step=1, flags=gs.FLAG_NO_UPDATE)
        with gs.Workspace(n, deg, w, h) as ws:
            ws.load(params, sh)
            ws.set_target(other)
            _, by_pointer = ws.train_views([frame], [buf.ptr])
            g_pointer = ws.download(gs.BUF_GRAD2D, np.float32, 9 * n)
            kept = ws.download(gs.BUF_TARGET, np.uint8)
            _, on_other = ws.train_views([frame], None)
            ws.set_target(photo)
            _, on_target = ws.train_views([frame], None)
            g_target = ws.download(gs.BUF_GRAD2D, np.float32, 9 * n)
    finally:
        buf.free()
    np.testing.assert_array_equal(kept, other.reshape(-1))
    np.testing.assert_allclose(np.asarray(by_pointer), np.asarray(on_target), rtol=1e-6)
    assert abs(on_other[0][0] - on_target[0][0]) > 1e-3 * abs(on_target[0][0])
    scale = np.abs(g_target).max()
    assert scale > 0 and np.abs(g_pointer - g_target).max() <= 1e-5 * scale


def test_work_order_is_every_tile_once_longest_first():
    """The rasterizers' work list (TILE_ORDER) after a 3-view step: 2x2 quads
    of neighbouring tiles of one view as four consecutive entries
    (view << 16 | tile, 0xffffffff for a member outside the tile grid),
    holding every (view, tile) exactly once, the quads bucketed by
    floor(log2(their longest list)) longest first, empty quads last."""
    case("work order: the 3 views' tiles in 2x2 quads, longest lists first")
    n, deg, w, h = 2500, 1, 200, 150
    params, sh = make_scene(n, deg, 78)
    cams = [make_camera(w, h, 0.4 + 1.1 * v) for v in range(3)]
    bufs = [gs.DeviceBuffer(smooth_target(w, h, 78 + v)) for v in range(3)]
    try:
        with gs.Workspace(n, deg, w, h) as ws:
            ws.set_views(3)
            ws.load(params, sh)
            ws.train_views([gs.make_frame(c, n, deg, tp=train_params(), step=1) for c in cams], [b.ptr for b in bufs])
            tx, ty = ws.tiles_x, ws.tiles_y
            T = tx * ty
            quads = ((tx + 1) // 2) * ((ty + 1) // 2)
            order = ws.download(gs.BUF_TILE_ORDER, np.uint32, 3 * 4 * quads)
            lengths = []
            for v in range(3):
                r = ws.download(gs.BUF_RANGES, np.uint32, (T + 1) * 2, view=v).reshape(-1, 2)[:T]
                lengths.append((r[:, 1] - r[:, 0]).astype(np.int64))
    finally:
        for b in bufs:
            b.free()
    assert tx % 2 == 1 or ty % 2 == 1, "the case covers quads that stick out of the grid"
    entries = [x for x in order.tolist() if x != 0xFFFFFFFF]
    want = sorted((v << 16) | t for v in range(3) for t in range(T))
    assert sorted(entries) == want
    assert order.size - len(entries) == 3 * (4 * quads - T)
    longest = []
    for g in order.reshape(-1, 4).tolist():
        members = [x for x in g if x != 0xFFFFFFFF]
        assert members, "a quad with no tile"
        views = {x >> 16 for x in members}
        assert len(views) == 1, "a quad over two views"
        xs = {(x & 0xFFFF) % tx for x in members}
        ys = {(x & 0xFFFF) // tx for x in members}
        assert max(xs) - min(xs) <= 1 and max(ys) - min(ys) <= 1 and min(xs) % 2 == 0 and min(ys) % 2 == 0
        longest.append(max(lengths[x >> 16][x & 0xFFFF] for x in members))
    lens = np.array(longest)
    bucket = np.where(lens > 0, 31 - np.floor(np.log2(np.maximum(lens, 1))).astype(np.int64), 32)
    assert np.all(np.diff(bucket) >= 0), "buckets out of order"
    assert lens.max() > 0
