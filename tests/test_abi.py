"""The C ABI without a GPU: the library loads, exports every symbol the
header declares, its host helpers match the oracle / reference fixtures, and
bad shapes are refused before anything touches a device."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

import chx
import oracle as O


def test_exports_every_header_symbol():
    syms = chx.header_symbols()
    assert len(syms) >= 18
    out = subprocess.run(["nm", "-D", "--defined-only", chx.LIB_PATH], check=True,
                         capture_output=True, text=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    missing = [s for s in syms if s not in exported]
    assert not missing, missing
    for s in syms:
        assert hasattr(chx.lib(), s)


def test_num_blocks_and_maps():
    for N in (32, 64, 256, 2048, 4096):
        assert chx.num_blocks(N) == O.num_blocks(N)
        xm, ym = chx.block_maps(N)
        ox, oy = O.block_maps(N)
        np.testing.assert_array_equal(xm, ox)
        np.testing.assert_array_equal(ym, oy)
    assert chx.num_blocks(48) == -1 and chx.num_blocks(0) == -1


def test_blocks_to_triangle_matches_reference_reorg(golden):
    for N, F in ((96, 2), (32, 1), (256, 1)):
        nb = chx.num_blocks(N)
        blocks = np.arange(F * nb * 2048, dtype=np.int32)
        np.testing.assert_array_equal(chx.blocks_to_triangle(blocks, N),
                                      golden[f"reorg_N{N}_F{F}__tri"])


def test_strerror_and_caps():
    for s in range(0, 10):
        assert chx.lib().chx_strerror(s)
    assert chx.lib().chx_max_time_accum(0) == 288
    assert chx.lib().chx_max_time_accum(1) == 2272
    assert chx.lib().chx_max_time_accum(2) == 0
    assert chx.lib().chx_max_time_accum(3) == 0
    assert chx.lib().chx_max_time_accum(7) == -1
    assert b"gfx950" in chx.lib().chx_version()


def _create(**kw):
    base = dict(num_elements=256, num_freq=1, time_steps=256, time_accum=256,
                upper_triangle_convention=1, kernel_batch=0, device=0)
    base.update(kw)
    cfg = chx.Config(**base)
    h = ctypes.c_void_p()
    rc = chx.lib().chx_plan_create(ctypes.byref(cfg), ctypes.byref(h))
    if rc == 0:
        chx.lib().chx_plan_destroy(h)
    return rc


@pytest.mark.parametrize("kw,status", [
    (dict(num_elements=48), chx.ERR_SHAPE),
    (dict(num_elements=0), chx.ERR_SHAPE),
    (dict(num_freq=0), chx.ERR_SHAPE),
    (dict(time_steps=272, time_accum=8), chx.ERR_SHAPE),     # T % 32
    (dict(time_steps=0), chx.ERR_SHAPE),
    (dict(time_accum=12), chx.ERR_TIME_ACCUM),               # ta % 8
    (dict(time_steps=320, time_accum=256), chx.ERR_TIME_ACCUM),  # T % ta
    (dict(time_steps=296 * 4, time_accum=296), chx.ERR_TIME_ACCUM),  # conference cap
    (dict(kernel_batch=1, time_steps=2280 * 4, time_accum=2280), chx.ERR_TIME_ACCUM),
    (dict(kernel_batch=4), chx.ERR_INVALID_ARGUMENT),
    (dict(upper_triangle_convention=2), chx.ERR_INVALID_ARGUMENT),
    (dict(num_freq=2 ** 20, num_elements=4096), chx.ERR_SHAPE),  # grid bound
])
def test_validation_is_host_side(kw, status):
    assert _create(**kw) == status


def test_null_arguments():
    L = chx.lib()
    assert L.chx_plan_create(None, None) == chx.ERR_INVALID_ARGUMENT
    h = ctypes.c_void_p()
    assert L.chx_plan_create(None, ctypes.byref(h)) == chx.ERR_INVALID_ARGUMENT
    assert L.chx_plan_destroy(None) == chx.ERR_INVALID_ARGUMENT
    assert L.chx_correlate(None, None, None, None) == chx.ERR_INVALID_ARGUMENT
    assert L.chx_blocks_to_triangle(1, 64, None, None) == chx.ERR_INVALID_ARGUMENT


def test_valid_config_without_gpu_reports_no_device():
    if chx.device_count() > 0:
        pytest.skip("a GPU is visible")
    assert _create() == chx.ERR_NO_DEVICE
    assert _create(kernel_batch=2, time_accum=8) == chx.ERR_NO_DEVICE
    with pytest.raises(chx.ChxError) as e:
        chx.Plan(256, 1, 256)
    assert e.value.status == chx.ERR_NO_DEVICE


def test_stream_event_and_device_helpers_host_side():
    L = chx.lib()
    vp = ctypes.c_void_p
    L.chx_stream_create.argtypes = [ctypes.c_int32, ctypes.POINTER(vp)]
    L.chx_event_create.argtypes = [ctypes.POINTER(vp)]
    L.chx_device_alloc.argtypes = [ctypes.c_int32, ctypes.c_size_t, ctypes.POINTER(vp)]
    for name in ("chx_stream_destroy", "chx_event_destroy", "chx_device_free"):
        getattr(L, name).argtypes = [vp]
    L.chx_event_record.argtypes = [vp, vp]
    L.chx_stream_wait_event.argtypes = [vp, vp]
    assert L.chx_stream_create(0, None) == chx.ERR_INVALID_ARGUMENT
    assert L.chx_event_create(None) == chx.ERR_INVALID_ARGUMENT
    assert L.chx_device_alloc(0, 16, None) == chx.ERR_INVALID_ARGUMENT
    assert L.chx_stream_destroy(None) == chx.ERR_INVALID_ARGUMENT
    assert L.chx_event_destroy(None) == chx.ERR_INVALID_ARGUMENT
    assert L.chx_device_free(None) == chx.ERR_INVALID_ARGUMENT
    assert L.chx_event_record(None, None) == chx.ERR_INVALID_ARGUMENT
    assert L.chx_stream_wait_event(None, None) == chx.ERR_INVALID_ARGUMENT
    assert L.chx_device_arch(0, None, 0) == chx.ERR_INVALID_ARGUMENT
    if chx.device_count() == 0:
        h = vp()
        assert L.chx_stream_create(0, ctypes.byref(h)) == chx.ERR_NO_DEVICE and not h.value
        assert L.chx_device_alloc(0, 16, ctypes.byref(h)) == chx.ERR_NO_DEVICE
        assert chx.device_arch(0) is None
        # the tuning cache falls back to the gfx950 key without a device
        assert chx.job_signature([dict(num_elements=64, num_freq=1, time_steps=64)]) \
            .startswith("gfx950/")


def test_noalloc_refusals_are_host_side():
    # chx_correlate_noalloc refuses a request its fixed scratch cannot hold
    # and the matrix-core methods before touching a device (the pointers are
    # never used).
    fake = 1 << 40
    with pytest.raises(chx.ChxError) as e:
        chx.correlate_noalloc(4096, 64, 256, fake, fake)  # 2 MiB of sums + the maps
    assert e.value.status == chx.ERR_BUDGET and "limit" in str(e.value)
    with pytest.raises(chx.ChxError) as e:
        chx.correlate_noalloc(256, 1, 256, fake, fake, kernel_batch=chx.FP6_MFMA)
    assert e.value.status == chx.ERR_INVALID_ARGUMENT and "plan" in str(e.value)
    with pytest.raises(chx.ChxError) as e:
        chx.correlate_noalloc(256, 1, 250, fake, fake)
    assert e.value.status == chx.ERR_SHAPE
    # the largest request that fits: 4096 x 32 (config 5's array)
    if chx.device_count() == 0:
        with pytest.raises(chx.ChxError) as e:
            chx.correlate_noalloc(4096, 32, 256, fake, fake)
        assert e.value.status == chx.ERR_NO_DEVICE


def _batch_create(jobs, tuning=None, n=None):
    arr = (chx.JobDesc * max(1, len(jobs)))()
    for k, j in enumerate(jobs):
        base = dict(num_elements=256, num_freq=1, time_steps=4096,
                    upper_triangle_convention=1, reserved=0, d_input=0x1000, d_vis=0x2000)
        base.update(j)
        arr[k] = chx.JobDesc(**base)
    tp = ctypes.byref(chx.BatchTuning(**tuning)) if tuning else None
    h = ctypes.c_void_p()
    rc = chx.lib().chx_batch_create(0, arr, len(jobs) if n is None else n, tp, ctypes.byref(h))
    if rc == 0:
        chx.lib().chx_batch_destroy(h)
    return rc


@pytest.mark.parametrize("jobs,tuning,status", [
    ([dict(num_elements=48)], None, chx.ERR_SHAPE),
    ([dict(), dict(num_freq=0)], None, chx.ERR_SHAPE),
    ([dict(time_steps=0)], None, chx.ERR_SHAPE),
    ([dict(upper_triangle_convention=2)], None, chx.ERR_INVALID_ARGUMENT),
    ([dict(reserved=1)], None, chx.ERR_INVALID_ARGUMENT),
    ([dict(d_input=None)], None, chx.ERR_INVALID_ARGUMENT),
    ([dict(d_vis=None)], None, chx.ERR_INVALID_ARGUMENT),
    ([dict()], dict(workgroups=12), chx.ERR_INVALID_ARGUMENT),
    ([dict()], dict(max_item_stages=-4), chx.ERR_INVALID_ARGUMENT),
    ([dict()], dict(method=1), chx.ERR_INVALID_ARGUMENT),
    # int8: one ring (256 samples) of 2 planes x 256 elements is 128 KiB
    ([dict()], dict(workspace_bytes=128 * 1024 - 1, method=2), chx.ERR_BUDGET),
    ([dict(num_elements=160), dict(num_elements=32)],
     dict(workspace_bytes=(512 + 256) * 256 - 1, method=2), chx.ERR_BUDGET),
    # fp6 (the default): one ring (256 samples) at 1.5 bytes per value, 96 KiB
    ([dict()], dict(workspace_bytes=96 * 1024 - 1), chx.ERR_BUDGET),
    ([dict(num_elements=160), dict(num_elements=32)],
     dict(workspace_bytes=384 * (256 + 128) - 1, method=3), chx.ERR_BUDGET),
])
def test_batch_validation_is_host_side(jobs, tuning, status):
    assert _batch_create(jobs, tuning) == status


def test_batch_null_arguments_and_no_device():
    L = chx.lib()
    h = ctypes.c_void_p()
    assert L.chx_batch_create(0, None, 1, None, ctypes.byref(h)) == chx.ERR_INVALID_ARGUMENT
    assert _batch_create([dict()], n=0) == chx.ERR_INVALID_ARGUMENT
    assert L.chx_batch_run(None, 0, None) == chx.ERR_INVALID_ARGUMENT
    assert L.chx_batch_get_info(None, None) == chx.ERR_INVALID_ARGUMENT
    assert L.chx_batch_destroy(None) == chx.ERR_INVALID_ARGUMENT
    if chx.device_count() == 0:
        assert _batch_create([dict(), dict(num_elements=2048, num_freq=8)]) == chx.ERR_NO_DEVICE
        assert _batch_create([dict()], dict(workspace_bytes=128 * 1024, method=2)) == \
            chx.ERR_NO_DEVICE
        assert _batch_create([dict()], dict(workspace_bytes=96 * 1024)) == chx.ERR_NO_DEVICE


def _fake(jobs):
    return [dict(j, d_input=0x1000, d_vis=0x2000) for j in jobs]


def _tiles(nt, method):
    """Work items of one channel (both matrix-core methods): tiles of
    128 x 256, row tile i holding ceil((nt - i) / 2) of them, with the half
    tiles (the last one of each row tile with nt - i odd) two to an item."""
    wide = sum((nt - i + 1) // 2 for i in range(nt))
    return wide - (nt + 1) // 2 // 2


def _full_tile_order(nt, band=8):
    """(row tile, first column tile) of the full 128 x 256 tiles in the
    kernels' item order: bands of 8 row tiles; in a band, column tile by
    column tile from the band's first row on, and in a column the band's
    rows of its parity up to it (ct >= ti, ct - ti even, ct <= nt - 2)."""
    out = []
    for r0 in range(0, nt, band):
        for ct in range(r0, nt - 1):
            out += [(ti, ct) for ti in range(r0, min(r0 + band, nt, ct + 1))
                    if (ct - ti) % 2 == 0]
    return out


def test_full_tile_order_is_bands_of_row_tiles():
    # Every full tile once; and any 32 consecutive items (the workgroups of
    # one XCD) read the planes of at most 22 tiles of 128 elements however
    # large the array (row by row: every column tile of the channel).
    for nt in range(1, 70):
        order = _full_tile_order(nt)
        assert sorted(order) == [(r, r + 2 * k) for r in range(nt) for k in range((nt - r) // 2)]

    def touched(win):
        return len({t for r, c in win for t in (r, c, c + 1)})
    for nt in (32, 64, 128):
        order = _full_tile_order(nt)
        assert max(touched(order[i:i + 32]) for i in range(len(order) - 31)) <= 22


def _item_blocks(nt):
    """The 32x32 blocks each work item of one channel stores, following
    the kernel's item order: full tiles block by block, then pairs of half
    tiles (rows nt-1-4q and nt-3-4q, column tile nt-1), an odd last one
    alone."""
    items = []
    for r, ct in _full_tile_order(nt):
        items.append({(4 * r + i, 4 * ct + j) for i in range(4) for j in range(8)})
    halves = list(range(nt - 1, -1, -2))
    for q in range(0, len(halves), 2):
        rows = halves[q:q + 2]
        items.append({(4 * r + i, 4 * (nt - 1) + j) for r in rows for i in range(4)
                      for j in range(4)})
    return items


def _small_item_blocks(nt):
    """The 32x32 blocks each 128 x 128 work item stores: tiles row-major over
    the upper triangle of 128-element tiles, wave w computing rows
    64 (w/2).. and columns 64 (w%2).. of its tile as 2x2 blocks."""
    items = []
    for ti in range(nt):
        for ct in range(ti, nt):
            items.append({(4 * ti + 2 * (w >> 1) + i, 4 * ct + 2 * (w & 1) + j)
                          for w in range(4) for i in range(2) for j in range(2)})
    return items


def _expected_items(jobs, chunk, max_st, unit, method, small=0):
    """Work items of the plan: per round, per active job, F x tiles x pieces
    (pieces hold at most 2048 stages, FP6 1024; FP6 jobs of at most `small`
    elements on 128 x 128 tiles)."""
    n, rounds = 0, -(-max(j["time_steps"] for j in jobs) // chunk)
    for r in range(rounds):
        for j in jobs:
            t0 = r * chunk
            if t0 >= j["time_steps"]:
                continue
            nst = -(-min(chunk, j["time_steps"] - t0) // unit) * (unit // 64)
            per = max_st if 0 < max_st < nst else nst
            per = min(per, 1024 if method == 3 else 2048)
            nt = -(-j["num_elements"] // 128)
            tiles = (nt * (nt + 1) // 2 if method == 3 and j["num_elements"] <= small
                     else _tiles(nt, method))
            n += j["num_freq"] * tiles * -(-nst // per)
    return rounds, n


def test_tile_count_closed_forms():
    for nt in range(1, 40):
        m = nt // 2
        wide = (m + 1) ** 2 if nt % 2 else m * (m + 1)
        assert _tiles(nt, 2) == _tiles(nt, 3) == wide - (nt + 1) // 2 // 2


def test_items_cover_the_upper_triangle_once():
    # Every upper-triangle 32x32 block of an nt-tile array is stored by
    # exactly one work item (blocks below the diagonal are computed and
    # not stored; a lone half tile's columns past the array are not listed,
    # as they are never stored), and the item count is the planner's.
    for nt in range(1, 24):
        items = _item_blocks(nt)
        assert len(items) == _tiles(nt, 3)
        seen = {}
        for k, blocks in enumerate(items):
            for (by, bx) in blocks:
                if bx >= by and bx < 4 * nt:
                    assert (by, bx) not in seen, (nt, by, bx)
                    seen[(by, bx)] = k
        assert len(seen) == 4 * nt * (4 * nt + 1) // 2
        for blocks in items:
            assert all(bx < 4 * nt for (_, bx) in blocks)


def test_small_tiles_cover_the_upper_triangle_once():
    # With 128 x 128 tiles every upper-triangle block is stored by exactly
    # one work item, and nothing is computed past the array.
    for nt in range(1, 24):
        items = _small_item_blocks(nt)
        assert len(items) == nt * (nt + 1) // 2
        seen = {}
        for k, blocks in enumerate(items):
            assert all(bx < 4 * nt and by < 4 * nt for (by, bx) in blocks)
            for (by, bx) in blocks:
                if bx >= by:
                    assert (by, bx) not in seen, (nt, by, bx)
                    seen[(by, bx)] = k
        assert len(seen) == 4 * nt * (4 * nt + 1) // 2


def _batch_workspace(jobs, chunk, unit, method):
    """Workspace of a batch plan at a given chunk: every job's planes, then
    per job the most code-sum rows (runs per channel and element group) any
    round writes, 4 bytes per element, and two 8-byte corrections per
    element, each region 256-aligned."""
    r256 = lambda x: -(-x // 256) * 256
    ws = 0
    nps = [-(-j["num_elements"] // 128) * 128 for j in jobs]
    for j, Np in zip(jobs, nps):
        stages = min(chunk, -(-j["time_steps"] // unit) * unit) // 64
        ws += (128 if method == 2 else 96) * j["num_freq"] * Np * stages
    for j, Np in zip(jobs, nps):
        T, F, rows = j["time_steps"], j["num_freq"], 0
        for t0 in range(0, T, chunk):
            stages = -(-min(chunk, T - t0) // unit) * unit // 64
            if method == 2:
                units = 4 * stages
            else:
                hpu = 1 if Np >= 1024 else 1024 // Np
                units = -(-2 * stages // hpu)
            per = min(max(units // 8, 1), 16)
            rows = max(rows, -(-units // per))
        ws += r256(rows * F * Np * 4) + r256(2 * F * Np * 8)
    return ws


@pytest.mark.parametrize("method", [2, 3])
@pytest.mark.parametrize("tuning", [None, dict(max_item_stages=64), dict(max_item_stages=6),
                                    dict(workspace_bytes=1 << 30),
                                    dict(workspace_bytes=3 << 30, max_item_stages=512),
                                    dict(small_tile_elements=512),
                                    dict(small_tile_elements=4096, max_item_stages=256)])
def test_batch_plan_host_only(tuning, method):
    import workloads as W
    jobs = W.uneven_batch_jobs()
    t = dict(tuning or {}, method=method)
    info = chx.batch_plan(_fake(jobs), t)
    ring = 4
    unit = 64 * ring
    bytes_per_value = 2.0 if method == 2 else 1.5
    per_unit = sum(int(bytes_per_value * unit) * j["num_freq"] * -(-j["num_elements"] // 128)
                   * 128 for j in jobs)
    budget = t.get("workspace_bytes", 16 << 30)
    maxT = max(j["time_steps"] for j in jobs)
    chunk = min(budget // per_unit * unit, -(-maxT // unit) * unit)
    # The chunk then shrinks until planes, code-sum rows and offset
    # corrections fit the budget together.
    ws = _batch_workspace(jobs, chunk, unit, method)
    while ws > budget and chunk > unit:
        c = int(chunk * budget / ws) // unit * unit
        chunk = max(min(c, chunk - unit), unit)
        ws = _batch_workspace(jobs, chunk, unit, method)
    max_st = -(-t.get("max_item_stages", 0) // ring) * ring
    small = t.get("small_tile_elements", 0)
    rounds, items = _expected_items(jobs, chunk, max_st, unit, method, small)
    assert info["chunk_samples"] == chunk
    assert info["rounds"] == rounds and info["items"] == items
    assert info["max_item_stages"] == max_st
    assert info["workspace_bytes"] <= budget
    assert info["cmacs"] == sum(W.cmacs(j["num_elements"], j["num_freq"], j["time_steps"])
                                for j in jobs)
    assert info["num_jobs"] == 64
    assert info["method"] == method
    assert info["small_tile_jobs"] == (sum(j["num_elements"] <= small for j in jobs)
                                       if method == 3 else 0)


@pytest.mark.parametrize("method, cap", [(2, 2048), (3, 1024)])
def test_items_hold_at_most_their_cap(method, cap):
    """A long job is cut into work items of <= 2048 stages (int8: its int32
    sums of 64x products stay <= 2^30) or <= 1024 (FP6: its FP32 sums of
    code products, from -2^24, stay within 2^24, where they are exact)."""
    job = [dict(num_elements=256, num_freq=3, time_steps=5000 * 64)]  # 5000 stages
    plan = chx.batch_plan(_fake(job), dict(workspace_bytes=64 << 30, method=method))
    assert plan["rounds"] == 1 and plan["chunk_samples"] == 5000 * 64
    # channels x tiles (nt = 2) x ceil(5000 / cap) pieces
    assert plan["items"] == 3 * 2 * -(-5000 // cap)


def test_batch_plan_validates_like_create():
    with pytest.raises(chx.ChxError) as e:
        chx.batch_plan(_fake([dict(num_elements=40, num_freq=1, time_steps=64)]))
    assert e.value.status == chx.ERR_SHAPE
    with pytest.raises(chx.ChxError) as e:
        chx.batch_plan(_fake([dict(num_elements=256, num_freq=1, time_steps=64)]),
                       dict(workspace_bytes=1000))
    assert e.value.status == chx.ERR_BUDGET
    one = chx.batch_plan(_fake([dict(num_elements=32, num_freq=1, time_steps=1)]))
    assert one["rounds"] == 1 and one["items"] == 1 and one["chunk_samples"] == 256
    one = chx.batch_plan(_fake([dict(num_elements=32, num_freq=1, time_steps=1)]),
                         dict(method=2))
    assert one["rounds"] == 1 and one["items"] == 1 and one["chunk_samples"] == 256


def test_serve_validation_is_host_side():
    L = chx.lib()
    h = ctypes.c_void_p()
    assert L.chx_serve_create(None, ctypes.byref(h)) == chx.ERR_INVALID_ARGUMENT

    def create(**kw):
        base = dict(device=0, frame_slots=4, slot_bytes=1 << 20, workspace_bytes=0,
                    hbm_budget_bytes=0, max_batch_frames=0, max_item_stages=0)
        base.update(kw)
        cfg = chx.ServeConfig(**base)
        rc = L.chx_serve_create(ctypes.byref(cfg), ctypes.byref(h))
        if rc == 0:
            L.chx_serve_destroy(h)
        return rc

    assert create(frame_slots=0) == chx.ERR_INVALID_ARGUMENT
    assert create(slot_bytes=-1) == chx.ERR_INVALID_ARGUMENT
    assert create(max_item_stages=-4) == chx.ERR_INVALID_ARGUMENT
    # 4 slots of 1 MiB + 2 * 4 * 1 MiB of workspace = 12 MiB
    assert create(hbm_budget_bytes=12 * (1 << 20) - 1) == chx.ERR_BUDGET
    if chx.device_count() == 0:
        assert create(hbm_budget_bytes=12 * (1 << 20)) == chx.ERR_NO_DEVICE
    sid = ctypes.c_int64()
    for fn, args in ((L.chx_serve_open, (None, 64, 1, 256, 1, 0, ctypes.byref(sid))),
                     (L.chx_serve_submit, (None, 1, 0x1000, None)),
                     (L.chx_serve_submit_device, (None, 1, 0x1000, None)),
                     (L.chx_serve_wait_frame, (None, 1)),
                     (L.chx_serve_read, (None, 1, None)),
                     (L.chx_serve_close, (None, 1)),
                     (L.chx_serve_get_stats, (None, None)),
                     (L.chx_serve_reset_stats, (None,)),
                     (L.chx_serve_latencies, (None, None, None, 0, None)),
                     (L.chx_serve_destroy, (None,))):
        assert fn(*args) == chx.ERR_INVALID_ARGUMENT, fn.__name__


def test_blocks_to_full_host_matches_reference(golden):
    for N, F in ((96, 2), (64, 1)):
        nb = chx.num_blocks(N)
        blocks = (np.arange(F * nb * 2048, dtype=np.int64) * 37 - 100000).astype(np.int32)
        np.testing.assert_array_equal(chx.blocks_to_full(blocks, N),
                                      golden[f"reorgfull_N{N}_F{F}__full"])
    assert chx.lib().chx_blocks_to_full(1, 48, None, None) == chx.ERR_INVALID_ARGUMENT


def test_split_paired_feeds_matches_reference(golden):
    for F, E in ((4, 16), (5, 16), (6, 7), (1, 16)):
        buf = golden[f"split_F{F}_E{E}__in"].copy()
        assert chx.lib().chx_split_paired_feeds(F, E, buf.ctypes.data) == 0
        np.testing.assert_array_equal(buf, golden[f"split_F{F}_E{E}__out"], err_msg=f"F{F} E{E}")
    full = np.random.default_rng(3).integers(-2**31, 2**31, size=(3, 24, 24, 2), dtype=np.int32)
    np.testing.assert_array_equal(chx.split_paired_feeds(full, 7, 12),
                                  O.split_paired_feeds(full, 7, 12))
    L = chx.lib()
    assert L.chx_split_paired_feeds(2, 16, None) == chx.ERR_INVALID_ARGUMENT
    assert L.chx_split_paired_feeds(-2, 16, full.ctypes.data) == chx.ERR_INVALID_ARGUMENT
    assert L.chx_split_paired_feeds(2, -1, full.ctypes.data) == chx.ERR_INVALID_ARGUMENT
    before = full.copy()
    assert L.chx_split_paired_feeds(1, 12, full.ctypes.data) == 0  # F/2 = 0: nothing
    np.testing.assert_array_equal(full, before)


def test_feed16_to_triangle_matches_reference(golden):
    for F_out, F_in in ((4, 4), (6, 3), (2, 5), (3, 0)):
        src = golden[f"tri16_out{F_out}_in{F_in}__in"]
        out = np.full(F_out * 272, 0x5A5A5A5A, np.int32)
        assert chx.lib().chx_feed16_to_triangle(F_out, F_in, src.ctypes.data,
                                                out.ctypes.data) == 0
        np.testing.assert_array_equal(out, golden[f"tri16_out{F_out}_in{F_in}__out"])
    full16 = np.random.default_rng(4).integers(-2**31, 2**31, size=(5, 16, 16, 2), dtype=np.int32)
    for F_out in (5, 2, 9):
        np.testing.assert_array_equal(chx.feed16_to_triangle(full16, F_out),
                                      O.feed16_to_triangle(full16, F_out))
    L = chx.lib()
    assert L.chx_feed16_to_triangle(1, 1, None, full16.ctypes.data) == chx.ERR_INVALID_ARGUMENT
    assert L.chx_feed16_to_triangle(1, 1, full16.ctypes.data, None) == chx.ERR_INVALID_ARGUMENT
    assert L.chx_feed16_to_triangle(-1, 1, full16.ctypes.data,
                                    full16.ctypes.data) == chx.ERR_INVALID_ARGUMENT


_SKIP_PROBE = r"""
#include "chx_internal.h"
#include <cstdio>
#include <set>
// Every item of one channel, every wave: the blocks tile_skip_mask leaves
// out must be blocks the flush never stores (bx < by, or bx >= nb), with
// the wave layouts the GEMM tiles use (by0 = 4 ti + gr0, bx0 = 4 ct + 2 cw).
int main() {
  std::set<int> masks[2];
  for (int nt = 1; nt <= 40; nt++) {
    const int nb = 4 * nt, Np = 128 * nt;
    for (int small = 0; small < 2; small++) {
      const int items = small ? (int)chx::small_tiles(nt) : (int)chx::paired_tiles(nt);
      for (int tb = 0; tb < items; tb++) {
        int ti_a, ti_b, ct;
        if (small) {
          int r = 0, t = tb;
          while (r < nt - 1 && t >= nt - r) { t -= nt - r; r++; }
          ti_a = ti_b = r;
          ct = r + t;
        } else {
          chx::tile_coords(tb, nt, ti_a, ti_b, ct);
        }
        const bool pair = !small && ti_b >= 0 && ti_b != ti_a;
        for (int w = 0; w < 4; w++) {
          const int m = chx::tile_skip_mask(pair, small, ti_a, ti_b, ct, Np, w);
          masks[small].insert(m);
          const int ti = pair && w >= 2 ? ti_b : ti_a;
          const int gr0 = small ? 2 * (w >> 1) : 0;
          const int cw = pair || small ? w & 1 : w;
          const int ng = small ? 2 : 4;
          for (int i = 0; i < ng; i++)
            for (int j = 0; j < 2; j++) {
              const int by = 4 * ti + gr0 + i, bx = 4 * ct + 2 * cw + j;
              const bool stored = bx >= by && bx < nb;
              if (((m >> (2 * i + j)) & 1) && stored) {
                printf("BAD nt=%d small=%d tb=%d w=%d block=(%d,%d) mask=%x\n", nt, small, tb, w,
                       by, bx, m);
                return 1;
              }
            }
          if (m >> (small ? 4 : 8)) { printf("BAD mask %x\n", m); return 1; }
        }
      }
    }
  }
  for (int s = 0; s < 2; s++) {
    printf("masks%d", s);
    for (int m : masks[s]) printf(" %x", m);
    printf("\n");
  }
  return 0;
}
"""


def test_tile_skip_masks_leave_out_only_unstored_blocks(tmp_path):
    # The GEMMs' per-wave skip masks (chx_internal.h, compiled here for the
    # host) never leave out a block that a flush stores, over every item of
    # arrays of 1-40 tiles of 128 elements; and the masks that occur are the
    # ones the kernels have variants for (0xF4, 0x40 and, for a wave past the
    # array, 0xFF; 128 x 128 tiles: 0x04 and 0x0F) or no skip at all.
    import shutil
    gxx = shutil.which("g++")
    if gxx is None or not os.path.exists("/opt/rocm/include/hip/hip_runtime.h"):
        pytest.skip("no host C++ compiler or HIP headers")
    src = tmp_path / "skip_probe.cpp"
    src.write_text(_SKIP_PROBE)
    exe = tmp_path / "skip_probe"
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    subprocess.run([gxx, "-std=c++17", "-O1", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
                    "-I" + os.path.join(root, "include"),
                    "-I" + os.path.join(root, "chime-x_amd", "csrc"), str(src), "-o", str(exe)],
                   check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True)
    assert out.returncode == 0, out.stdout
    lines = dict(l.split(" ", 1) for l in out.stdout.strip().splitlines())
    assert set(lines["masks0"].split()) == {"0", "f4", "40", "ff"}
    assert set(lines["masks1"].split()) == {"0", "4", "f"}
