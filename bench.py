"""ORB benchmark on one MI355X.

Workload (one step): the 1920x1080 uint8 synthetic pair (80 shapes, seed 1;
B = warped A), 8 pyramid levels at scale 1.2, FAST-9 threshold 20, edge 31,
Harris k 0.04 block 7, 7x7 Gaussian sigma 2, 256-bit rBRIEF, <= 10000
keypoints per image: detect + orient + describe both images in one batched
call, then the exact brute-force Hamming match of A's descriptors against B's
with the VALU popcount matcher, counts read on the device.  Inputs are
resident in HBM before the timed region.  value = frames per second of that
whole step (2 frames per step, matching time included).

Prints one JSON line.  `python bench.py [--steps K] [--warmup W] [--gpus 1]`.
"""
import argparse
import json
import math
import os
import pathlib
import sys
import time

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "cuda-orb_amd"))
sys.path.insert(0, str(ROOT))

W, H, NSHAPES, SEED, CAP, LEVELS, SCALE = 1920, 1080, 80, 1, 10000, 8, 1.2
CENTROID_TEMPERATURE = 4.0   # centroid-mode weights exp(-Hamming distance / 4) of the mass rule
HBM_PEAK_GBS = 8000.0
EVENT_EVERY = 8   # timed steps per event-timed launch of the dominant kernel
GRAPH_STEPS = 8   # pair steps recorded into one graph for the multi-step replay
METRIC = "ORB frames/s @1080p; exact Hamming Gpairs/s; learned-match speedup @99% recall"


def geometry(w, h, nlevels=LEVELS, scale=SCALE):
    """Level sizes and padded bytes of the flat pyramid (the library's rule)."""
    lv = []
    sc = 1.0
    for l in range(nlevels):
        if l:
            sc *= scale
        wl, hl = int(math.floor(w / sc + 0.5)), int(math.floor(h / sc + 0.5))
        if min(wl, hl) < 16:
            break
        p = wl if wl % 128 == 0 else wl - wl % 128 + 128
        lv.append((wl, hl, p))
    return lv


def stage_bytes(stage, frames, npts):
    """Algorithmic HBM bytes of one launch of a stage over `frames` ((w,h) list) and npts keypoints."""
    tot = 0
    for (w, h) in frames:
        lv = geometry(w, h)
        padded = sum(hl * p for (_, hl, p) in lv)
        p0 = lv[0][2]
        if stage == "pyramid":          # read the input once, write raw + blurred levels
            tot += h * p0 + 2 * padded
        elif stage == "detect":         # read every raw level once
            tot += padded
        elif stage == "nms":            # read the level-0 corner bitmap (1 bit per pixel) and clear it
            tot += 2 * h * p0 // 8
    if stage == "describe":             # per keypoint: 749-px raw disc, 37x37 blurred patch, point, descriptor
        tot += npts * (749 + 37 * 37 + 28 + 32)
    return tot


COUNTERS_FILE = ROOT / "profiles" / "orb_1080p_pair_hbm_counters.md"


def stage_kernel(stage, strip_height):
    """rocprofv3 kernel name of a stage's launch."""
    return {"pyramid": f"void k_pyramid<{strip_height}>", "detect": "k_detect", "nms": "k_nms", "select": "k_select",
            "describe": "void k_describe<true>", "match": "void k_match_pop<false>",
            "match_merge": "k_match_merge"}.get(stage)


def measured_traffic(kernel):
    """HBM bytes per dispatch of `kernel` from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary of this
    workload (rows: | kernel | dispatches | FETCH_SIZE KiB | WRITE_SIZE KiB | bytes |), or None."""
    try:
        text = COUNTERS_FILE.read_text()
    except OSError:
        return None
    for line in text.splitlines():
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) >= 5 and cells[0] == kernel:
            try:
                return int(float(cells[4]))
            except ValueError:
                return None
    return None


def dense_match(ctx, torch, dev, reps=3):
    """65536 queries against 262144 random 256-bit descriptors (queries = database rows with
    Binomial(256, 0.05) bit flips): the matrix-core matcher (FP4 operands, one code per bit, FP32
    accumulate) and the popcount matcher, timed with events on the library's stream, outputs
    compared element for element.  int_tops_equiv counts 512 operations per (query, entry) pair
    (256 multiply-adds), the contraction the matrix cores perform."""
    import orb_amd
    rng = np.random.default_rng(2)
    ndb, nq = 262144, 65536
    db = rng.integers(0, 256, (ndb, 32), dtype=np.uint8)
    src = rng.integers(0, ndb, nq)
    flips = np.packbits(rng.random((nq, 256)) < 0.05, axis=1, bitorder="little")
    q = db[src] ^ flips
    dq, ddb = torch.from_numpy(q).to(dev), torch.from_numpy(db).to(dev)
    res = {}
    outs = {}
    for name, method in (("popcount", orb_amd.ORB_MATCH_POPCOUNT), ("mfma", orb_amd.ORB_MATCH_MFMA)):
        out = ctx.match_top2(dq, ddb, method=method)
        torch.cuda.synchronize()
        ctx.profile_enable((1 << orb_amd.STAGES.index("match")) | (1 << orb_amd.STAGES.index("match_merge")))
        for _ in range(reps):
            ctx.match_top2(dq, ddb, method=method, out=out)
        prof = ctx.profile_read()
        ctx.profile_enable(0)
        ms_main = prof["match"][0] / prof["match"][1]
        mg = prof["match_merge"]   # the popcount matcher merges inside its one launch
        ms = ms_main + (mg[0] / mg[1] if mg[1] else 0.0)
        outs[name] = [o.cpu().numpy() for o in out]
        res[name] = {"ms": round(ms, 4), "gpairs_per_s": round(nq * ndb / ms / 1e6, 1),
                     "kernel_ms": round(ms_main, 4),
                     "int_tops_equiv": round(512.0 * nq * ndb / (ms_main * 1e-3) / 1e12, 1)}
    identical = all(np.array_equal(a, b) for a, b in zip(outs["popcount"], outs["mfma"]))
    recall_src = float((outs["mfma"][0] == src).mean())
    return {"queries": nq, "database": ndb, "popcount": res["popcount"], "mfma": res["mfma"],
            "mfma_speedup": round(res["popcount"]["ms"] / res["mfma"]["ms"], 3),
            "identical": identical, "idx1_is_source": round(recall_src, 5),
            "mfma_frac_of_int8_peak": round(res["mfma"]["int_tops_equiv"] / 5000.0, 4),
            "mfma_frac_of_fp4_peak": round(res["mfma"]["int_tops_equiv"] / 10100.0, 4)}


def batch_frames(ctx, torch, dev, reps=5):
    """64 frames of sizes drawn from {640x480, 1280x720, 1920x1080, 3840x2160} (shape count scaled to
    area), detected and described in one launch per stage, each matched to a warped partner (the
    partners are the second half of a 128-frame batch; the 64 matches in one launch); <= 5000
    keypoints per frame."""
    import orb_amd
    rng = np.random.default_rng(3)
    sizes = [(640, 480), (1280, 720), (1920, 1080), (3840, 2160)]
    pick = [sizes[i] for i in rng.integers(0, 4, 64)]
    imgs = []
    for i, (w, h) in enumerate(pick):
        a = orb_amd.synth_image(300 + i, w, h, max(1, int(40 * w * h / 307200)))
        b, _ = orb_amd.synth_warp(300 + i, a, w, h)
        imgs.append((a, b))
    frames = [torch.from_numpy(a).to(dev) for a, _ in imgs] + [torch.from_numpy(b).to(dev) for _, b in imgs]
    ws = [w for w, _ in pick] * 2
    hs = [h for _, h in pick] * 2
    cap = 5000
    ctx.init(noctaves=LEVELS, edge_threshold=31, wta_k=2, score_type=0, patch_size=31, fast_threshold=20,
             max_pts=cap, scale_factor=SCALE)
    pts = ctx.alloc_points(cap, frames=128)
    desc = ctx.alloc_desc(cap, frames=128)
    counts = torch.zeros(128, dtype=torch.int32, device=dev)
    pl = [pts[i] for i in range(128)]
    dl = [desc[i] for i in range(128)]

    def step():
        ctx.detect_and_compute_batch(frames, ws, hs, pl, dl, counts)
        ctx.match_batch(pl[:64], dl[:64], counts[:64], cap, dl[64:], counts[64:], cap)

    step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        step()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / reps
    kp = counts.cpu().numpy()
    # each stage timed on its own, after the timed batches (the match stage of a batch is one launch)
    stage_ms = {}
    for si, name in enumerate(orb_amd.STAGES):
        ctx.profile_enable(1 << si)
        for _ in range(2):
            step()
        torch.cuda.synchronize()
        v = ctx.profile_read()[name]
        if v[1]:
            stage_ms[name] = round(v[0] / v[1], 4)
    ctx.profile_enable(0)
    ctx.init(noctaves=LEVELS, edge_threshold=31, wta_k=2, score_type=0, patch_size=31, fast_threshold=20,
             max_pts=CAP, scale_factor=SCALE)
    return {"frames": 128, "pairs": 64, "ms_per_batch": round(dt * 1e3, 3),
            "frames_per_s": round(128 / dt, 1), "mean_keypoints": round(float(kp.mean()), 1),
            "megapixels": round(sum(w * h for w, h in zip(ws, hs)) / 1e6, 2), "stage_ms": stage_ms}


def synth_frames(seeds, w, h, nshapes, warp_seed=None, threads=16):
    """Seeded synthetic frames (host threads; the generator releases the GIL)."""
    import orb_amd
    from concurrent.futures import ThreadPoolExecutor

    def make(s):
        a = orb_amd.synth_image(s, w, h, nshapes)
        if warp_seed is None:
            return a
        return orb_amd.synth_warp(warp_seed * 1000003 + s, a, w, h)[0]

    with ThreadPoolExecutor(threads) as ex:
        return list(ex.map(make, seeds))


def extract(ctx, torch, dev, seeds, w, h, nshapes, cap, warp_seed=None, limit=None, batch=64):
    """Detect + describe seeded frames in batches of 64; descriptors of all frames, in order."""
    out, total = [], 0
    for i in range(0, len(seeds), batch):
        imgs = synth_frames(seeds[i:i + batch], w, h, nshapes, warp_seed)
        n = len(imgs)
        frames = [torch.from_numpy(a).to(dev) for a in imgs]
        pts = ctx.alloc_points(cap, frames=n)
        desc = ctx.alloc_desc(cap, frames=n)
        counts = torch.zeros(n, dtype=torch.int32, device=dev)
        ctx.detect_and_compute_batch(frames, [w] * n, [h] * n, [pts[j] for j in range(n)],
                                     [desc[j] for j in range(n)], counts)
        cnt = counts.cpu().numpy()
        for j in range(n):
            out.append(desc[j, :cnt[j]].clone())
            total += int(cnt[j])
        if limit is not None and total >= limit:
            break
    d = torch.cat(out)
    return d[:limit] if limit is not None else d


def learned_match(ctx, torch, dev, log=None):
    """Learned-bucket matching: a 1,048,576-descriptor database extracted by the pipeline from 2000
    seeded 1280x720 frames, 65536 queries from warped copies; MLP 256->512->512->1024 buckets (BF16,
    FP32 accumulate, GELU, plus the centroid term at scale 0.125) trained on exact matches of other
    warped copies; top-8 bucket probe, exact fallback when max softmax < 0.3; recall@1 against exact
    brute force, and the trained network against nearest-centroid probing at equal probe counts."""
    import orb_amd
    w, h, nshapes, per = 1280, 720, 120, 525
    ndb, nq = 1 << 20, 65536
    t_all = time.perf_counter()
    ctx.init(noctaves=LEVELS, edge_threshold=31, wta_k=2, score_type=0, patch_size=31, fast_threshold=20,
             max_pts=per, scale_factor=SCALE)
    seeds = list(range(10000, 12000))
    note = lambda msg: print(f"[learned] {msg} ({time.perf_counter() - t_all:.1f} s)", file=sys.stderr, flush=True)
    db = extract(ctx, torch, dev, seeds, w, h, nshapes, per, limit=ndb)
    note(f"database {db.shape[0]} descriptors")
    more = 12000
    while db.shape[0] < ndb:   # frames with fewer than `per` keypoints: extend the seed list
        extra = extract(ctx, torch, dev, list(range(more, more + 64)), w, h, nshapes, per)
        db = torch.cat([db, extra])[:ndb]
        more += 64
    t_db = time.perf_counter()
    # training queries: four warped copies of every database frame (warp seeds 1, 3, 5, 7)
    train_q = torch.cat([extract(ctx, torch, dev, seeds, w, h, nshapes, per, warp_seed=ws) for ws in (1, 3, 5, 7)])
    train_q = train_q[: (train_q.shape[0] // 4096) * 4096]
    note(f"training queries {train_q.shape[0]}")
    test_q = extract(ctx, torch, dev, seeds, w, h, nshapes, per, warp_seed=2, limit=nq)
    t_q = time.perf_counter()
    note(f"test queries {test_q.shape[0]}")
    ctx.init(noctaves=LEVELS, edge_threshold=31, wta_k=2, score_type=0, patch_size=31, fast_threshold=20,
             max_pts=CAP, scale_factor=SCALE)
    # exact labels of the training queries
    labels = ctx.match_top2(train_q, db, method=orb_amd.ORB_MATCH_MFMA)[0]
    torch.cuda.synchronize()
    lx = orb_amd.LearnedIndex(ctx, nbuckets=1024, hidden=512, nprobe=8, fallback=0.3, seed=4, centroid_scale=0.125)
    t0 = time.perf_counter()
    lx.build(db, iters=8)
    t_build = time.perf_counter() - t0
    off, _, _ = lx.bucket_info()
    sizes = np.diff(off)
    note("buckets built")

    def timed(fn, reps=3):
        fn()
        torch.cuda.synchronize()
        best = 1e30
        for _ in range(reps):
            t0 = time.perf_counter()
            r = fn()
            torch.cuda.synchronize()
            best = min(best, time.perf_counter() - t0)
        return r, best * 1e3

    exact, ms_exact = timed(lambda: ctx.match_top2(test_q, db, method=orb_amd.ORB_MATCH_MFMA))
    d_ex = exact[1].cpu().numpy()
    rec = lambda o: float((o[1].cpu().numpy() == d_ex).mean())
    rec_idx = lambda o: float((o[0].cpu().numpy() == exact[0].cpu().numpy()).mean())
    # the network's forward pass alone (descriptor expansion + three BF16 products with the centroid
    # term, logits rounded to FP16 and written as FP32): 2 (256*512 + 512*512 + 512*1024) FLOP per query
    # on the matrix cores
    _, ms_fwd = timed(lambda: lx.logits(test_q))
    fwd_flop = 2.0 * (256 * 512 + 512 * 512 + 512 * 1024) * test_q.shape[0]
    # untrained, the network (last layer zero + centroid term) ranks buckets as the centroids do
    lx.set_probe(8, 0.0)
    (u0, _), _ = timed(lambda: lx.query(test_q, mode=orb_amd.ORB_LINDEX_MLP), reps=1)
    (c0, _), _ = timed(lambda: lx.query(test_q, mode=orb_amd.ORB_LINDEX_CENTROIDS), reps=1)
    untrained_same = all(bool(torch.equal(a_, b_)) for a_, b_ in zip(u0, c0))
    t0 = time.perf_counter()
    losses = lx.train(train_q, labels, epochs=3, batch=4096, lr=1e-3)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    note(f"trained, losses {losses}")
    # the trained network against nearest-centroid probing, same probe counts, no fallback
    versus = []
    for npb in (4, 8, 16, 20, 24, 28, 32):
        lx.set_probe(npb, 0.0)
        (o, _), ms = timed(lambda: lx.query(test_q, mode=orb_amd.ORB_LINDEX_MLP), reps=2)
        (oc, _), msc = timed(lambda: lx.query(test_q, mode=orb_amd.ORB_LINDEX_CENTROIDS), reps=2)
        versus.append({"nprobe": npb, "network_recall_at_1": round(rec(o), 5), "network_recall_idx": round(rec_idx(o), 5),
                       "network_ms": round(ms, 3), "centroid_recall_at_1": round(rec(oc), 5),
                       "centroid_recall_idx": round(rec_idx(oc), 5), "centroid_ms": round(msc, 3)})
    # the learned matcher as specified: top-8 probe, exact fallback below max softmax 0.3
    lx.set_probe(8, 0.3)
    (lq, nfb), ms_learned = timed(lambda: lx.query(test_q, mode=orb_amd.ORB_LINDEX_MLP))
    sweep = []
    for npb in (8, 16, 20, 24, 28, 32):
        for fb in ((0.0, 0.3, 0.5) if npb in (8, 16, 32) else (0.0,)):
            lx.set_probe(npb, fb)
            (o, nf), ms = timed(lambda: lx.query(test_q, mode=orb_amd.ORB_LINDEX_MLP), reps=2)
            sweep.append({"nprobe": npb, "fallback": fb, "ms": round(ms, 3), "recall_at_1": round(rec(o), 5),
                          "recall_idx": round(rec_idx(o), 5), "fallback_queries": int(nf),
                          "speedup": round(ms_exact / ms, 3)})
    # per-query probe counts: of the nprobe best buckets, the best-ranked prefix whose probability
    # mass reaches `mass` (network: its softmax; centroids: exp(-distance / 4), the temperature
    # that fits the training queries' nearest-neighbour buckets best in likelihood)
    mass_rows = []
    for npb in (32, 40, 48):
        for mass in (0.97, 0.98, 0.985, 0.99, 0.995):
            lx.set_probe(npb, 0.0)
            lx.set_mass(mass, CENTROID_TEMPERATURE)
            (o, _), ms = timed(lambda: lx.query(test_q, mode=orb_amd.ORB_LINDEX_MLP), reps=2)
            (oc, _), msc = timed(lambda: lx.query(test_q, mode=orb_amd.ORB_LINDEX_CENTROIDS), reps=2)
            pr = (lx.probes(test_q)[0] >= 0).sum(1).float().mean().item()
            prc = (lx.probes(test_q, mode=orb_amd.ORB_LINDEX_CENTROIDS)[0] >= 0).sum(1).float().mean().item()
            mass_rows.append({"nprobe": npb, "mass": mass, "network_ms": round(ms, 3),
                              "network_recall_at_1": round(rec(o), 5), "network_recall_idx": round(rec_idx(o), 5),
                              "network_mean_probes": round(pr, 2), "centroid_ms": round(msc, 3),
                              "centroid_recall_at_1": round(rec(oc), 5), "centroid_recall_idx": round(rec_idx(oc), 5),
                              "centroid_mean_probes": round(prc, 2)})
    lx.set_mass(1.0, CENTROID_TEMPERATURE)
    # the fastest settings whose nearest-neighbour index (not only its distance) matches exact search
    # for >= 99% of the queries, over the fixed probe counts and the mass rule
    ok = [dict(r, mass=1.0) for r in sweep if r["recall_idx"] >= 0.99]
    ok += [{"nprobe": r["nprobe"], "mass": r["mass"], "fallback": 0.0, "ms": r["network_ms"],
            "recall_at_1": r["network_recall_at_1"], "recall_idx": r["network_recall_idx"],
            "mean_probes": r["network_mean_probes"], "fallback_queries": 0,
            "speedup": round(ms_exact / r["network_ms"], 3)} for r in mass_rows if r["network_recall_idx"] >= 0.99]
    best = min(ok, key=lambda r: r["ms"]) if ok else None
    ok_c = [{"nprobe": r["nprobe"], "mass": 1.0, "ms": r["centroid_ms"], "recall_at_1": r["centroid_recall_at_1"],
             "recall_idx": r["centroid_recall_idx"], "speedup": round(ms_exact / r["centroid_ms"], 3)}
            for r in versus if r["centroid_recall_idx"] >= 0.99]
    ok_c += [{"nprobe": r["nprobe"], "mass": r["mass"], "ms": r["centroid_ms"], "recall_at_1": r["centroid_recall_at_1"],
              "recall_idx": r["centroid_recall_idx"], "mean_probes": r["centroid_mean_probes"],
              "speedup": round(ms_exact / r["centroid_ms"], 3)} for r in mass_rows if r["centroid_recall_idx"] >= 0.99]
    best_c = min(ok_c, key=lambda r: r["ms"]) if ok_c else None
    # the two winners once more, back to back (network, then centroids): 1 + 3 query batches each, the
    # last launches of the run, so a kernel trace of it ends with them
    for row, mode in ((best, orb_amd.ORB_LINDEX_MLP), (best_c, orb_amd.ORB_LINDEX_CENTROIDS)):
        if row is None:
            continue
        lx.set_probe(row["nprobe"], row.get("fallback", 0.0))
        lx.set_mass(row.get("mass", 1.0), CENTROID_TEMPERATURE)
        _, row["replayed_ms"] = timed(lambda: lx.query(test_q, mode=mode), reps=3)
        row["replayed_ms"] = round(row["replayed_ms"], 3)
    lx.set_mass(1.0, CENTROID_TEMPERATURE)
    res = {
        "database": int(db.shape[0]), "queries": int(test_q.shape[0]), "train_queries": int(train_q.shape[0]),
        "buckets": 1024, "centroid_scale": 0.125,
        "bucket_size": {"min": int(sizes.min()), "median": int(np.median(sizes)), "max": int(sizes.max())},
        "exact_mfma_ms": round(ms_exact, 3),
        "forward": {"ms": round(ms_fwd, 4), "gflop": round(fwd_flop / 1e9, 1),
                    "tflops": round(fwd_flop / (ms_fwd * 1e-3) / 1e12, 1),
                    "frac_of_bf16_peak": round(fwd_flop / (ms_fwd * 1e-3) / 2.5e15, 4)},
        "top8_fallback_0.3": {"ms": round(ms_learned, 3), "recall_at_1": round(rec(lq), 5),
                              "recall_idx": round(rec_idx(lq), 5), "fallback_queries": int(nfb),
                              "speedup_vs_exact": round(ms_exact / ms_learned, 3)},
        "untrained_equals_centroid_probe": untrained_same,
        "network_vs_centroids": versus,
        "sweep": sweep, "mass_rule": mass_rows, "centroid_temperature": CENTROID_TEMPERATURE,
        "fastest_at_99_recall": best, "centroid_fastest_at_99_recall": best_c,
        "train_loss_per_epoch": [round(float(x), 4) for x in losses],
        "build_s": round(t_build, 2), "train_s": round(t_train, 2),
        "data_s": round(t_q - t_all, 1),
    }
    lx.close()
    return res


def cpu_baseline(a, b, min_seconds=8.0):
    """The C oracle on the host doing the same step (detect+describe both images, reference-rule match):
    once on one thread and once on every core this process may use (OpenMP over pyramid and
    detection rows, keypoints and match queries; identical outputs), each repeated for about
    min_seconds.  `value` is the all-core rate."""
    from oracle import oracle as O
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)

    def leg(threads):
        O.set_threads(threads)
        reps = 0
        t0 = time.perf_counter()
        while True:
            pa, da, _ = O.detect_and_compute(a[:, :W].copy(), noctaves=LEVELS, scale=SCALE, max_pts=CAP)
            pb, db, _ = O.detect_and_compute(b[:, :W].copy(), noctaves=LEVELS, scale=SCALE, max_pts=CAP)
            O.match(da, db)
            reps += 1
            dt = time.perf_counter() - t0
            if dt >= min_seconds or reps >= 400:
                return 2.0 * reps / dt, reps, dt, len(pa), len(pb)

    try:
        v1, r1, t1, na, nb = leg(1)
        vn, rn, tn, _, _ = leg(cores)
    finally:
        O.set_threads(1)
    return {"value": round(vn, 4), "unit": "frames/s", "cores": cores, "kind": "port",
            "single_thread_value": round(v1, 4),
            "sample": f"the 1920x1080 pair step (oracle detect+describe of both images, reference-rule match "
                      f"{na}x{nb}): {rn} repeats in {tn:.1f} s on {cores} threads, {r1} repeats in {t1:.1f} s on 1"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the dense-matching, batched-frames and learned-matching measurements")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-matching measurement")
    ap.add_argument("--learned-only", action="store_true",
                    help="only the learned-matching measurement (its JSON object), e.g. under a profiler")
    ap.add_argument("--batch-only", action="store_true",
                    help="only the 128-frame batch measurement (its JSON object), e.g. under a profiler")
    ap.add_argument("--dense-only", action="store_true",
                    help="only the dense-matching measurement (its JSON object), e.g. under a profiler")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run a 64 MiB elementwise copy kernel first (PMC byte-counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported (frames would shard across "
              f"GPUs with no exchange, the matcher's database by rows)", file=sys.stderr)
        sys.exit(2)

    import torch
    import orb_amd

    dev = torch.device("cuda:0")
    if args.copy_calibration:
        x = torch.randint(0, 255, (64 << 20,), dtype=torch.uint8, device=dev)
        y = torch.empty_like(x)
        for _ in range(3):
            torch.bitwise_not(x, out=y)
        torch.cuda.synchronize()
        del x, y
    if args.learned_only or args.batch_only or args.dense_only:
        ctx = orb_amd.Context(0).init(noctaves=LEVELS, max_pts=CAP, scale_factor=SCALE)
        if args.learned_only:
            out = {"learned": learned_match(ctx, torch, dev)}
        elif args.batch_only:
            out = {"batch64": batch_frames(ctx, torch, dev, reps=max(args.steps, 1))}
        else:
            out = {"dense_match": dense_match(ctx, torch, dev, reps=max(args.steps, 1))}
        print(json.dumps(out), flush=True)
        ctx.close()
        return
    a = orb_amd.synth_image(SEED, W, H, NSHAPES)
    b, _ = orb_amd.synth_warp(SEED, a, W, H)
    imgs = [torch.from_numpy(a).to(dev), torch.from_numpy(b).to(dev)]
    ctx = orb_amd.Context(0).init(noctaves=LEVELS, edge_threshold=31, wta_k=2, score_type=0, patch_size=31,
                                  fast_threshold=20, max_pts=CAP, scale_factor=SCALE)
    pts = ctx.alloc_points(CAP, frames=2)
    desc = ctx.alloc_desc(CAP, frames=2)
    counts = torch.zeros(2, dtype=torch.int32, device=dev)

    def step():
        ctx.detect_and_compute_batch(imgs, [W, W], [H, H], [pts[0], pts[1]], [desc[0], desc[1]], counts)
        ctx.match_dcount(pts[0], desc[0], counts[0:1], CAP, desc[1], counts[1:2], CAP)

    # warm-up, then every stage timed on its own (one event pair per step, so the host keeps
    # ahead of the GPU and no launch gap falls inside a pair), to find the dominant kernel
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    stage_ms = {}
    for si, name in enumerate(orb_amd.STAGES):
        ctx.profile_enable(1 << si)
        for _ in range(max(args.warmup, 10)):
            step()
        torch.cuda.synchronize()
        v = ctx.profile_read()[name]
        stage_ms[name] = v[0] / v[1] if v[1] else 0.0
    ctx.profile_enable(0)
    dominant = max(stage_ms, key=stage_ms.get)
    n1, n2 = (int(v) for v in counts.cpu())

    # timed region: the dominant kernel carries an event pair on every EVENT_EVERY-th step (an
    # event pair costs the stream a few microseconds, so events on every step would slow it)
    dom_mask = 1 << orb_amd.STAGES.index(dominant)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        sampled = i % EVENT_EVERY == 0
        if sampled:
            ctx.profile_enable(dom_mask)
        step()
        if sampled:
            ctx.profile_enable(0)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    dom = ctx.profile_read()[dominant]
    ms_step = (t1 - t0) * 1e3 / args.steps
    dom_ms = dom[0] / max(dom[1], 1)

    # the same step recorded once into a HIP graph and replayed (kernel-only protocol)
    ctx.capture_begin()
    step()
    graph = ctx.capture_end()
    graph.launch()
    torch.cuda.synchronize()
    g0 = time.perf_counter()
    for _ in range(args.steps):
        graph.launch()
    torch.cuda.synchronize()
    g1 = time.perf_counter()
    graph.close()
    graph_ms = (g1 - g0) * 1e3 / args.steps
    # GRAPH_STEPS consecutive steps recorded into one graph: one hipGraphLaunch (its host cost is about
    # a step's GPU time) per GRAPH_STEPS steps
    ctx.capture_begin()
    for _ in range(GRAPH_STEPS):
        step()
    graph = ctx.capture_end()
    graph.launch()
    torch.cuda.synchronize()
    reps = max(args.steps // GRAPH_STEPS, 1)
    g0 = time.perf_counter()
    for _ in range(reps):
        graph.launch()
    torch.cuda.synchronize()
    g1 = time.perf_counter()
    graph.close()
    graph_multi_ms = (g1 - g0) * 1e3 / (reps * GRAPH_STEPS)

    if dominant in ("match", "match_merge"):
        pairs = n1 * n2
        achieved = 2 * 256 * pairs / (dom_ms * 1e-3) / 1e12
        roof = {"bound": "mfma", "achieved": round(achieved, 3), "peak": 5000.0, "unit": "TFLOP/s",
                "frac": round(achieved / 5000.0, 5), "traffic": None,
                "note": "popcount matcher: 512 int ops/pair (+/-1 dot equivalent) against the INT8 MFMA peak"}
    else:
        nbytes = stage_bytes(dominant, [(W, H), (W, H)], n1 + n2)
        achieved = nbytes / (dom_ms * 1e-3) / 1e9
        roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None, "kernel": dominant,
                "bytes_per_launch": nbytes}
    strip_h = orb_amd.PYRAMID_STRIP_HEIGHTS[ctx.tune_get(orb_amd.TUNE_PYRAMID)[0]]
    roof["kernel"] = stage_kernel(dominant, strip_h)
    roof["avg_launch_ms"] = round(dom_ms, 5)
    roof["timed_launches"] = int(dom[1])
    roof["traffic"] = measured_traffic(roof["kernel"])

    out = {
        "metric": METRIC,
        "value": round(2 * args.steps / (t1 - t0), 2),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic (seeded 1920x1080 pair, 80 shapes + value noise, warped partner)",
        "config": {"workload": "1920x1080 pair: detect+describe both (8 levels @1.2, FAST 20, Harris, "
                               "<=10000 kp) + exact popcount BF match",
                   "frames_per_step": 2, "keypoints": [n1, n2], "match_pairs": n1 * n2},
        "graph_replay": {"ms_per_step": round(graph_ms, 4), "frames_per_s": round(2e3 / graph_ms, 2),
                         f"ms_per_step_{GRAPH_STEPS}_steps_per_graph": round(graph_multi_ms, 4)},
        "stage_ms": {k: round(v, 5) for k, v in stage_ms.items()},
        "match_gpairs_per_s": round(n1 * n2 / (stage_ms["match"] + stage_ms["match_merge"]) / 1e6, 2)
        if stage_ms["match"] > 0 else None,
        "roofline": roof,
        "autotune": {"pyramid_strip_height": orb_amd.PYRAMID_STRIP_HEIGHTS[ctx.tune_get(orb_amd.TUNE_PYRAMID)[0]],
                     "pyramid_ms_per_variant": dict(zip(orb_amd.PYRAMID_STRIP_HEIGHTS,
                                                        [round(t, 5) for t in ctx.tune_get(orb_amd.TUNE_PYRAMID)[1]]))},
    }
    if not args.no_extras:
        out["dense_match"] = dense_match(ctx, torch, dev)
        out["batch64"] = batch_frames(ctx, torch, dev)
        if not args.no_learned:
            out["learned"] = learned_match(ctx, torch, dev)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(a, b)
    ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
