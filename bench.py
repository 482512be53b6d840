"""Benchmark of the bit-exact bisulfite ksw_extend2 path on one MI355X.

Workload (default cfg1): a 50 Mb synthetic bisulfite genome, 1M paired-end
2x150 read pairs simulated on it, and every ksw_extend2 call mem_chain2aln
(bsbolt/External/BWA/bwamem.c:636-790) issues for them -- left/right flanks of
the seeds, right-flank h0 = left score, w<<1 retries -- built in rounds on the
GPU.  A step is one pass of the extension kernels over that whole batch, with
inputs and outputs resident in device memory.

Prints ONE JSON line: extensions/s (value), GCUPS, the integer-VALU roofline
of the extension kernels, their HBM figure, and the CPU baseline: the
reference's own ksw_extend2 (compiled from bsbolt/External/BWA/ksw.c) run over
the same batch on the host cores, with its outputs compared to the GPU's.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# MI355X (gfx950): 256 CUs, 2.4 GHz peak engine clock, 64 int32 lanes/CU/clock.
# One extension cell is 12 integer VALU operations in the reference inner loop
# (ksw.c:433-447: 1 add + 1 select for M, 2 max for H, 2 sub + 2 max each for E and F).
OPS_PER_CELL = 12
PEAK_INT32_TOPS = 256 * 64 * 2.4e9 / 1e12          # 39.3 T int32 ops/s
PEAK_PK16_TOPS = 2 * PEAK_INT32_TOPS               # 78.6 T ops/s with packed 16-bit VALU
PEAK_HBM_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def host_cpu():
    """The host cores the CPU baseline may use and how they were counted: the
    CPUs this process may run on (sched_getaffinity), capped by OMP_NUM_THREADS
    where the job environment sets its CPU share there."""
    aff = len(os.sched_getaffinity(0))
    share = os.environ.get("OMP_NUM_THREADS")
    threads = min(aff, int(share)) if share and share.isdigit() and int(share) > 0 else aff
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"threads": threads, "nproc": os.cpu_count(), "affinity": aff,
            "omp_num_threads": share, "model": model}


def cpu_fields(args):
    """Host description carried by every cpu_baseline object."""
    h = host_cpu()
    return {"cores": args.cpu_threads, "nproc": h["nproc"], "affinity_cpus": h["affinity"],
            "job_cpu_share": h["omp_num_threads"], "cpu_model": h["model"]}


def profiled_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary (calibrated against a copy of known size), or None."""
    import re
    path = os.path.join(ROOT, "profiles", f"hbm_traffic_{workload}.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        m = re.match(r"traffic_bytes_per_launch\s+(.+?)\s*=\s*(\d+)", line.strip())
        if m and m.group(1) == kernel:
            return int(m.group(2))
    return None


def time_steps(eng, ptrs, steps):
    eng.sync()
    eng.timer_start()
    for _ in range(steps):
        eng.extend_device(ptrs)
    ms = eng.timer_stop()
    eng.sync()
    return ms / steps


def learned_leg(args, eng, ptrs, outs, cells_d, plain_out, plain_cells, n, total_cells):
    """Times the learned-band form on the resident batch; its outputs must equal the plain form's."""
    from bsbolt_amd import bsx, band
    from bsbolt_amd.tasks import RESULT_FIELDS
    path = args.band_model or band.DEFAULT_MODEL
    if os.path.exists(path):
        params, src = band.load_model(path), os.path.relpath(path, ROOT)
    else:
        from bsbolt_amd.workload import CONFIGS, Workload
        log(f"no band model at {path}: training one on cfg2 (20% of the reads)")
        w2 = Workload(CONFIGS["cfg2"].scaled(0.2))
        b2, _, _, _ = w2.build_ext_batch(lambda t: eng.extend(t, cells=False))
        w2.close()
        X, y = band.features_and_labels(eng, b2)
        params, info = band.train(eng, X, y, steps=800, log=log)
        src = "trained in this run on cfg2 x 0.2"
    res = {"model": src}
    # the learned form with the trained and an untrained network, then the
    # 24-column frame for every task with no predictor (same frame kernels and
    # plain fallback)
    for name, p in (("trained", params), ("untrained", band.init_params(0)), ("fixed24", None)):
        if p is not None:
            eng.set_band_model(p)
            eng.set_band_policy(bsx.BAND_LEARNED)
        else:
            eng.set_band_policy(bsx.BAND_FIXED, 0)
        try:
            for _ in range(max(1, args.warmup)):
                eng.extend_device(ptrs)
            eng.sync()
            mism = int(sum(int((outs[k].numpy() != plain_out[k]).sum()) for k in RESULT_FIELDS))
            mism += int((cells_d.numpy() != plain_cells).sum())
            ms = time_steps(eng, ptrs, args.steps)
            st = eng.band_stats()
            # kernel breakdown with the launches in order (each kernel alone on the GPU)
            eng.set_side_streams(False)
            eng.set_kernel_timing(args.steps)
            ms_in_order = time_steps(eng, ptrs, args.steps)
            kms, kcalls = eng.kernel_times()
            eng.set_kernel_timing(0)
            eng.set_side_streams(True)
        finally:
            eng.set_band_policy(bsx.BAND_PLAIN)
        routed = {w: int(st[c, 0]) for c, w in enumerate(("f24", "f40", "f80", "f144", "plain"))}
        rejected = {w: int(st[c, 1]) for c, w in enumerate(("f24", "f40", "f80", "f144"))}
        narrow = int(st[:4, 0].sum())
        entry = {"value": n / (ms / 1e3), "unit": "extensions/s", "ms_per_step": ms,
                 "gcups": total_cells / (ms / 1e3) / 1e9, "mismatches_vs_plain": mism,
                 "routed": routed, "rejected": rejected,
                 "narrow_accepted": narrow - int(st[:4, 1].sum()),
                 "fallback_rate": (int(st[:4, 1].sum()) / narrow) if narrow else 0.0,
                 "ms_per_step_launches_in_order": ms_in_order,
                 "kernel_ms_per_step": {"predict": float(kms[15] / max(kcalls, 1)),
                                        "narrow": float(kms[11:15].sum() / max(kcalls, 1)),
                                        "plain": float(kms[:11].sum() / max(kcalls, 1))}}
        log(f"{'learned (' + name + ')' if p is not None else '24-column frame, no predictor'}: {ms:.3f} ms/step, "
            f"routed {routed}, rejected {rejected}, mismatches {mism}")
        if name == "trained":
            res.update(entry)
        else:
            res[name] = entry
    eng.sync()
    return res


def cigar_leg(args, eng, gt):
    """Times ksw_global2 with CIGARs (bsx_global_batch_device) on every try of a
    CIGAR pass (mem_reg2aln's band inference and retries), as one resident batch
    of 4*md_tile copies.  The baseline is the compiled reference's ksw_global2
    over the same batch on the host threads (the oracle's restatement where the
    reference was not built), which is also the parity check."""
    from bsbolt_amd.tasks import GlbTasks
    m = len(gt)
    reps = max(1, 4 * args.md_tile)  # few regions need DP, so more copies fill the GPU
    qtot, ttot = len(gt.query), len(gt.target)
    t = GlbTasks(np.tile(gt.query, reps), np.concatenate([gt.q_off + r * qtot for r in range(reps)]),
                 np.tile(gt.qlen, reps), np.tile(gt.target, reps),
                 np.concatenate([gt.t_off + r * ttot for r in range(reps)]), np.tile(gt.tlen, reps),
                 np.tile(gt.w, reps))
    n = len(t)
    slot = int((gt.qlen + gt.tlen).max())
    keep = {k: eng.to_device(np.ascontiguousarray(getattr(t, k))) for k in
            ("query", "q_off", "qlen", "target", "t_off", "tlen", "w")}
    res = {"score": eng.empty((n,), np.int32), "n_cigar": eng.empty((n,), np.int32),
           "cigar": eng.empty((n, slot), np.uint32)}
    dev = {k: v.ptr for k, v in {**keep, **res}.items()}
    dev.update(n=n, cigar_slot=slot)
    for _ in range(max(1, args.warmup)):
        eng.global_device(dev)
    eng.sync()
    eng.timer_start()
    for _ in range(args.steps):
        eng.global_device(dev)
    ms = eng.timer_stop() / args.steps
    # banded cells of ksw_global2 (ksw.c:530-560): rows x min(qlen, 2w+1)
    band = np.minimum(gt.qlen.astype(np.int64), 2 * gt.w.astype(np.int64) + 1)
    cells = int((band * gt.tlen.astype(np.int64)).sum()) * reps
    log(f"CIGAR pass (ksw_global2): {n} tasks ({m} distinct x {reps}), {ms:.3f} ms")
    entry = {"value": n / (ms / 1e3), "unit": "alignments/s", "tasks": n, "ms_per_pass": ms,
             "gcups_banded": cells / (ms / 1e3) / 1e9, "cpu_baseline": None}
    if not args.no_cpu:
        from oracle import oracle as O
        use_ref = O.ref_available()
        # the whole resident batch on the host threads: the compiled reference's
        # ksw_global2 where it was built, else the oracle's restatement
        sc_, nc_, cg_, dt = O.global2_batch(t, slot, nthreads=args.cpu_threads, use_ref=use_ref)
        got_s, got_n = res["score"].numpy(), res["n_cigar"].numpy()
        got_c = np.empty((m, slot), np.uint32)  # the distinct tasks' CIGAR slots only
        eng._check(eng.L.bsx_memcpy(eng.ctx, got_c.ctypes.data, res["cigar"].ptr, got_c.nbytes, 2))
        mism = int((got_s != sc_).sum() + (got_n != nc_).sum())
        mism += sum(not np.array_equal(got_c[i, :nc_[i]], cg_[i, :nc_[i]]) for i in range(m) if nc_[i] > 0)
        kind = "reference" if use_ref else "port"
        entry["cpu_baseline"] = {"value": n / dt, "unit": "alignments/s", **cpu_fields(args), "kind": kind,
                                 "sample": f"the whole batch ({n} tasks) on {args.cpu_threads} threads, dynamic "
                                           f"4096-task chunks; scores and CIGAR counts compared on every task, "
                                           f"CIGARs on the {m} distinct ones",
                                 "seconds": dt, "mismatches_vs_gpu": int(mism)}
        log(f"CIGAR pass CPU baseline ({kind}): {n / dt / 1e6:.3f} M tasks/s on {args.cpu_threads} threads, "
            f"{mism} mismatches vs GPU")
    del keep, res
    return entry


def md_leg(args, eng):
    """Times the NM / MD:Z / XB:Z pass (the tail of bwa_gen_cigar2, bwa.c:250-343)
    on the GPU: the final CIGAR of every alignment region of the same workload
    at reduced scale (its CIGAR pass runs ksw_global2 on the GPU first)."""
    from bsbolt_amd.workload import CONFIGS, Workload, build_glb_batch, md_inputs
    cfg = CONFIGS[args.workload].scaled(args.md_scale)
    wl = Workload(cfg)
    try:
        wl.build_ext_batch(lambda t: eng.extend(t, cells=False))
        out = build_glb_batch(wl, lambda t: eng.global_(t), return_regions=True)
        idx, rb, re, oq, oq_off, lq, nc, cig = md_inputs(wl, out[7], out[4], out[3])
        opac, l_pac, crick_l = wl.opac(), wl.l_pac, cfg.genome_len
    finally:
        wl.close()
    glb = cigar_leg(args, eng, out[0])
    # the regions' tasks repeated md_tile times: a batch the size of a full
    # chunk of reads, so the pass is timed at full occupancy
    reps = max(1, args.md_tile)
    rb, re, oq_off, lq, nc = (np.tile(x, reps) for x in (rb, re, oq_off, lq, nc))
    cig = np.tile(cig, (reps, 1))
    n = len(rb)
    md_slot, xb_slot = 4 * int(lq.max()) + 4 * cig.shape[1] + 16, 4 * int(lq.max()) + 16
    host = {"opac": opac, "rb": rb, "re": re, "oquery": oq, "oq_off": oq_off, "l_query": lq, "n_cigar": nc,
            "cigar": cig}
    keep = {k: eng.to_device(np.ascontiguousarray(v)) for k, v in host.items()}
    outs = {"md": eng.empty((n, md_slot), np.uint8), "md_len": eng.empty((n,), np.int32),
            "xb": eng.empty((n, xb_slot), np.uint8), "xb_len": eng.empty((n,), np.int32),
            "nm": eng.empty((n,), np.int32), "meth": eng.empty((n, 4), np.int32), "status": eng.empty((n,), np.int32)}
    dev = {k: v.ptr for k, v in {**keep, **outs}.items()}
    dev.update(n=n, l_pac=l_pac, crick_l=crick_l, cigar_slot=cig.shape[1], md_slot=md_slot, xb_slot=xb_slot)
    for _ in range(max(1, args.warmup)):
        eng.md_device(dev)
    eng.sync()
    status = outs["status"].numpy()
    ml, xl = outs["md_len"].numpy(), outs["xb_len"].numpy()
    eng.timer_start()
    for _ in range(args.steps):
        eng.md_device(dev)
    ms = eng.timer_stop() / args.steps
    # algorithmic bytes: the 2-bit context window, the read, the CIGAR, the
    # task fields and the strings and counts written
    alg = int(((re - rb + 4 + 3) // 4).sum() + lq.astype(np.int64).sum() + 4 * nc.astype(np.int64).sum()
              + ml.astype(np.int64).sum() + xl.astype(np.int64).sum()) + n * (8 + 8 + 8 + 4 + 4 + 4 * 8)
    log(f"NM/MD/XB pass: {n} regions ({len(idx)} distinct x {reps}), {ms:.3f} ms, "
        f"status {np.bincount(status + 3, minlength=5).tolist()} (-3..1)")
    cpu = None
    if not args.no_cpu:  # the oracle's restatement on the host threads, over the distinct regions
        from oracle import oracle as O
        m = len(idx)
        ref, dt = O.gen_md_batch(l_pac, opac, crick_l, rb[:m], re[:m], oq, oq_off[:m], lq[:m], nc[:m], cig[:m],
                                 md_slot, xb_slot, nthreads=args.cpu_threads)
        gmd, gxb = outs["md"].numpy()[:m], outs["xb"].numpy()[:m]
        mism = int((ref["status"] != status[:m]).sum() + (ref["nm"] != outs["nm"].numpy()[:m]).sum()
                   + (ref["meth"] != outs["meth"].numpy()[:m]).any(axis=1).sum())
        mism += sum(ref["md"][i] != bytes(gmd[i, :ml[i]]).decode() or ref["xb"][i] != bytes(gxb[i, :xl[i]]).decode()
                    for i in range(m))
        cpu = {"value": m / dt, "unit": "alignments/s", **cpu_fields(args), "kind": "port",
               "sample": f"the {m} distinct regions, oracle restatement of bwa.c:250-343", "seconds": dt,
               "mismatches_vs_gpu": int(mism)}
        log(f"NM/MD/XB CPU restatement: {m / dt / 1e6:.2f} M regions/s on {args.cpu_threads} threads, "
            f"{mism} mismatches vs GPU")
    del keep, outs
    return {"value": n / (ms / 1e3), "unit": "alignments/s", "regions": n, "ms_per_pass": ms, "cpu_baseline": cpu,
            "cigar_pass": glb,
            "workload": f"{args.workload} x {args.md_scale}: final CIGAR of every alignment region "
                        f"({len(idx)} regions, batch of {reps} copies)",
            "status_ok": int((status == 0).sum()), "status_bridging": int((status == 1).sum()),
            "roofline_hbm": {"bound": "hbm", "achieved": alg / (ms / 1e3) / 1e9, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                             "frac": alg / (ms / 1e3) / 1e9 / PEAK_HBM_GBS, "algorithmic_bytes": alg}}


EXT_ITEM = {"query": 1, "q_off": 8, "qlen": 4, "target": 1, "t_off": 8, "tlen": 4, "h0": 4, "w": 4, "end_bonus": 4}


def chunk_ptrs(ptrs, lo, n, item):
    """Device addresses of tasks [lo, lo+n) of a resident batch: per-task arrays
    offset by lo, the concatenated sequences shared (offsets are absolute)."""
    p = dict(ptrs)
    for k, sz in item.items():
        if k not in ("query", "target"):
            p[k] = ptrs[k] + lo * sz
    p["n"] = n
    return p


def workload_leg(args, eng, name, cigar=False):
    """Another workload of the configs: its mem_chain2aln batch built in rounds
    on the GPU, then timed resident in HBM in the plain, learned and
    no-predictor frame forms (outputs must equal the plain form's).  With
    `cigar`, the batch is also timed as hipGraph-captured chunks of
    args.graph_batch tasks, and the workload's CIGAR pass (every try of
    mem_reg2aln's band doubling) runs as graph-captured chunks of the same size
    after it."""
    from bsbolt_amd.workload import CONFIGS, Workload, build_glb_batch
    from bsbolt_amd.tasks import RESULT_FIELDS
    cfg = CONFIGS[name]
    t0 = time.time()
    wl = Workload(cfg)
    batch, round_res, side, rnd = wl.build_ext_batch(lambda t: eng.extend(t, cells=False))
    glb = None
    if cigar:
        glb = build_glb_batch(wl, lambda t: eng.global_(t))
    wl.close()
    n = len(batch)
    log(f"{name}: {wl.n_reads} reads, {n} extension tasks built in {time.time() - t0:.1f}s")
    keep, ptrs = eng.tasks_to_device(batch)
    outs = {k: eng.empty((n,), np.int32) for k in RESULT_FIELDS}
    cells_d = eng.empty((n,), np.int64)
    ptrs.update({k: v.ptr for k, v in outs.items()})
    ptrs["cells"] = cells_d.ptr
    ptrs["n"] = n
    eng.extend_device(ptrs)
    eng.sync()
    got = {k: v.numpy() for k, v in outs.items()}
    for k in RESULT_FIELDS:
        if not np.array_equal(got[k], round_res[k]):
            raise SystemExit(f"{name}: resident run disagrees with the build rounds on {k}")
    cells = cells_d.numpy()
    total_cells = int(cells.sum())
    steps = max(1, args.extra_steps)
    ms = time_steps(eng, ptrs, steps)
    leg = {"workload": f"{name}: {cfg.genome_len // 1_000_000} Mb genome, {wl.n_reads} reads x {cfg.read_len} bp"
                       f"{' PE' if cfg.paired else ' SE'}{' RRBS' if cfg.rrbs else ''}",
           "tasks": n, "cells_ref": total_cells, "mean_qlen": float(batch.qlen.mean()),
           "plain": {"value": n / (ms / 1e3), "unit": "extensions/s", "ms_per_step": ms,
                     "gcups": total_cells / (ms / 1e3) / 1e9}}
    log(f"{name} plain: {ms:.3f} ms/step, {total_cells / (ms / 1e3) / 1e9:.1f} GCUPS")
    if not args.no_learned:
        a2 = argparse.Namespace(**vars(args))
        a2.steps, a2.warmup = steps, 1
        lb = learned_leg(a2, eng, ptrs, outs, cells_d, got, cells, n, total_cells)
        for x in (lb, lb.get("fixed24")):
            if x:
                x["speedup_vs_plain"] = ms / x["ms_per_step"]
        leg["learned_band"] = lb
    if cigar:
        gb = args.graph_batch
        # extension in graph-captured chunks over the resident batch: one graph
        # per chunk, replayed on one stream, then spread round-robin over
        # args.graph_streams contexts (one stream each) so that a chunk's tail
        # overlaps the next chunk's work
        from bsbolt_amd import bsx as _bsx
        ctxs = [eng] + [_bsx.Engine(0) for _ in range(max(0, args.graph_streams - 1))]
        chunks = [chunk_ptrs(ptrs, lo, min(gb, n - lo), {**EXT_ITEM, **{k: 4 for k in RESULT_FIELDS}, "cells": 8})
                  for lo in range(0, n, gb)]
        timing = {}
        for mode, pool in (("one_stream", [eng]), ("streams", ctxs)):
            for e in pool:  # with --graph-in-order the chunks' graphs hold no side-stream fork/join
                e.set_side_streams(not args.graph_in_order)
            graphs = [(pool[j % len(pool)], pool[j % len(pool)].capture_extend(p)) for j, p in enumerate(chunks)]
            for e in pool:
                e.set_side_streams(True)
            for e, g in graphs:
                g.launch()
            for e in pool:
                e.sync()
            for k in RESULT_FIELDS:
                if not np.array_equal(outs[k].numpy(), got[k]):
                    raise SystemExit(f"{name}: graph-batched extension ({mode}) disagrees on {k}")
            t1 = time.perf_counter()
            for _ in range(steps):
                for e, g in graphs:
                    g.launch()
            for e in pool:
                e.sync()
            timing[mode] = (time.perf_counter() - t1) * 1e3 / steps
            for e, g in graphs:
                g.close()
        for e in ctxs[1:]:
            e.close()
        ms_g = timing["streams"]
        graphs = chunks
        # the CIGAR pass: every try, as graph-captured chunks
        gt = glb[0]
        m = len(gt)
        slot = int((gt.qlen + gt.tlen).max())
        gkeep = {k: eng.to_device(np.ascontiguousarray(getattr(gt, k))) for k in
                 ("query", "q_off", "qlen", "target", "t_off", "tlen", "w")}
        gres = {"score": eng.empty((m,), np.int32), "n_cigar": eng.empty((m,), np.int32),
                "cigar": eng.empty((m, slot), np.uint32)}
        gp = {k: v.ptr for k, v in {**gkeep, **gres}.items()}
        gp.update(n=m, cigar_slot=slot)
        ggraphs = []
        for lo in range(0, m, gb):
            p = chunk_ptrs(gp, lo, min(gb, m - lo), {"q_off": 8, "qlen": 4, "t_off": 8, "tlen": 4, "w": 4,
                                                      "score": 4, "n_cigar": 4, "cigar": 4 * slot})
            ggraphs.append(eng.capture_global(p))
        for g in ggraphs:
            g.launch()
        eng.sync()
        mism = int((gres["score"].numpy() != glb[1]).sum() + (gres["n_cigar"].numpy() != glb[2]).sum())
        eng.timer_start()
        for _ in range(steps):
            for g in ggraphs:
                g.launch()
        ms_c = eng.timer_stop() / steps
        for g in ggraphs:
            g.close()
        band = np.minimum(gt.qlen.astype(np.int64), 2 * gt.w.astype(np.int64) + 1)
        gcells = int((band * gt.tlen.astype(np.int64)).sum())
        leg["graph_pass"] = {
            "graph_batch": gb, "extension_graphs": len(graphs), "cigar_graphs": len(ggraphs),
            "graph_streams": args.graph_streams, "graph_in_order": bool(args.graph_in_order), "extension_ms": ms_g,
            "extension_ms_one_stream": timing["one_stream"], "extension_ms_one_call": ms,
            "cigar_tasks": m, "cigar_ms": ms_c, "cigar_mismatches_vs_rounds": mism,
            "value": (n + m) / ((ms_g + ms_c) / 1e3), "unit": "tasks/s (ksw_extend2 + ksw_global2)",
            "ms_per_pass": ms_g + ms_c, "gcups_extension": total_cells / (ms_g / 1e3) / 1e9,
            "gcups_cigar_banded": gcells / (ms_c / 1e3) / 1e9}
        log(f"{name} graph pass: {len(graphs)} extension graphs {ms_g:.3f} ms on {args.graph_streams} streams "
            f"({timing['one_stream']:.3f} ms on one) + {len(ggraphs)} CIGAR graphs "
            f"({m} tasks) {ms_c:.3f} ms, {mism} CIGAR mismatches")
        del gkeep, gres
    if not args.no_cpu:
        from oracle import oracle as O
        kind = "reference" if O.ref_available() else "port"
        cpu_out, dt = O.extend2_batch(batch, nthreads=args.cpu_threads, use_ref=(kind == "reference"),
                                      want_cells=False)
        mism = int(sum(int((cpu_out[k] != got[k]).sum()) for k in RESULT_FIELDS))
        leg["cpu_baseline"] = {"value": n / dt, "unit": "extensions/s", **cpu_fields(args), "kind": kind,
                               "gcups": total_cells / dt / 1e9, "seconds": dt,
                               "sample": f"the full batch ({n} tasks), dynamic 4096-task chunks per thread",
                               "mismatches_vs_gpu": mism}
        log(f"{name} CPU baseline: {n / dt / 1e6:.2f} M ext/s, {mism} mismatches")
    del keep, outs, cells_d
    return leg


def host_buffer_leg(eng, batch, reps):
    """The host-memory entry point (bsx_extend_batch: H2D copies, the kernels,
    D2H copies) on the same batch: the PCIe-inclusive rate."""
    eng.extend(batch, cells=False)
    t0 = time.perf_counter()
    for _ in range(reps):
        eng.extend(batch, cells=False)
    dt = (time.perf_counter() - t0) / reps
    return {"value": len(batch) / dt, "unit": "extensions/s", "ms_per_call": dt * 1e3,
            "note": "host buffers in and out (bsx_extend_batch), plain form"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="cfg1")
    ap.add_argument("--scale", type=float, default=1.0, help="fraction of the workload's reads")
    ap.add_argument("--cpu-threads", type=int, default=host_cpu()["threads"],
                    help="CPU-baseline threads (default: every CPU this job may use)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-band leg")
    ap.add_argument("--no-md", action="store_true", help="skip the NM/MD/XB leg")
    ap.add_argument("--md-scale", type=float, default=0.05, help="fraction of the reads for the NM/MD/XB leg")
    ap.add_argument("--md-tile", type=int, default=16, help="copies of those regions in the timed NM/MD/XB batch")
    ap.add_argument("--band-model", default=None, help="band model (.npy); default bsbolt_amd/band_model.npy")
    ap.add_argument("--extra", default="cfg3,cfg4",
                    help="other workloads timed after the headline one (cfg4 with its graph-batched CIGAR pass)")
    ap.add_argument("--extra-steps", type=int, default=5)
    ap.add_argument("--graph-batch", type=int, default=262144, help="tasks per hipGraph-captured chunk")
    ap.add_argument("--graph-streams", type=int, default=4, help="contexts (streams) the chunks' graphs are spread over")
    ap.add_argument("--graph-in-order", type=int, default=0,
                    help="capture the chunks' graphs with the launches in order (1) or with side streams (0)")
    ap.add_argument("--copy-bytes", type=int, default=0,
                    help="only run one device-to-device copy of this many bytes (HBM counter calibration)")
    args = ap.parse_args()
    if args.copy_bytes:
        from bsbolt_amd import bsx
        import numpy as _np
        eng = bsx.Engine(0)
        a = eng.empty((args.copy_bytes,), _np.uint8)
        b = eng.empty((args.copy_bytes,), _np.uint8)
        eng._check(eng.L.bsx_memcpy(eng.ctx, b.ptr, a.ptr, args.copy_bytes, 3))
        eng.timer_start()
        eng._check(eng.L.bsx_memcpy(eng.ctx, b.ptr, a.ptr, args.copy_bytes, 3))
        ms = eng.timer_stop()
        print(json.dumps({"copy_bytes": args.copy_bytes, "ms": ms,
                          "GB/s": 2 * args.copy_bytes / (ms / 1e3) / 1e9}), flush=True)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on exactly one GPU; the extension batch shards "
                                   "by task index with no exchange step, multi-GPU runs are "
                                   "not part of this build", "n_gpus": args.gpus}), flush=True)
        sys.exit(2)

    from bsbolt_amd import bsx
    from bsbolt_amd.workload import CONFIGS, Workload
    from bsbolt_amd.tasks import RESULT_FIELDS

    cfg = CONFIGS[args.workload]
    if args.scale != 1.0:
        cfg = cfg.scaled(args.scale)
    eng = bsx.Engine(0)
    tune_table, tune_src, tune_key = eng.get_tuning()
    log(f"tuning table {tune_table} ({tune_src}, {tune_key})")
    t0 = time.time()
    wl = Workload(cfg)
    log(f"simulated {wl.n_reads} reads on a {cfg.genome_len / 1e6:.0f} Mb genome in {time.time() - t0:.1f}s")
    t0 = time.time()
    batch, round_res, side, rnd = wl.build_ext_batch(lambda t: eng.extend(t, cells=False),
                                                     on_round=lambda r, n: log(f"round {r}: {n} tasks"))
    wl.close()
    n = len(batch)
    log(f"built {n} extension tasks in {time.time() - t0:.1f}s")

    # resident batch
    keep, ptrs = eng.tasks_to_device(batch)
    outs = {k: eng.empty((n,), np.int32) for k in RESULT_FIELDS}
    cells_d = eng.empty((n,), np.int64)
    ptrs.update({k: v.ptr for k, v in outs.items()})
    ptrs["cells"] = cells_d.ptr
    ptrs["n"] = n
    for _ in range(max(1, args.warmup)):  # at least one run: its outputs are verified below
        eng.extend_device(ptrs)
    eng.sync()
    got = {k: v.numpy() for k, v in outs.items()}
    cells = cells_d.numpy()
    for k in RESULT_FIELDS:  # the resident run must agree with the rounds that built the batch
        if not np.array_equal(got[k], round_res[k]):
            raise SystemExit(f"resident run disagrees with the build rounds on {k}")
    total_cells = int(cells.sum())
    nominal = int((batch.qlen.astype(np.int64) * batch.tlen).sum())

    # timed region: K steps bracketed by stream synchronisation; HIP events on
    # the context's stream, which joins the side streams at the end of a step
    eng.sync()
    w0 = time.perf_counter()
    eng.timer_start()
    for _ in range(args.steps):
        eng.extend_device(ptrs)
    ms = eng.timer_stop()
    eng.sync()
    wall = time.perf_counter() - w0
    # per-kernel times for the roofline: the same steps with the launches in
    # order, so that each kernel has the GPU to itself (HIP events around every
    # kernel, on the stream it is launched on)
    eng.set_side_streams(False)
    eng.set_kernel_timing(args.steps)
    eng.timer_start()
    for _ in range(args.steps):
        eng.extend_device(ptrs)
    ms_in_order = eng.timer_stop() / args.steps
    kms, kcalls = eng.kernel_times()
    eng.set_kernel_timing(0)
    eng.set_side_streams(True)
    ms_step = ms / args.steps
    ext_s = n / (ms_step / 1e3)
    gcups = total_cells / (ms_step / 1e3) / 1e9
    tops = total_cells * OPS_PER_CELL / (ms_step / 1e3) / 1e12
    alg_bytes = int(batch.qlen.astype(np.int64).sum() + batch.tlen.astype(np.int64).sum()) + n * (36 + 24 + 8)
    gbs = alg_bytes / (ms_step / 1e3) / 1e9
    log(f"{args.steps} steps: {ms_step:.3f} ms/step (events), wall {wall * 1e3 / args.steps:.3f} ms/step")
    # dominant kernel: the plain bin with the most time; its algorithmic work is
    # the cells of the tasks in that bin
    limits = bsx.ext_bin_limits()
    bin_of = np.searchsorted(limits, batch.qlen)
    per_launch = kms[:len(limits)] / max(kcalls, 1)
    dom = int(np.argmax(per_launch))
    lanes, cols = bsx.EXT_SHAPES[dom]
    dom_name = f"ext_lane_kernel<{cols}, true>" if lanes == 1 else f"ext_plain_kernel<{lanes}, {cols}, true>"
    dom_cells = int(cells[bin_of == dom].sum())
    dom_ms = float(per_launch[dom])
    dom_tops = dom_cells * OPS_PER_CELL / (dom_ms / 1e3) / 1e12
    m = bin_of == dom
    dom_bytes = int(batch.qlen[m].astype(np.int64).sum() + batch.tlen[m].astype(np.int64).sum()) + int(m.sum()) * 80
    traffic = profiled_traffic(dom_name, args.workload)
    log(f"dominant kernel {dom_name}: {dom_ms:.3f} ms/launch, {dom_tops:.2f} TOP/s, traffic {traffic}")

    # ---- learned-band form: same batch, same outputs required ----
    learned = None
    if not args.no_learned:
        learned = learned_leg(args, eng, ptrs, outs, cells_d, got, cells, n, total_cells)

    # roofline of the frame kernels the 24-column form runs (class 0: the lane
    # frame kernel for queries <= 128 bases, the 8x3 group frame kernel above):
    # useful work = the cells of the tasks whose run the 24-column frame holds
    # (band label 0 from the GPU's own labelling pass), over the class's kernel
    # time with the launches in order
    frame_roof = None
    if learned and learned.get("fixed24"):
        labels = eng.band_labels(batch)
        acc = labels == 0
        useful = int(cells[acc].sum())
        f24 = learned["fixed24"]
        t_ms = f24["kernel_ms_per_step"]["narrow"]
        tops_f = useful * OPS_PER_CELL / (t_ms / 1e3) / 1e12
        frame_roof = {"bound": "valu", "kernel": "ext_lane_frame_kernel<24> + ext_narrow_kernel<8, 3> (class 0)",
                      "achieved": tops_f, "peak": PEAK_PK16_TOPS, "unit": "TOP/s", "frac": tops_f / PEAK_PK16_TOPS,
                      "frac_int32": tops_f / PEAK_INT32_TOPS, "ms_per_launch": t_ms, "useful_cells": useful,
                      "tasks_held": int(acc.sum()), "tasks_routed": int(f24["routed"]["f24"]),
                      "tasks_held_by_stats": int(f24["narrow_accepted"])}
        log(f"24-column frame kernels: {t_ms:.3f} ms, {useful} useful cells, {tops_f:.2f} TOP/s "
            f"({tops_f / PEAK_PK16_TOPS:.4f} of packed peak); held {int(acc.sum())} (stats {f24['narrow_accepted']})")

    best_form = "plain"
    best_ms = ms_step
    for form, leg in (("learned", learned), ("fixed 24-column frame", (learned or {}).get("fixed24"))):
        if leg and leg.get("mismatches_vs_plain") == 0 and leg["ms_per_step"] < best_ms:
            best_form, best_ms = form, leg["ms_per_step"]
    line = {
        "metric": "extensions/s + GCUPS for bit-exact bisulfite ksw_extend2; learned-band vs plain",
        "value": n / (best_ms / 1e3),
        "unit": "extensions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": best_ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic (seeded genome + bisulfite read simulation; no external data)",
        "config": {"workload": f"{args.workload}: {cfg.genome_len // 1_000_000} Mb genome, {wl.n_reads} reads "
                               f"x {cfg.read_len} bp{' PE' if cfg.paired else ' SE'}, all mem_chain2aln "
                               f"ksw_extend2 calls", "tasks": n, "form": best_form,
                   "cells_ref": total_cells, "cells_nominal": nominal,
                   "mean_qlen": float(batch.qlen.mean()), "mean_tlen": float(batch.tlen.mean())},
        "plain": {"value": ext_s, "ms_per_step": ms_step, "gcups": gcups},
        "gcups": total_cells / (best_ms / 1e3) / 1e9,
        "gcups_nominal": nominal / (best_ms / 1e3) / 1e9,
        "roofline": {"bound": "valu", "kernel": dom_name, "achieved": dom_tops, "peak": PEAK_PK16_TOPS,
                     "unit": "TOP/s", "frac": dom_tops / PEAK_PK16_TOPS, "frac_int32": dom_tops / PEAK_INT32_TOPS,
                     "traffic": traffic, "ops_per_launch": dom_cells * OPS_PER_CELL, "ops_per_cell": OPS_PER_CELL,
                     "ms_per_launch": dom_ms, "share_of_step_in_order": dom_ms / ms_in_order,
                     "ms_per_step_launches_in_order": ms_in_order,
                     "algorithmic_bytes_per_launch": dom_bytes,
                     "kernel_ms_per_step": {f"bin<={int(l)}": float(t) for l, t in zip(limits, per_launch) if t > 0}},
        "roofline_step": {"bound": "valu", "achieved": tops, "peak": PEAK_PK16_TOPS, "unit": "TOP/s",
                          "frac": tops / PEAK_PK16_TOPS, "frac_int32": tops / PEAK_INT32_TOPS},
        "roofline_hbm": {"bound": "hbm", "achieved": gbs, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                         "frac": gbs / PEAK_HBM_GBS, "algorithmic_bytes": alg_bytes},
        "learned_band": learned,
        "roofline_frame24": frame_roof,
        "tuning": {"table": dict(zip(("tpb_plain", "tpb_narrow", "shape_var_65_96", "shape_var_97_128"), tune_table)),
                   "source": tune_src, "device": tune_key},
    }
    if not args.no_md:
        line["md_layer"] = md_leg(args, eng)
    line["host_buffers"] = host_buffer_leg(eng, batch, 3)
    log(f"host-buffer entry point: {line['host_buffers']['ms_per_call']:.1f} ms per call")
    del keep, outs, cells_d
    extra = [x for x in args.extra.split(",") if x and x != args.workload]
    line["workloads"] = {x: workload_leg(args, eng, x, cigar=(x == "cfg4")) for x in extra}
    if learned:  # the learned-band form over the plain form, and the no-predictor frame form
        learned["speedup_vs_plain"] = ms_step / learned["ms_per_step"]
        if learned.get("fixed24"):
            learned["fixed24"]["speedup_vs_plain"] = ms_step / learned["fixed24"]["ms_per_step"]

    if not args.no_cpu:
        from oracle import oracle as O
        kind = "reference" if O.ref_available() else "port"
        log(f"CPU baseline: {kind} ksw_extend2 over all {n} tasks on {args.cpu_threads} threads")
        cpu_out, dt = O.extend2_batch(batch, nthreads=args.cpu_threads, use_ref=(kind == "reference"),
                                      want_cells=False)
        mism = int(sum(int((cpu_out[k] != got[k]).sum()) for k in RESULT_FIELDS))
        line["cpu_baseline"] = {"value": n / dt, "unit": "extensions/s", **cpu_fields(args),
                                "kind": kind, "gcups": total_cells / dt / 1e9, "seconds": dt,
                                "sample": f"the full batch ({n} tasks), dynamic 4096-task chunks per thread",
                                "mismatches_vs_gpu": mism}
        line["speedup_vs_cpu"] = line["value"] / (n / dt)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
