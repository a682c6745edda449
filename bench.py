"""Exact-BC throughput on one MI355X: batched FP32-compensated path vs the
plain per-source path, on the 1M-vertex mixed graph (Chung-Lu power-law core
+ degree-2 chains + pendant trees).

A step is one pass of the hot path over one slice of the job's sources: the
job's sources (every vertex of every biconnected component with more than two
vertices) are taken in the library's tile order inside their component and
cut into blocks of 256 consecutive sources (a whole number of the library's
tiles of 64, 128 or 256 sources), and slice i holds every S-th block (i, i+S,
...), so each slice is a uniform sample of the job with the tile structure of
a full run.  A step is credited n*m * (slice sources / job sources) traversals
(MTEPS = n*m / t, n and m of the input graph).  Inputs are resident in HBM
before the timed region; each step is one bcx_run through the C-ABI.

One JSON line on stdout (rank 0, N=1).  --gpus N > 1 exits: the benchmark runs
on one GPU.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "centrality-gpu_amd"))
import bcx  # noqa: E402

HBM_PEAK_GBS = 8000.0


def make_slices(comps, offsets, orders, n_slices, tile=64):
    """Slot lists per slice, plus per-slice (bfs-like, chain) source counts
    per component for the algorithmic byte model.  orders[c] is the library's
    tile order of component c (bcx_batched_order), so a slice is a set of
    whole tiles of a full run."""
    tiles = []
    for ci, c in enumerate(comps):
        order = np.asarray(orders[ci], np.int64)
        for t0 in range(0, c.n, tile):
            tiles.append((ci, order[t0:t0 + tile]))
    slices = [[] for _ in range(n_slices)]
    for i, (ci, verts) in enumerate(tiles):
        slices[i % n_slices].append((ci, verts))
    out = []
    chain_flags = []
    for c in comps:
        f = np.zeros(c.n, bool)
        f[c.chain_nodes] = True
        chain_flags.append(f)
    for sl in slices:
        slots = np.concatenate([offsets[ci] + v for ci, v in sl]) if sl else np.zeros(0, np.int64)
        bfs = np.zeros(len(comps), np.int64)
        chn = np.zeros(len(comps), np.int64)
        for ci, v in sl:
            x = int(chain_flags[ci][v].sum())
            chn[ci] += x
            bfs[ci] += len(v) - x
        out.append((np.sort(slots), bfs, chn))
    return out


def bytes_alg(comps, bfs, chn):
    """Compulsory HBM bytes of the method: |S|(16a+80n) + |X|(8a+100n)."""
    tot = 0.0
    for c, s, x in zip(comps, bfs, chn):
        tot += s * (16.0 * c.arcs + 80.0 * c.n) + x * (8.0 * c.arcs + 100.0 * c.n)
    return tot


def copy_probe(mib):
    """A plain device-to-device copy of `mib` MiB (hipMemcpy), for scaling
    rocprofv3 FETCH_SIZE / WRITE_SIZE readings to bytes."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    nbytes = mib << 20
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(nbytes)) == 0
    for _ in range(3):
        assert hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3) == 0  # device to device
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"copy_probe_mib": mib, "copies": 3}))


BATCHED_KERNELS = ("bcx::k_b_", "bcx::k_small", "bcx::k_wg", "bcx::k_na_")


def _union(iv):
    """Total length of the union of [start, end) intervals."""
    tot, cur_s, cur_e = 0, None, None
    for s, e in sorted(iv):
        if cur_e is None or s > cur_e:
            if cur_e is not None:
                tot += cur_e - cur_s
            cur_s, cur_e = s, e
        else:
            cur_e = max(cur_e, e)
    if cur_e is not None:
        tot += cur_e - cur_s
    return tot


def summarize_trace(path):
    """Markdown summary of a rocprofv3 kernel-trace CSV: per kernel the calls,
    the summed duration and the busy time as the union of its launches'
    intervals (kernels of concurrent tiles overlap, so sums exceed the wall
    time); for the batched kernels together, their union and the span from the
    first start to the last end."""
    import csv
    rows = {}
    with open(path) as f:
        for r in csv.DictReader(f):
            s, e = int(r["Start_Timestamp"]), int(r["End_Timestamp"])
            rows.setdefault(r["Kernel_Name"], []).append((s, e))
    print("| Kernel | Calls | Sum ms | Union ms | Avg us |")
    print("|---|---:|---:|---:|---:|")
    tot = lambda iv: sum(e - s for s, e in iv)  # noqa: E731
    for k, iv in sorted(rows.items(), key=lambda kv: -tot(kv[1])):
        print(f"| `{k[:60]}` | {len(iv)} | {tot(iv) / 1e6:.1f} | {_union(iv) / 1e6:.1f} | "
              f"{tot(iv) / len(iv) / 1e3:.1f} |")
    b = [x for k, iv in rows.items() if any(p in k for p in BATCHED_KERNELS) for x in iv]
    if b:
        print(f"\nBatched kernels: {len(b)} launches, union {_union(b) / 1e6:.1f} ms, "
              f"span first start to last end {(max(e for _, e in b) - min(s for s, _ in b)) / 1e6:.1f} ms")


TRAFFIC_FILES = {2: "batched_config2_traffic.json", 3: "batched_config3_traffic.json",
                 4: "wg_config4_traffic.json", 5: "batched_config5_traffic.json"}
WORKLOADS = {
    1: "16 Erdos-Renyi blocks of 128 vertices on a tree of cut vertices, 20% of edges as chains, "
       "512 pendant vertices (10k vertices)",
    2: "1M-vertex Chung-Lu core + degree-2 chains + pendant trees",
    3: "519k-vertex Chung-Lu block + 40,000 small Erdos-Renyi blocks on a tree of cut vertices (12.8M vertices)",
    4: "256x256 grid with 5% of edges deleted and 2,000 spliced chains (74k vertices)",
    5: "R-MAT scale 20, edge factor 16, 10% of vertices as chains (750k vertices)",
}


def measured_traffic(config, src_per_step):
    """HBM bytes per step from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE
    summary of this workload (scaled per source when the slice differs);
    None for workloads without one."""
    name = TRAFFIC_FILES.get(config)
    p = os.path.join(ROOT, "profiles", name) if name else None
    if not p or not os.path.exists(p):
        return None
    with open(p) as f:
        t = json.load(f)
    per_src = t["hbm_bytes_per_step"] / t["sources_per_step"]
    return per_src * src_per_step


def timed_steps(g, slices, idx, warmup, **kw):
    for i in range(warmup):
        g.run(subset=slices[idx[i % len(idx)]][0], **kw)
    t_kernel = 0.0
    stats = []
    t0 = time.perf_counter()
    for i in idx:
        _, st = g.run(subset=slices[i][0], **kw)
        t_kernel += st["t_kernel_s"]
        stats.append(st)
    t = time.perf_counter() - t0
    return t, t_kernel, stats


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=8)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", type=int, default=2)
    ap.add_argument("--scale", type=float, default=1.0)
    ap.add_argument("--slices", type=int, default=0,
                    help="0 = per workload: 64 for the wide graphs, 1 for the deep grid (4)")
    ap.add_argument("--plain-sources", type=int, default=4096,
                    help="sources the plain path is timed on (whole tiles, evenly spaced over the job; "
                         "8x the plain path's persistent grid of 512 workgroups)")
    ap.add_argument("--cpu-sources", type=int, default=0, help="0 = 48 per host thread (10-30 s of CPU work)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--copy-probe", type=int, default=0, help="MiB; run only a device copy (profiling)")
    ap.add_argument("--summarize-trace", default="",
                    help="rocprofv3 kernel-trace CSV: print per-kernel calls, summed and union busy time")
    ap.add_argument("--slots", type=int, default=0, help="concurrent tiles of the batched path (0 = library default)")
    ap.add_argument("--bpc", type=int, default=0,
                    help="level-kernel blocks per CU of the multi-launch tiles (0 = tuned)")
    ap.add_argument("--spl", type=int, default=0,
                    help="sources per lane of the multi-launch tiles (1, 2, 4; 0 = tuned), for A/B measurements")
    ap.add_argument("--flags", type=int, default=0,
                    help="BCX_FLAG_* bits of the batched run, for A/B measurements (16 = no chain contraction)")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.summarize_trace:
        summarize_trace(args.summarize_trace)
        return
    if args.gpus != 1:
        print(f"bench.py: the benchmark runs on one GPU; --gpus {args.gpus} is not supported", file=sys.stderr)
        sys.exit(2)

    if not args.slices:
        # a step must hold enough tiles to fill the GPU: the deep grid's
        # tiles run one workgroup each, so its step is the whole job
        args.slices = {1: 1, 4: 1}.get(args.config, 64)
    t0 = time.perf_counter()
    n, uv = bcx.generate(args.config, 0, args.scale)
    host = bcx.Host.from_edges(n, uv).decompose()
    comps = host.components()
    m = len(uv)
    offsets = np.concatenate([[0], np.cumsum([c.n for c in comps])]).astype(np.int64)
    total_src = int(offsets[-1])
    t_prep = time.perf_counter() - t0
    print(f"# graph n={n} m={m} components={len(comps)} sources={total_src} prep={t_prep:.1f}s",
          file=sys.stderr, flush=True)

    eng = bcx.Engine(0)
    n_cu, arch = eng.info()
    t0 = time.perf_counter()
    g = eng.upload(comps)
    print(f"# upload of {len(comps)} components {time.perf_counter() - t0:.2f}s", file=sys.stderr, flush=True)
    # steps are slices of whole tiles in the library's tile order (the
    # tile-order tuner runs here, once per component, outside the timing)
    t0 = time.perf_counter()
    orders = [g.batched_order(ci) for ci in range(len(comps))]
    slices = make_slices(comps, offsets, orders, args.slices, tile=256)
    n_tiles = sum((c.n + 63) // 64 for c in comps)
    plain_slices = max(1, n_tiles // max(1, args.plain_sources // 64))
    pslices = make_slices(comps, offsets, orders, plain_slices)
    print(f"# tile orders {time.perf_counter() - t0:.2f}s", file=sys.stderr, flush=True)
    W = float(n) * float(m)

    # batched FP32-compensated path (the headline)
    K = args.steps
    idx = [i % args.slices for i in range(K)]
    t_b, tk_b, st_b = timed_steps(g, slices, idx, args.warmup, path=bcx.PATH_BATCHED,
                                  precision=bcx.PREC_FP32C, use_chains=True, batch_width=args.slots,
                                  flags=args.flags, sources_per_lane=args.spl, blocks_per_cu=args.bpc)
    src_b = sum(len(slices[i][0]) for i in idx)
    work_b = W * src_b / total_src
    mteps_b = work_b / t_b / 1e6
    balg_b = sum(bytes_alg(comps, slices[i][1], slices[i][2]) for i in idx)
    print(f"# batched: {K} steps {t_b:.3f}s (kernel {tk_b:.3f}s) {mteps_b:.1f} MTEPS", file=sys.stderr,
          flush=True)

    plain = None
    if not args.no_plain:
        # one slice of whole tiles (stride plain_slices over the tile order):
        # with chain reduction (the reference's method: chain vertices combine
        # two endpoint rows, which may need forward-only row passes), and
        # without (every vertex a BFS source); the faster one is the comparison
        sub = pslices[0][0]
        g.run(subset=sub[:64], path=bcx.PATH_PLAIN, precision=bcx.PREC_FP64)  # module load, plan buffers
        plain = {"sources": int(len(sub)), "tiles_in_sample": int((len(sub) + 63) // 64)}
        for key, chains in (("chains", True), ("no_chains", False)):
            t0 = time.perf_counter()
            _, st = g.run(subset=sub, path=bcx.PATH_PLAIN, precision=bcx.PREC_FP64, use_chains=chains)
            t_p = time.perf_counter() - t0
            plain[key] = {"mteps": W * len(sub) / total_src / t_p / 1e6, "s": t_p, "kernel_s": st["t_kernel_s"],
                          "bfs_sources": st["n_bfs_sources"], "chain_sources": st["n_chain_sources"],
                          "row_only_passes": st["n_row_only"],
                          "achieved_gbs": st["bytes_alg"] / st["t_kernel_s"] / 1e9 if st["t_kernel_s"] else None}
            print(f"# plain ({key}): {len(sub)} sources {t_p:.3f}s {plain[key]['mteps']:.1f} MTEPS "
                  f"(row-only passes {st['n_row_only']})", file=sys.stderr, flush=True)
        plain["mteps"] = max(plain["chains"]["mteps"], plain["no_chains"]["mteps"])

    # precision + CPU baseline on one bounded sample of the giant component
    cpu = None
    prec = None
    big = int(np.argmax([c.n for c in comps]))
    c = comps[big]
    if not args.no_cpu:
        sys.path.insert(0, ROOT)
        import oracle  # CPU baseline and reference result only
        threads = oracle.num_threads()
        k = args.cpu_sources or 48 * threads
        rng = np.random.default_rng(1234)
        sample = np.sort(rng.choice(c.n, size=min(k, c.n), replace=False)).astype(np.int32)
        tc = time.perf_counter()
        want = oracle.component_bc(c.row_ptr, c.col, c.reach, sources=sample, threads=threads)
        tc = time.perf_counter() - tc
        cpu = {"value": W * len(sample) / total_src / tc / 1e6, "unit": "MTEPS", "cores": threads,
               "kind": "port",
               "sample": f"{len(sample)} random sources of the {c.n}-vertex giant component, "
                         f"FP64 Brandes (oracle), {tc:.1f}s"}
        gb = eng.upload([c])
        try:
            got_b, _ = gb.run(subset=sample, path=bcx.PATH_BATCHED, precision=bcx.PREC_FP32C)
            got_p, _ = gb.run(subset=sample, path=bcx.PATH_PLAIN, precision=bcx.PREC_FP64)
        finally:
            gb.free()
        pos = want > 0
        rb = np.abs(got_b[pos] - want[pos]) / want[pos]
        rp = np.abs(got_p[pos] - want[pos]) / want[pos]
        prec = {"batched_fp32c_max_rel": float(rb.max()), "batched_fp32c_mean_rel": float(rb.mean()),
                "plain_fp64_max_rel": float(rp.max()),
                "zero_exact": bool(np.all(got_b[~pos] == 0) and np.all(got_p[~pos] == 0)),
                "sources": int(len(sample))}

    achieved = balg_b / tk_b / 1e9 if tk_b > 0 else None
    traffic = measured_traffic(args.config, src_b / K) if args.scale == 1.0 else None
    launches = sum(st["n_launches"] for st in st_b) / K
    tiles_run = sum(st["n_tiles"] for st in st_b)
    tiles = {"flags": args.flags, "contracted_vertices": max(st["n_contracted"] for st in st_b),
             "components": {"small": st_b[0]["n_comp_small"], "wg": st_b[0]["n_comp_wg"],
                            "multi": st_b[0]["n_comp_multi"]},
             "multi_blocks_per_cu": st_b[0]["sched_blocks_per_cu"], "multi_slots": st_b[0]["sched_slots"],
             "multi_sources_per_lane": st_b[0]["sched_sources_per_lane"],
             "multi_launch_tiles_per_step": tiles_run / K,
             "list_entries_per_tile": round(sum(st["n_list_entries"] for st in st_b) / tiles_run, 1)
             if tiles_run else None}
    line = {
        "metric": "exact-BC MTEPS on 1 MI355X: batched FP32-comp vs plain, <=1e-6 rel err",
        "value": round(mteps_b, 3),
        "unit": "MTEPS",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * t_b / K, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32 levels, u64/f64 sigma, fp32-kahan delta, f64 bc",
        "data": f"synthetic (seeded generator, workload {args.config}, scale {args.scale})",
        "config": {"workload": WORKLOADS.get(args.config, f"workload {args.config}"), "generator": args.config,
                   "n": n, "m": m, "sources": total_src, "slices": args.slices,
                   "slice_sources_timed": src_b, "gpu": arch, "cus": n_cu},
        "tiles": tiles,
        "plain": plain,
        "speedup_vs_plain": round(mteps_b / plain["mteps"], 3) if plain else None,
        "precision": prec,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1) if achieved else None,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4) if achieved else None,
                     "traffic": round(traffic) if traffic else None,
                     "traffic_unit": "HBM bytes per step (rocprofv3 FETCH_SIZE+WRITE_SIZE, scaled)",
                     # the same bytes over the same time: the memory system's actual rate
                     "frac_measured_bytes": round(traffic / (tk_b / K) / 1e9 / HBM_PEAK_GBS, 4)
                     if traffic and tk_b > 0 else None,
                     "launches_per_step": round(launches, 1),
                     "algorithmic_bytes_per_step": round(balg_b / K),
                     "kernel": "batched kernels of a step (k_b_expand, k_b_pull, k_b_back, the chain kernels; "
                               "k_small; k_wg)"},
        "cpu_baseline": cpu,
    }
    g.free()
    eng.close()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
