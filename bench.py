"""Benchmark: BC7 encode throughput on one MI355X.

Workload: BC7 (BPTC) at quality 50 (50 simulated-annealing steps, the
reference default) on one 4096x4096 RGBA8 image from the seeded generator
(seed 3).  One step = one full encode of the image, input already resident
in HBM, output written to HBM.  Prints one JSON line.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

The CPU baseline is the CPU oracle (a restatement of FasTC's BC7 encoder;
the reference's own BC7 sources need CMake-generated headers and are not
built here), timed on this host's cores on a tile of the same image.
"""
import argparse
import ctypes
import json
import os
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
KERNELS = ["bc7_scan", "bc7_rank", "bc7_select", "bc7_setup", "bc7_pack", "bc7_anneal",
           "bc7_guard"]
DEFAULT_RANKER = os.path.join(ROOT, "fastc_amd", "data", "bc7_ranker.npy")


def psnr(a, b):
    """Image::ComputePSNR (Base/src/Image.cpp:205-255): premultiplied-alpha
    RGB, peak 3 * 255^2 per pixel."""
    a = a.view(np.uint8).reshape(-1, 4).astype(np.float64)
    b = b.view(np.uint8).reshape(-1, 4).astype(np.float64)
    d = a[:, :3] * (a[:, 3:4] / 255.0) - b[:, :3] * (b[:, 3:4] / 255.0)
    mse = (d * d).sum() / a.shape[0]
    return float(10.0 * np.log10(3.0 * 255.0 * 255.0 / mse))


def headline_search(headline, full):
    return "full (every mode and selected partition, annealed)" if headline is full else \
        "learned-pruned (ranker top-k, the full search below the threshold)"
TEXEL_IN_BYTES = 4
BLOCK_OUT_BYTES = 16


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--size", type=int, default=4096)
    ap.add_argument("--quality", type=int, default=50)
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--cpu-tile", type=int, default=512,
                    help="side of the image tile the CPU baseline encodes")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-other-formats", action="store_true",
                    help="leave the DXT1/DXT5/ETC1 ring figures out of the BC7 line")
    ap.add_argument("--no-batch", action="store_true",
                    help="leave the 512-texture batch figures out of the BC7 line")
    ap.add_argument("--no-learned", action="store_true",
                    help="time only the full search")
    ap.add_argument("--ranker", default=DEFAULT_RANKER, help="trained ranker weights (.npy)")
    ap.add_argument("--topk", type=int, default=4)
    ap.add_argument("--threshold", type=float, default=0.6)
    ap.add_argument("--workload", default="bc7",
                    choices=["bc7", "batch", "dxt1", "dxt5", "etc1", "decode", "stats"],
                    help="bc7: the headline BC7 q=50 4096^2 line (default); batch: the "
                         "512-texture mixed batch; dxt1/dxt5/etc1: HBM-bound encodes; "
                         "decode: decoders and fused error reductions, all formats; "
                         "stats: BC7 CompressWithStats (statistics search and records)")
    ap.add_argument("--textures", type=int, default=512)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a device-to-device copy of 512 MiB (3 times), the "
                         "known-size kernel that scales rocprofv3 FETCH_SIZE/WRITE_SIZE")
    return ap.parse_args()


def copy_calibration():
    import torch
    n = 512 << 20
    src = torch.ones(n, dtype=torch.uint8, device="cuda:0")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_per_launch": n, "launches": 3}))


def host_info():
    """Cores this process may use and the CPU model (from /proc/cpuinfo)."""
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": cores, "cpu_model": model}


def _bands(rows, parts):
    return [(i * rows // parts, (i + 1) * rows // parts) for i in range(parts)]


def cpu_baseline(img, quality, tile, seed, runs=5):
    """The CPU oracle (the reference BC7 algorithm with the reference's LCG
    annealing stream, built -O3 like the reference's Release) on the top-left
    tile x tile region, split into row bands over every host core (the oracle
    releases the GIL).  `runs` runs with different LCG seeds, since the
    reference seeds its stream with time(NULL): throughput from the mean
    wall time, PSNR as mean and spread, under both decoders."""
    from oracle import pyoracle
    pyoracle.lib()
    host = host_info()
    cores = host["nproc"]
    sub = np.ascontiguousarray(img[:tile, :tile])
    bands = _bands(tile // 4, cores)
    walls, ps, ps_spec = [], [], []
    out = None
    for run in range(runs):
        out = np.zeros((tile // 4) * (tile // 4) * 16, dtype=np.uint8)

        def work(k):
            r0, r1 = bands[k]
            if r1 > r0:
                pyoracle.bc7_compress(sub, quality=quality, rng=pyoracle.RNG_LCG,
                                      lcg_seed=1000003 * (run + 1) + k,
                                      rect=(0, 4 * r0, tile, 4 * r1), out=out)

        t0 = time.perf_counter()
        th = [threading.Thread(target=work, args=(k,)) for k in range(cores)]
        for t in th:
            t.start()
        for t in th:
            t.join()
        walls.append(time.perf_counter() - t0)
        ps.append(psnr(sub, pyoracle.bc7_decompress(out, tile, tile)))
        ps_spec.append(psnr(sub, pyoracle.bc7_decompress_spec(out, tile, tile)))
        print(f"[bench] CPU BC7 run {run + 1}/{runs}: {walls[-1]:.1f} s", file=sys.stderr,
              flush=True)
    texels = tile * tile
    return {
        "value": texels / float(np.mean(walls)) / 1e6,
        "unit": "Mtexels/s",
        "cores": cores,
        "kind": "port",
        "sample": f"BC7 q={quality} on the top-left {tile}x{tile} tile of the {img.shape[1]}x"
                  f"{img.shape[0]} seed-{seed} image ({texels} texels), {runs} runs with "
                  f"different reference LCG seeds, {cores} threads (row bands), mean wall "
                  f"{np.mean(walls):.2f} s",
        "host": host,
        "wall_s": walls,
        "psnr_db": float(np.mean(ps)), "psnr_db_sd": float(np.std(ps)),
        "psnr_spec_db": float(np.mean(ps_spec)), "psnr_spec_db_sd": float(np.std(ps_spec)),
    }


def cpu_reference_simple(fmt, img, runs=3):
    """The reference's own stb_dxt (DXT1/DXT5) or rg_etc1 (ETC1) encoder,
    compiled from the FasTC sources (oracle/_ref), on one image split into row
    bands over every host core; median of `runs` runs."""
    from oracle import pyoracle
    L = pyoracle.ref_lib()
    if L is None:
        return None
    host = host_info()
    cores = host["nproc"]
    h, w = img.shape
    bb = 8 if fmt in (0, 2) else 16
    out = np.zeros((h // 4) * (w // 4) * bb, np.uint8)
    if fmt != 2:
        # stb_dxt builds its tables on first use behind an unsynchronised
        # flag: one call before the threads start
        L.ref_dxt_compress(np.zeros(16, np.uint32).ctypes.data, 4, 4, 1, out.ctypes.data)
    bands = _bands(h // 4, cores)

    def work(k):
        r0, r1 = bands[k]
        if r1 <= r0:
            return
        src = img[4 * r0:4 * r1].ctypes.data
        dst = out[r0 * (w // 4) * bb:].ctypes.data
        if fmt == 2:
            L.ref_etc1_compress(src, w, 4 * (r1 - r0), dst)
        else:
            L.ref_dxt_compress(src, w, 4 * (r1 - r0), int(fmt == 1), dst)

    walls = []
    for _ in range(runs):
        t0 = time.perf_counter()
        th = [threading.Thread(target=work, args=(k,)) for k in range(cores)]
        for t in th:
            t.start()
        for t in th:
            t.join()
        walls.append(time.perf_counter() - t0)
    wall = float(np.median(walls))
    texels = h * w
    label = {0: "DXT1 (stb_dxt)", 1: "DXT5 (stb_dxt)", 2: "ETC1 (rg_etc1, low quality)"}[fmt]
    return {"value": texels / wall / 1e6, "unit": "Mtexels/s", "cores": cores,
            "kind": "reference",
            "sample": f"{label} built from the reference sources, one {w}x{h} seed-2 image "
                      f"with hard edges, {cores} threads (row bands), median of {runs} runs "
                      f"({wall * 1e3:.1f} ms)",
            "host": host, "gbs": (texels * 4 + texels // 16 * bb) / wall / 1e9}


def batch_leg(args, per_surface_steps=None, batch_steps=None):
    """Texture-batch workload: 512 textures of power-of-two sides 64..2048
    (log-uniform per axis), 40% RGB-only, full mip chains, 60% BPTC q=50 /
    25% DXT5 / 15% ETC1, seed 4.  The batch (one graph launch, autotuned
    launch shapes) against one library call per surface on the same stream.
    Returns the JSON object of the workload."""
    import torch
    import fastc_amd
    from fastc_amd import batch_workload as bw

    fastc_amd.lib()
    t0 = time.perf_counter()
    tex = bw.make_batch(count=args.textures, seed=4)
    texels = bw.texel_count(tex)
    texels_by_format = {f: bw.texel_count([t for t in tex if t["format"] == f])
                        for f in (1, 2, 3)}
    print(f"[bench] batch of {len(tex)} textures, {texels} texels "
          f"({time.perf_counter() - t0:.1f} s)", file=sys.stderr, flush=True)
    surfs, keep = [], []
    for t in tex:
        for lv in t["levels"]:
            h, w = lv.shape
            d_in = torch.from_numpy(lv.view(np.int32)).cuda()
            d_out = torch.zeros(fastc_amd.compressed_size(t["format"], w, h), dtype=torch.uint8,
                                device="cuda")
            keep.append((d_in, d_out))
            surfs.append((t["format"], d_in.data_ptr(), d_out.data_ptr(), w, h))
    algo_bytes = sum(lv.size * 4 + keep_i[1].numel()
                     for lv, keep_i in zip([lv for t in tex for lv in t["levels"]], keep))
    stream = torch.cuda.current_stream()
    sh = stream.cuda_stream
    q = args.quality

    def timed(fn, steps, label):
        fn()
        torch.cuda.synchronize()
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        ev0.record(stream)
        for i in range(steps):
            fn()
            print(f"[bench] {label} step {i + 1}/{steps}", file=sys.stderr, flush=True)
        ev1.record(stream)
        torch.cuda.synchronize()
        return ev0.elapsed_time(ev1) / steps

    def per_surface():
        for fmt, pin, pout, w, h in surfs:
            fastc_amd.compress_device(pin, pout, fastc_amd.compressed_size(fmt, w, h), w, h,
                                      fmt=fmt, quality=q, seed=args.seed, stream=sh)

    ms_single = timed(per_surface, per_surface_steps or args.steps, "per-surface")
    # autotune on the first textures' surfaces, then the whole batch
    nsub = sum(len(t["levels"]) for t in tex[:max(1, len(tex) // 8)])
    sub = fastc_amd.Batch(surfs[:nsub], quality=q, seed=args.seed)
    tune = sub.autotune(stream=sh)
    del sub
    batch = fastc_amd.Batch(surfs, quality=q, seed=args.seed)
    batch.set_tune(tune)
    ms_batch = timed(lambda: batch.run(use_graph=True, stream=sh), batch_steps or args.steps,
                     "batch graph")
    blocks = {name: batch.blocks(f) for name, f in (("dxt5", 1), ("etc1", 2), ("bptc", 3))}
    line = {
        "metric": "texture batch Mtexels/s (512 textures with mips, BPTC q=50 / DXT5 / ETC1)",
        "value": texels / (ms_batch / 1e3) / 1e6, "unit": "Mtexels/s", "n_gpus": 1,
        "steps": args.steps, "ms_per_step": ms_batch, "higher_is_better": True,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"{len(tex)} textures, seed 4, full mip chains, "
                               f"{len(surfs)} surfaces", "texels": texels, "blocks": blocks,
                   "search": "full (reference BoxSelection search)"},
        "per_surface_launches": {"value": texels / (ms_single / 1e3) / 1e6,
                                 "ms_per_step": ms_single, "calls": len(surfs)},
        "speedup_vs_per_surface": ms_single / ms_batch,
        "autotune": tune,
        "roofline": {"bound": "hbm", "achieved": algo_bytes / (ms_batch / 1e3) / 1e9,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": algo_bytes / (ms_batch / 1e3) / 1e9 / HBM_PEAK_GBS,
                     "traffic": None, "kernel": "whole batch",
                     "algorithmic_bytes": algo_bytes},
        "formats_texels": {name: int(texels_by_format[f]) for name, f in
                           (("bptc", 3), ("dxt5", 1), ("etc1", 2))},
    }
    return line


def run_batch(args):
    print(json.dumps(batch_leg(args)), flush=True)


def run_decode(args):
    """Decoders and fused error reductions on 4096x4096 surfaces of every
    format (the seed-3 image encoded at quality 0).  HBM-bound: algorithmic
    bytes are the encoded surface plus 4 B per texel (decoded image written,
    or original image read), reported against 8 TB/s."""
    import torch
    import fastc_amd
    from fastc_amd import synth

    fastc_amd.lib()
    n = args.size
    img = synth.generate(n, n, args.seed)
    d_img = torch.from_numpy(img.view(np.int32)).cuda()
    dec = torch.zeros(n * n, dtype=torch.int32, device="cuda")
    blk = torch.zeros(n * n // 16, dtype=torch.int64, device="cuda")
    sse = torch.zeros(2, dtype=torch.int64, device="cuda")
    stream = torch.cuda.current_stream()
    sh = stream.cuda_stream
    steps = max(args.steps, 20)
    texels = n * n
    results = {}
    for label, fmt, bb in (("dxt1", 0, 8), ("dxt5", 1, 16), ("etc1", 2, 8), ("bptc", 3, 16)):
        enc = torch.zeros(texels // 16 * bb, dtype=torch.uint8, device="cuda")
        fastc_amd.compress_device(d_img.data_ptr(), enc.data_ptr(), enc.numel(), n, n, fmt=fmt,
                                  stream=sh)
        ops = {
            "decode": lambda: fastc_amd.decompress_device(enc.data_ptr(), dec.data_ptr(), n, n,
                                                          fmt=fmt, stream=sh),
            "block_error": lambda: fastc_amd.block_error_device(
                d_img.data_ptr(), enc.data_ptr(), n, n, sse.data_ptr(), blk.data_ptr(),
                fmt=fmt, stream=sh),
        }
        res = {}
        for op, fn in ops.items():
            fn()
            torch.cuda.synchronize()
            fastc_amd.set_kernel_timing(True)
            for _ in range(steps):
                fn()
            tot, launches = fastc_amd.kernel_time(op)
            fastc_amd.set_kernel_timing(False)
            ms = tot / launches
            algo = texels * 4 + enc.numel() + (texels // 16 * 8 if op == "block_error" else 0)
            res[op] = {"kernel_avg_ms": ms, "algorithmic_bytes": algo,
                       "achieved_gbs": algo / (ms / 1e3) / 1e9,
                       "frac": algo / (ms / 1e3) / 1e9 / HBM_PEAK_GBS}
        s = sse.cpu().numpy().view(np.uint64)
        res["psnr_db"] = fastc_amd.psnr_from_sse(s, n, n)
        results[label] = res
        print(f"[bench] decode {label} done", file=sys.stderr, flush=True)
    head = results["bptc"]["block_error"]
    line = {
        "metric": "fused BC7 decode+error GB/s (4096x4096)",
        "value": head["achieved_gbs"], "unit": "GB/s", "n_gpus": 1, "steps": steps,
        "ms_per_step": head["kernel_avg_ms"], "higher_is_better": True, "dtype": "int32",
        "data": "synthetic",
        "config": {"workload": f"decoders and error reductions, {n}x{n}, seed {args.seed}, "
                               "quality-0 encodes", "texels": texels},
        "formats": results,
        "roofline": {"bound": "hbm", "achieved": head["achieved_gbs"], "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": head["frac"], "traffic": None,
                     "kernel": "block_error (BPTC)",
                     "algorithmic_bytes": head["algorithmic_bytes"]},
    }
    print(json.dumps(line), flush=True)


def run_stats(args):
    """BPTCC::CompressWithStats on the GPU (fh_compress_with_stats_device) on
    the headline image (4096x4096, q=50, seed 3): the statistics shape scan,
    CompressClusters over its candidates (modes 4 and 5 on opaque blocks too,
    each single-subset mode once per shape entry), annealing, packing and one
    1168-byte BlockStat record per block, all device-resident."""
    import torch
    import fastc_amd
    from fastc_amd import synth

    fastc_amd.lib()
    n = args.size
    img = synth.generate(n, n, args.seed)
    d_img = torch.from_numpy(img.view(np.int32)).cuda()
    nb = n * n // 16
    enc = torch.zeros(n * n, dtype=torch.uint8, device="cuda")
    rec = torch.zeros(nb * fastc_amd.BLOCK_STATS_DTYPE.itemsize, dtype=torch.uint8, device="cuda")
    stream = torch.cuda.current_stream()
    sh = stream.cuda_stream

    def step():
        fastc_amd.compress_with_stats_device(d_img.data_ptr(), enc.data_ptr(), enc.numel(),
                                             rec.data_ptr(), nb, n, n, quality=args.quality,
                                             seed=args.seed, stream=sh)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(args.steps):
        step()
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / args.steps
    print("[bench] stats timed", file=sys.stderr, flush=True)
    fastc_amd.set_kernel_timing(True)
    step()
    torch.cuda.synchronize()
    kernels = {}
    for k in ("bc7_scan", "bc7_select", "bc7_setup", "bc7_anneal", "bc7_pack", "bc7_guard"):
        tot, launches = fastc_amd.kernel_time(k)
        if launches:
            kernels[k] = tot
    fastc_amd.set_kernel_timing(False)
    blk = torch.zeros(nb, dtype=torch.int64, device="cuda")
    sse = torch.zeros(2, dtype=torch.int64, device="cuda")
    fastc_amd.block_error_device(d_img.data_ptr(), enc.data_ptr(), n, n, sse.data_ptr(),
                                 blk.data_ptr(), fmt=3, stream=sh)
    st = rec.cpu().numpy().view(fastc_amd.BLOCK_STATS_DTYPE)
    psnr_db = fastc_amd.psnr_from_sse(sse.cpu().numpy().view(np.uint64), n, n)
    line = {
        "metric": "BC7 CompressWithStats Mtexels/s (4096², q=50)",
        "value": n * n / (ms / 1e3) / 1e6, "unit": "Mtexels/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"BC7 CompressWithStats q={args.quality}, {n}x{n} RGBA8, "
                               f"seed {args.seed}, records on the device"},
        "psnr_db": psnr_db, "kernels_ms": kernels,
        "records": {"bytes": int(rec.numel()),
                    "path": np.bincount(st["path"], minlength=4).tolist(),
                    "mode": np.bincount(st["mode"], minlength=9).tolist(),
                    "mean_shape_lines": float((st["num_two_shape"] +
                                               st["num_three_shape"]).mean())},
    }
    print(json.dumps(line), flush=True)


def bc7_q0_leg(with_cpu=True, steps=50, warmup=5):
    """BC7 at quality 0 on the 256x256 seed-1 image, the reference's
    single-thread configuration (CompressImageData, BPTC, iQuality 0, -t 1):
    the GPU encode with the input resident in HBM (4096 blocks: launch-bound,
    a parity configuration rather than a throughput one), its bytes against
    the oracle's, PSNR under both decoders, and the oracle on one thread."""
    import torch
    import fastc_amd
    from fastc_amd import synth
    W = H = 256
    img = synth.generate(W, H, 1)
    out_bytes = W * H
    dev = torch.device("cuda:0")
    d_in = torch.from_numpy(img.view(np.int32)).to(dev)
    d_out = torch.zeros(out_bytes, dtype=torch.uint8, device=dev)
    stream = torch.cuda.current_stream()

    def step():
        fastc_amd.compress_device(d_in.data_ptr(), d_out.data_ptr(), out_bytes, W, H,
                                  quality=0, stream=stream.cuda_stream)

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    ev0.record(stream)
    for _ in range(steps):
        step()
    ev1.record(stream)
    torch.cuda.synchronize()
    ms = ev0.elapsed_time(ev1) / steps
    # bytes of one encode with the reference's initial watermark counter
    got = fastc_amd.compress(img, quality=0, wm_counter=ctypes.c_uint32(0xFFFFFFFF))
    leg = {"workload": "BC7 encode q=0, 256x256 RGBA8, seed 1", "ms_per_step": ms,
           "steps": steps, "value": W * H / (ms / 1e3) / 1e6, "unit": "Mtexels/s",
           "psnr_db": psnr(img, fastc_amd.decompress(got, W, H)),
           "psnr_spec_db": psnr(img, fastc_amd.decompress(
               got, W, H, fmt=fastc_amd.eCompressionFormat_BPTC | fastc_amd.DECODE_SPEC))}
    if with_cpu:
        from oracle import pyoracle
        pyoracle.lib()
        walls = []
        ref = None
        for _ in range(3):
            t0 = time.perf_counter()
            ref, _ = pyoracle.bc7_compress(img, quality=0)
            walls.append(time.perf_counter() - t0)
        wall = float(np.median(walls))
        leg["bit_exact_vs_oracle"] = bool(np.array_equal(got, ref))
        leg["cpu_baseline"] = {
            "value": W * H / wall / 1e6, "unit": "Mtexels/s", "cores": 1, "kind": "port",
            "sample": f"the whole 256x256 image at q=0 on one thread (the reference's -t 1), "
                      f"median of 3 runs ({wall * 1e3:.1f} ms)",
            "host": host_info()}
    return leg


def simple_leg(args, fmt, with_cpu=True):
    """DXT1 / DXT5 / ETC1 on a ring of 16 distinct 2048x2048 images (seeds
    2..17, 5% hard edges): one batch graph of 16 surfaces, so the working set
    (256 MiB of input) exceeds the last-level cache.  HBM-bound: reported as
    GB/s of algorithmic traffic against 8 TB/s, next to the reference's own
    encoder on the host cores.  Returns the JSON object of the workload."""
    import torch
    import fastc_amd
    from fastc_amd import synth

    fastc_amd.lib()
    n, side = 16, 2048
    bb = 8 if fmt in (0, 2) else 16
    surfs, keep = [], []
    for i in range(n):
        img = synth.generate(side, side, 2 + i, edges=True)
        d_in = torch.from_numpy(img.view(np.int32)).cuda()
        d_out = torch.zeros(side * side // 16 * bb, dtype=torch.uint8, device="cuda")
        keep.append((d_in, d_out))
        surfs.append((fmt, d_in.data_ptr(), d_out.data_ptr(), side, side))
        if i == 0:
            first = img
    stream = torch.cuda.current_stream()
    sh = stream.cuda_stream
    batch = fastc_amd.Batch(surfs)
    tune = batch.autotune(stream=sh)
    batch.run(use_graph=True, stream=sh)
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    steps = max(args.steps, 20)
    fastc_amd.set_kernel_timing(True)
    ev0.record(stream)
    for _ in range(steps):
        batch.run(use_graph=True, stream=sh)
    ev1.record(stream)
    torch.cuda.synchronize()
    ms = ev0.elapsed_time(ev1) / steps
    name = {0: "dxt", 1: "dxt", 2: "etc1"}[fmt]
    fastc_amd.set_kernel_timing(False)
    texels = n * side * side
    algo = texels * 4 + texels // 16 * bb
    label = {0: "DXT1", 1: "DXT5", 2: "ETC1"}[fmt]
    line = {
        "metric": f"{label} encode GB/s (16 x 2048x2048, ring larger than the LLC)",
        "value": algo / (ms / 1e3) / 1e9, "unit": "GB/s", "n_gpus": 1, "steps": steps,
        "ms_per_step": ms, "higher_is_better": True, "dtype": "int32", "data": "synthetic",
        "config": {"workload": f"{label}, 16 distinct 2048x2048 images, seeds 2-17, hard edges",
                   "texels": texels},
        "mtexels_per_s": texels / (ms / 1e3) / 1e6,
        "autotune": tune,
        "roofline": {"bound": "hbm", "achieved": algo / (ms / 1e3) / 1e9, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": algo / (ms / 1e3) / 1e9 / HBM_PEAK_GBS,
                     "traffic": None, "kernel": f"{name} batch kernel",
                     "algorithmic_bytes": algo},
    }
    if with_cpu:
        cb = cpu_reference_simple(fmt, first)
        if cb is not None:
            line["cpu_baseline"] = cb
    return line


def run_simple(args, fmt):
    print(json.dumps(simple_leg(args, fmt)), flush=True)


def main():
    args = parse()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build targets one GPU; --gpus must be 1",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    if args.workload != "bc7":
        if args.workload == "batch":
            run_batch(args)
        elif args.workload == "decode":
            run_decode(args)
        elif args.workload == "stats":
            run_stats(args)
        else:
            run_simple(args, {"dxt1": 0, "dxt5": 1, "etc1": 2}[args.workload])
        return
    import torch
    import fastc_amd
    from fastc_amd import synth

    fastc_amd.lib()
    if not torch.cuda.is_available() or not fastc_amd.device_available():
        print(json.dumps({"error": "no HIP device"}))
        sys.exit(1)
    W = H = args.size
    img = synth.generate(W, H, args.seed)
    texels = W * H
    nblocks = texels // 16
    out_bytes = nblocks * BLOCK_OUT_BYTES
    dev = torch.device("cuda:0")
    d_in = torch.from_numpy(img.view(np.int32)).to(dev)
    d_out = torch.zeros(out_bytes, dtype=torch.uint8, device=dev)
    stream = torch.cuda.current_stream()
    sh = stream.cuda_stream
    sse = torch.zeros(2, dtype=torch.int64, device=dev)
    blk_err = torch.zeros(nblocks, dtype=torch.int64, device=dev)

    ranker = None
    if not args.no_learned and os.path.exists(args.ranker):
        ranker = fastc_amd.Ranker()
        ranker.set_params(np.load(args.ranker))

    def image_error(fmt):
        """Per-block premultiplied errors and PSNR of the encoded surface,
        decoded on the device (fmt: the reference decoder, or with
        DECODE_SPEC the BC7 specification's)."""
        fastc_amd.block_error_device(d_in.data_ptr(), d_out.data_ptr(), W, H, sse.data_ptr(),
                                     blk_err.data_ptr(), fmt=fmt, stream=sh)
        torch.cuda.synchronize()
        s = sse.cpu().numpy().view(np.uint64)
        be = blk_err.cpu().numpy().view(np.uint64).reshape(H // 4, W // 4)
        return fastc_amd.psnr_from_sse(s, W, H), be

    def run(rk, label, threshold):
        def step():
            fastc_amd.compress_device(d_in.data_ptr(), d_out.data_ptr(), out_bytes, W, H,
                                      quality=args.quality, seed=args.seed, stream=sh,
                                      ranker=rk, topk=args.topk, threshold=threshold)

        for i in range(args.warmup):
            step()
            torch.cuda.synchronize()
            print(f"[bench] {label} warmup {i + 1}/{args.warmup} done", file=sys.stderr,
                  flush=True)
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        # the library brackets each of its kernels with HIP events on `stream`
        # (the stream it launches on) for the per-kernel durations below
        fastc_amd.set_kernel_timing(True)
        t0 = time.perf_counter()
        ev0.record(stream)
        for i in range(args.steps):
            step()
            print(f"[bench] {label} step {i + 1}/{args.steps} queued", file=sys.stderr,
                  flush=True)
        ev1.record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        ms = ev0.elapsed_time(ev1) / args.steps
        kernels = {}
        for name in KERNELS:
            tot, n = fastc_amd.kernel_time(name)
            if n:
                kernels[name] = {"avg_ms": tot / n, "launches": n,
                                 "ms_per_step": tot / args.steps}
        fastc_amd.set_kernel_timing(False)
        p_ref, be = image_error(fastc_amd.eCompressionFormat_BPTC)
        p_spec, be_spec = image_error(fastc_amd.eCompressionFormat_BPTC | fastc_amd.DECODE_SPEC)
        return {"ms_per_step": ms, "value": texels / (ms / 1e3) / 1e6, "kernels": kernels,
                "psnr_db": p_ref, "psnr_spec_db": p_spec, "wall_s": wall,
                "block_errors": be, "block_errors_spec": be_spec}

    full = run(None, "full search", args.threshold)
    learned = None
    confident = None
    if ranker is not None:
        learned = run(ranker, "learned", args.threshold)
        # the share of searched blocks whose top softmax probability reaches
        # the threshold (the rest run the full search)
        blocks = img.reshape(H // 4, 4, W // 4, 4).transpose(0, 2, 1, 3).reshape(-1, 16)
        searched = ~(blocks == blocks[:, :1]).all(1) & ((blocks >> 24) != 0).any(1)
        sb = np.ascontiguousarray(blocks[searched])
        d_b = torch.from_numpy(sb.view(np.int32)).to(dev)
        z = torch.empty((sb.shape[0], 512), dtype=torch.float32, device=dev)
        ranker.forward(d_b.data_ptr(), sb.shape[0], z.data_ptr(), sh)
        valid = torch.zeros(512, dtype=torch.bool, device=dev)
        for m, k in enumerate([16, 64, 64, 64, 8, 4, 1, 64]):
            valid[64 * m:64 * m + k] = True
        z = torch.where(valid, z, torch.full_like(z, -float("inf")))
        confident = float((torch.softmax(z, 1).max(1).values >= args.threshold).float().mean())
        del z, d_b
    headline = full
    if learned is not None and learned["psnr_db"] >= full["psnr_db"] - 0.1:
        headline = learned
    kernels = headline["kernels"]
    dominant = max(kernels, key=lambda k: kernels[k]["ms_per_step"])

    # roofline of the dominant kernel: algorithmic HBM bytes of the launch
    # (4 B read + 1 B written per texel; one launch covers the whole image)
    # over its own average launch duration
    algo_bytes = texels * TEXEL_IN_BYTES + out_bytes
    dom_ms = kernels[dominant]["avg_ms"]
    achieved = algo_bytes / (dom_ms / 1e3) / 1e9
    traffic = None
    valu = None
    prof_name = os.path.join("profiles", "bc7_q50_4096_counters.json")
    prof = os.path.join(ROOT, prof_name)
    if os.path.exists(prof):
        with open(prof) as f:
            pc = json.load(f)
        if pc.get("kernel") == dominant and pc.get("search") == headline_search(headline, full):
            traffic = pc.get("hbm_bytes_per_launch")
            valu = dict(pc.get("valu") or {})
            valu["source"] = f"{prof_name}: rocprofv3 --pmc passes of this command (not this run)"
    search = headline_search(headline, full)

    def leg(r):
        return {"value": r["value"], "ms_per_step": r["ms_per_step"], "psnr_db": r["psnr_db"],
                "psnr_spec_db": r["psnr_spec_db"]}

    line = {
        "metric": "BC7 Mtexels/s (4096\u00b2, q=50); learned-pruned vs full GPU search at "
                  "\u0394PSNR\u22640.1dB",
        "value": headline["value"],
        "unit": "Mtexels/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": headline["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"BC7 encode q={args.quality}, {W}x{H} RGBA8, seed {args.seed}",
                   "search": search, "texels": texels, "blocks": nblocks},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_source": f"{prof_name} (FETCH_SIZE/WRITE_SIZE passes, copy-"
                                       "calibrated; not this run)" if traffic else None,
                     "kernel": dominant, "kernel_avg_ms": dom_ms,
                     "algorithmic_bytes": algo_bytes},
        "kernels": kernels,
        "psnr_db": headline["psnr_db"],
        "psnr_spec_db": headline["psnr_spec_db"],
        "psnr_note": "psnr_db: blocks decoded by the reference decoder (BPTCC::Decompress); "
                     "psnr_spec_db: by the BC7 specification, which differs for mode-4 "
                     "index-mode-1 blocks",
        "full_search": dict(leg(full), kernels=full["kernels"]),
        "wall_s": headline["wall_s"],
    }
    if learned is not None:
        line["learned"] = dict(leg(learned), topk=args.topk, threshold=args.threshold,
                               confident_share=confident,
                               fallback="the full search (below the threshold)",
                               speedup_vs_full=full["ms_per_step"] / learned["ms_per_step"],
                               delta_psnr_db=learned["psnr_db"] - full["psnr_db"],
                               delta_psnr_spec_db=learned["psnr_spec_db"] - full["psnr_spec_db"],
                               weights=os.path.relpath(args.ranker, ROOT))
    if valu is not None:
        line["valu"] = valu
    # useful-work count of the annealing (one extra encode per search, not
    # timed): the reference's QuantizedError point evaluations for this
    # encode and the GPU's pixel scores, with VALU instructions per
    # evaluation from the committed counter passes
    def counted(rk):
        fastc_amd.set_work_counting(True)
        try:
            fastc_amd.compress_device(d_in.data_ptr(), d_out.data_ptr(), out_bytes, W, H,
                                      quality=args.quality, seed=args.seed, stream=sh, ranker=rk,
                                      topk=args.topk, threshold=args.threshold)
            torch.cuda.synchronize()
            return fastc_amd.work_counts()
        finally:
            fastc_amd.set_work_counting(False)

    pc_valu = {}
    if os.path.exists(prof):
        with open(prof) as f:
            pcj = json.load(f)
        pc_valu["full"] = (pcj.get("full_search_anneal") or {}).get("SQ_INSTS_VALU")
        pc_valu["learned"] = (pcj.get("valu") or {}).get("SQ_INSTS_VALU")
    op = {"definition": "anneal_point_evals: QuantizedError point evaluations the reference's "
                        "annealing makes for this encode (counted on the GPU; equal to the "
                        "oracle's count on test images); gpu_pixel_scores: pixel scores the "
                        "GPU's annealing makes (16 per candidate step); valu_*: annealing "
                        "kernel SQ_INSTS_VALU from " + prof_name + " (not this run) x 64 lanes "
                        "per count"}
    for name, rk in (("full", None), ("learned", ranker)):
        if name == "learned" and ranker is None:
            continue
        n_ref, n_gpu = counted(rk)
        entry = {"anneal_point_evals": n_ref, "gpu_pixel_scores": n_gpu,
                 "point_evals_per_block": n_ref / nblocks}
        v = pc_valu.get(name)
        if v and n_ref:
            entry["valu_lane_insts_per_point_eval"] = v * 64 / n_ref
            entry["valu_lane_insts_per_pixel_score"] = v * 64 / n_gpu
        op[name] = entry
    line["op_count"] = op
    if not args.no_cpu_baseline:
        cb = cpu_baseline(img, args.quality, args.cpu_tile, args.seed)
        t = args.cpu_tile
        for key, errs in (("gpu_psnr_same_tile_db", headline["block_errors"]),
                          ("gpu_psnr_spec_same_tile_db", headline["block_errors_spec"])):
            tile_sse = int(errs[:t // 4, :t // 4].sum(dtype=np.uint64))
            cb[key] = fastc_amd.psnr_from_sse((tile_sse, 0), t, t)
        line["cpu_baseline"] = cb
    line["bc7_q0_256"] = bc7_q0_leg(with_cpu=not args.no_cpu_baseline)
    print("[bench] BC7 q=0 256x256 done", file=sys.stderr, flush=True)
    if not args.no_other_formats:
        formats = {}
        for name, fmt in (("dxt1", 0), ("dxt5", 1), ("etc1", 2)):
            formats[name] = simple_leg(args, fmt, with_cpu=not args.no_cpu_baseline)
            print(f"[bench] {name} ring done", file=sys.stderr, flush=True)
        line["formats"] = formats
    if not args.no_batch:
        b = batch_leg(args, per_surface_steps=1, batch_steps=3)
        if "cpu_baseline" in line and "formats" in line and "cpu_baseline" in \
                line["formats"]["dxt5"]:
            # CPU time of the batch's texels at the per-format CPU rates above
            rates = {"bptc": line["cpu_baseline"]["value"],
                     "dxt5": line["formats"]["dxt5"]["cpu_baseline"]["value"],
                     "etc1": line["formats"]["etc1"]["cpu_baseline"]["value"]}
            secs = sum(b["formats_texels"][k] / (rates[k] * 1e6) for k in rates)
            b["cpu_baseline"] = {
                "value": sum(b["formats_texels"].values()) / secs / 1e6, "unit": "Mtexels/s",
                "cores": line["cpu_baseline"]["cores"], "kind": "port",
                "sample": "estimated: the batch's texels of each format at the CPU rates "
                          "measured above (BC7 q=50 oracle port, stb_dxt DXT5 and rg_etc1 "
                          "ETC1 reference builds), all host cores"}
        line["texture_batch"] = b
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
