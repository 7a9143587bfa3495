"""Benchmark of the MI355X Wavelets compression hot path.

Default workload (one GPU): the batched launch -- a mixed stream of 2000
cubelets, sizes drawn from seed from {32^3, 64^3, 128^3, 256x256x64,
512x512x16}, dtypes uint8 / int16 / float32, CDF 9/7 or Haar (50/50) at
min(3, log2 dim) steps per axis, log quantizer with 256 bins, threshold
fraction 0.9, contents a smooth blob field plus noise generated on the
device.  One step is the whole device-resident batch (transform, threshold
select, quantize + histogram) with the inputs already in HBM; value is the
stream's input bytes per second.  Beside it, on the same stream:

  plain     the single-cubelet drop-in (wv_compress_device) run cubelet by
            cubelet over the first cubelets of the stream, and whether its
            bins equal the batch's;
  stages    one eager run of the batch serialised on one stream, HIP events
            between its stages: each stage's algorithmic bytes (transform
            s_in + 20, select 4, quantize + histogram 8 per sample) over its
            time.  The transform (the fused DWT pass over every group) is
            the dominant kernel: the line's roofline;
  copy      the HBM rate of a 128-bit streaming copy kernel (the measured
            ceiling beside the 8 TB/s peak);
  single    the 512^3 float32 cubelet (40 ellipsoids + N(0,20^2) noise),
            CDF 9/7 3,3,3, keep 5%, uniform 1024 bins, with its fused DWT
            pass's roofline: algorithmic bytes (read 4 B + write 4 B per
            sample) over the average pass duration from HIP events recorded
            on the pipeline's own stream, and measured HBM traffic;
  cpu       the oracle (a C restatement of the reference CPU path) on a
            bounded sample of the mixed stream, one thread and one worker
            per host core, and single-threaded on a 512^3 sub-volume.

Prints one JSON line.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
ITEMSIZE = {"uint8": 1, "int16": 2, "float32": 4}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", choices=["batched", "single"], default="batched")
    ap.add_argument("--count", type=int, default=2000, help="cubelets in the mixed stream")
    ap.add_argument("--plain-count", type=int, default=400,
                    help="cubelets of the stream timed through the plain path")
    ap.add_argument("--size", type=int, default=512, help="edge of the single cubelet")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="first run a plain device copy of 4 N bytes (N = samples of the "
                         "single cubelet) a few times: the reference kernel for scaling "
                         "FETCH_SIZE / WRITE_SIZE counters")
    ap.add_argument("--extras", default="e2e,serving,split,opt",
                    help="comma list: e2e (single-cubelet one-call), serving, split "
                         "(the single cubelet's volume split into overlapping cubelets, "
                         "all compressed by the batched path), opt (the -opt sweep on "
                         "the 2048^2 u8 image)")
    ap.add_argument("--split-cube", type=int, default=64, help="split: cubelet edge")
    ap.add_argument("--split-overlap", type=int, default=8, help="split: overlap")
    ap.add_argument("--serve-count", type=int, default=300, help="requests of the serving run")
    ap.add_argument("--serve-slots", type=int, default=8, help="serving device workers")
    ap.add_argument("--serve-edge", type=int, default=0,
                    help="serving: equal int16 cubes of this edge instead of the lognormal mix")
    ap.add_argument("--serve-load", type=float, default=0.7,
                    help="offered rate as a fraction of the measured saturation rate")
    return ap.parse_args()


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(vol_sample):
    """The oracle (a single-threaded C restatement of the reference CPU
    path) on a bounded sample: DWT + threshold + quantize, the stages the
    GPU step covers."""
    tests = os.path.join(ROOT, "tests")
    sys.path.insert(0, tests)
    import _oracle as O
    out = O.compress(vol_sample, O.CDF97, O.UNIFORM, 1024, [3, 3, 3], 0.95,
                     want_file=False)
    r = out["res"]
    t = r.t_dwt + r.t_thresh + r.t_quant
    nbytes = vol_sample.nbytes
    return {"value": round(nbytes / t / 1e9, 4), "unit": "GB/s", "cores": 1,
            "kind": "port",
            "sample": "%dx%dx%d float32 sub-volume of the 512^3 cubelet, CDF97 3,3,3, "
                      "thresh 0.95, uniform 1024; DWT+threshold(sort)+quantize "
                      "= %.2f s single-threaded (Huffman %.2f s not counted)" % (
                          vol_sample.shape[2], vol_sample.shape[1], vol_sample.shape[0],
                          t, r.t_huff)}


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def host_cores():
    """The host cores this process may run on (the GPU box's share)."""
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def cpu_baseline_stream(cubes, target_s=12.0):
    """The oracle on a bounded sample of the mixed stream -- the first
    cubelets, copied to the host, until about target_s of single-threaded
    work -- timed two ways: one thread over the whole sample, and one worker
    per host core over disjoint cubelets of it.  Stages timed: DWT +
    threshold + quantize (the GPU step's work; the oracle's Huffman coding
    is not counted)."""
    import concurrent.futures as cf
    tests = os.path.join(ROOT, "tests")
    sys.path.insert(0, tests)
    import _oracle as O
    items, t_single, nbytes = [], 0.0, 0
    for c in cubes:
        v = c["vol"].cpu().numpy()
        item = (v.astype(np.int32) if v.dtype == np.int16 else v, c)
        r = O.compress(item[0], c["wavelet"], O.LOG, c["bins"], c["steps"], c["frac"],
                       want_file=False)["res"]
        t_single += r.t_dwt + r.t_thresh + r.t_quant
        nbytes += v.nbytes
        items.append(item)
        if t_single >= target_s:
            break
    cores = host_cores()

    def work(k):
        t = 0.0
        for v, c in items[k::cores]:
            r = O.compress(v, c["wavelet"], O.LOG, c["bins"], c["steps"], c["frac"],
                           want_file=False)["res"]
            t += r.t_dwt + r.t_thresh + r.t_quant
        return t

    t0 = time.time()
    with cf.ThreadPoolExecutor(max_workers=cores) as ex:
        per_worker = list(ex.map(work, range(cores)))
    wall = time.time() - t0
    t_multi = max(per_worker)
    model = cpu_model()
    return {"value": round(nbytes / t_multi / 1e9, 4), "unit": "GB/s", "cores": cores,
            "kind": "port",
            "sample": "the first %d cubelets of the mixed stream (%.2f GB of input, copied to "
                      "the host), DWT + threshold (sort) + log 256 quantize per cubelet; one "
                      "worker per host core over disjoint cubelets: %.2f s of stage time in the "
                      "busiest worker (%.2f s wall with Huffman)" % (
                          len(items), nbytes / 1e9, t_multi, wall),
            "single_thread": {"value": round(nbytes / t_single / 1e9, 4), "cores": 1,
                              "seconds": round(t_single, 2)},
            "cpu_model": model, "nproc": os.cpu_count(), "host_cores_used": cores}


def copy_ceiling(wv, torch, dev, nbytes=1 << 30, reps=10):
    """The HBM rate a streaming copy reaches on this GPU: wv_copy_stream
    (128-bit accesses, 16 loads in flight per thread) over 1 GiB, the best
    of its access hints x blocks per CU; bytes moved = read + write."""
    src = torch.empty(nbytes // 4, dtype=torch.float32, device=dev).uniform_()
    dst = torch.empty_like(src)
    best = None
    for nt in (0, 1):
        for vpt in (4, 8, 16):
            for bpc in (1, 2, 4, 8, 16):
                for _ in range(3):
                    wv.copy_stream(src, dst, nt, bpc, vpt)
                e0, e1 = wv.Event(), wv.Event()
                e0.record()
                for _ in range(reps):
                    wv.copy_stream(src, dst, nt, bpc, vpt)
                e1.record()
                torch.cuda.synchronize()
                ms = e0.elapsed_ms(e1) / reps
                gbps = 2.0 * nbytes / (ms * 1e-3) / 1e9
                if best is None or gbps > best[0]:
                    best = (gbps, nt, bpc, vpt, ms)
    ok = bool(torch.equal(src, dst))
    del src, dst
    torch.cuda.empty_cache()
    gbps, nt, bpc, vpt, ms = best
    log("copy ceiling: %.1f GB/s (%s, %d x 16 B per thread, %d blocks per CU)" % (
        gbps, "non-temporal" if nt else "plain", vpt, bpc))
    return {"gbps": round(gbps, 1), "frac_of_peak": round(gbps / HBM_PEAK_GBS, 4),
            "ms": round(ms, 4), "bytes_copied": nbytes, "nontemporal": bool(nt),
            "blocks_per_cu": bpc, "vec_per_thread": vpt, "copy_verified": ok,
            "kernel": "wv_copy_stream (copy_kernel: 4, 8 or 16 x 16-byte loads in flight per "
                      "thread; the best of hints x depth x blocks per CU)"}


def measured_traffic():
    """HBM bytes per DWT pass launch on the 512^3 cubelet, from the two
    rocprofv3 --pmc runs (FETCH_SIZE, WRITE_SIZE) scaled by a copy of known
    size, recorded with their commands in profiles/."""
    import re
    path = os.path.join(ROOT, "profiles", "dwt_pass_traffic_512cube_cdf97.md")
    try:
        m = re.search(r"traffic_bytes_per_launch = (\d+)", open(path).read())
        return int(m.group(1)) if m else None
    except OSError:
        return None


def single_cubelet(args, wv, torch, dev, steps, warmup):
    """The 512^3 single-cubelet device pipeline and the DWT pass roofline."""
    from wavelets_amd import synth
    L = wv.lib()
    n = args.size
    t0 = time.time()
    vol = synth.ellipsoids_f32(n=n, seed=2)
    log("generated %d^3 float32 in %.1f s" % (n, time.time() - t0))
    vol_d = torch.from_numpy(vol).to(dev)
    N = vol.size
    if args.calibrate_copy:
        dst = torch.empty_like(vol_d)
        for _ in range(3):
            dst.copy_(vol_d)        # reads 4 N bytes, writes 4 N bytes
        torch.cuda.synchronize()
        del dst
    g = wv.geom([n, n, n], [3, 3, 3], wv.CDF97, wv.FLOAT32)
    coef = torch.empty(g.count, dtype=torch.float32, device=dev)
    ws = torch.empty(max(g.ws_floats, L.wv_threshold_ws_bytes(g.count) // 4 + 1),
                     dtype=torch.float32, device=dev)
    res_d = torch.zeros(64, dtype=torch.uint8, device=dev)
    bins = torch.empty(g.count, dtype=torch.int32, device=dev)
    counts = torch.empty(1024 + 128, dtype=torch.int32, device=dev)
    stream = torch.cuda.current_stream()
    sp = C.c_void_p(stream.cuda_stream)
    npass = L.wv_dwt3d_num_passes(C.byref(g))

    def step(evs):
        arr = (C.c_void_p * (npass + 1))(*[e.h for e in evs])
        wv.check(L.wv_dwt3d_forward_ev(C.byref(g), C.c_void_p(vol_d.data_ptr()),
                                       C.c_void_p(coef.data_ptr()), C.c_void_p(ws.data_ptr()),
                                       None, arr, sp))
        wv.check(L.wv_threshold(C.c_void_p(coef.data_ptr()), g.count, C.c_float(0.95),
                                None, C.c_void_p(ws.data_ptr()),
                                C.c_void_p(res_d.data_ptr()), sp))
        wv.check(L.wv_quantize_hist_uniform_dev(
            C.c_void_p(coef.data_ptr()), g.count, 1024, C.c_void_p(res_d.data_ptr()),
            C.c_void_p(bins.data_ptr()), C.c_void_p(counts.data_ptr()), sp))

    # W warmup steps, continued until 2 s of work have run, so that the timed
    # steps see settled clocks (the roofline figure comes from them)
    t_w = time.time()
    k = 0
    while k < warmup or time.time() - t_w < 2.0:
        step([wv.Event() for _ in range(npass + 1)])
        torch.cuda.synchronize()
        k += 1
    evsets = [[wv.Event() for _ in range(npass + 1)] for _ in range(steps)]
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    start.record(stream)
    for k in range(steps):
        step(evsets[k])
    stop.record(stream)
    torch.cuda.synchronize()
    ms_step = start.elapsed_time(stop) / steps
    pass_ms = np.zeros(npass)
    for evs in evsets:
        for i in range(npass):
            pass_ms[i] += evs[i].elapsed_ms(evs[i + 1])
    pass_ms /= steps
    pass_bytes = 8.0 * N        # read 4 B + write 4 B per sample (float32 input)
    avg_pass_ms = float(pass_ms.mean())
    achieved = pass_bytes / (avg_pass_ms * 1e-3) / 1e9
    _, r = wv.threshold(coef, 0.95)
    out = {
        "workload": "512^3 float32 cubelet (40 ellipsoids + N(0,20^2)), CDF97 3,3,3, keep 5%% "
                    "(thresh 0.95), uniform 1024 bins" if n == 512 else "%d^3 float32" % n,
        "gbps_input": round(vol.nbytes / (ms_step * 1e-3) / 1e9, 3),
        "ms_per_step": round(ms_step, 4),
        "threshold": float(r.threshold), "max_abs": float(r.max_abs),
        "select_fallback": bool(r.n_candidates < 0),
    }
    traffic = measured_traffic() if n == 512 else None
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
            "traffic_source": "profiles/dwt_pass_traffic_512cube_cdf97.md (rocprofv3 FETCH_SIZE "
                              "+ WRITE_SIZE, scaled by what they read for a 1 GiB copy kernel)"
                              if traffic else None,
            "kernel": "dwt_pass_kernel (fused axis pass) on the 512^3 cubelet",
            "pass_ms": [round(float(v), 4) for v in pass_ms],
            "algorithmic_bytes_per_launch": pass_bytes}
    if "e2e" in args.extras:
        opt = wv.Options(wv.CDF97, wv.UNIFORM, 1024, (C.c_int * 3)(3, 3, 3),
                         C.c_float(0.95), 1, 30, 1)
        orig = (C.c_int * 3)(n, n, n)
        wb = L.wv_compress_work_bytes(orig, C.byref(opt))
        work = torch.empty(wb, dtype=torch.uint8, device=dev)
        cap = g.count * 4 + (1 << 22)
        fbuf = np.empty(cap, np.uint8)
        info = wv.CompressInfo()
        tot = []
        for _ in range(6):      # the first call is a warmup
            wv.check(L.wv_compress(C.c_void_p(vol_d.data_ptr()), wv.FLOAT32, orig, C.byref(opt),
                                   C.c_void_p(work.data_ptr()), wb, None,
                                   fbuf.ctypes.data_as(C.c_void_p), cap, C.byref(info), sp))
            tot.append(info.ms_total)
        ms_e2e = float(np.median(tot[1:]))
        out["e2e_one_call"] = {
            "gbps_input": round(vol.nbytes / (ms_e2e * 1e-3) / 1e9, 3),
            "ms_total": round(ms_e2e, 3),
            "file_bytes": int(info.file_bytes),
            "payload_bytes": int(info.payload_bytes),
            "note": "wv_compress (median of 5 after a warmup): device stages, Huffman coding on "
                    "the device, payload D2H, .cube image on the host"}
        # the device coder alone on the pipeline's bins: run statistics, code
        # book on the host, payload kernels, payload D2H
        counts_full = torch.zeros(1025, dtype=torch.int32, device=dev)
        counts_full.copy_(counts[:1025])
        zb = 511
        t_enc = []
        for _ in range(6):
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            words, book = wv.encode_device(bins, counts_full, 1024, zb)
            host_words = words.cpu()
            t_enc.append((time.perf_counter() - t1) * 1e3)
            nbits = book.payload_bits
            book.close()
        out["device_huffman"] = {
            "ms": round(float(np.median(t_enc[1:])), 3),
            "payload_bytes": int(host_words.numel() * 4), "payload_bits": int(nbits),
            "note": "wv_encode_device on the 512^3 bins + payload D2H (median of 5)"}
        del work
        out["decompress"] = decompress_leg(wv, torch, dev, fbuf[:info.file_bytes].tobytes(),
                                           vol.nbytes)
    return out, roof, vol


def decompress_leg(wv, torch, dev, file_bytes, out_bytes):
    """Decompression of the 512^3 file: device-resident (payload on the device:
    parallel Huffman decode, dequantize, inverse transform, float output on
    the device) and the one call from a host file image to a host volume."""
    L = wv.lib()
    info, table, bnd, val = wv.cube_parse(file_bytes)
    nwords = info.payload_bytes // 4
    words = np.frombuffer(file_bytes, np.uint32, nwords, info.payload_offset)
    words_d = torch.from_numpy(words.view(np.int32).copy()).to(dev)
    count = info.size[0] * info.size[1] * info.size[2]
    bins_d = torch.empty(count, dtype=torch.int32, device=dev)
    dws = torch.empty(L.wv_huff_decode_ws_bytes(nwords, table.size), dtype=torch.uint8, device=dev)
    wb = L.wv_decompress_work_bytes(C.byref(info))
    work = torch.empty(wb, dtype=torch.uint8, device=dev)
    out_d = torch.empty(out_bytes // 4, dtype=torch.float32, device=dev)
    sp = wv._stream()
    rounds = C.c_int()
    t_dec, t_all = [], []
    for _ in range(6):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        wv.check(L.wv_huff_decode_device(table.ctypes.data_as(C.c_void_p), table.size,
                                         info.dup_key, info.dup_value,
                                         C.c_void_p(words_d.data_ptr()), nwords,
                                         C.c_void_p(bins_d.data_ptr()), count,
                                         C.c_void_p(dws.data_ptr()), C.byref(rounds), sp))
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        wv.check(L.wv_decompress_device(C.byref(info), C.c_void_p(bins_d.data_ptr()), None, None,
                                        C.c_void_p(work.data_ptr()), wb,
                                        C.c_void_p(out_d.data_ptr()), sp))
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        t_dec.append((t1 - t0) * 1e3)
        t_all.append((t2 - t0) * 1e3)
    ms_dec = float(np.median(t_dec[1:]))
    ms_all = float(np.median(t_all[1:]))
    host_out = np.empty(out_bytes, np.uint8)
    buf = np.frombuffer(file_bytes, np.uint8)
    t_one = []
    for _ in range(3):
        t0 = time.perf_counter()
        n = L.wv_decompress(buf.ctypes.data_as(C.c_void_p), buf.size, C.c_void_p(work.data_ptr()),
                            wb, host_out.ctypes.data_as(C.c_void_p), out_bytes, None, None, sp)
        wv.check(n if n < 0 else 0)
        t_one.append((time.perf_counter() - t0) * 1e3)
    ms_one = float(np.median(t_one[1:]))
    return {
        "device_resident_ms": round(ms_all, 3),
        "gbps_output": round(out_bytes / (ms_all * 1e-3) / 1e9, 3),
        "huffman_decode_ms": round(ms_dec, 3), "sync_rounds": rounds.value,
        "one_call_ms": round(ms_one, 3),
        "note": "512^3 float32 file (7.6 MB): device decode (incl. its host sync rounds) + "
                "dequantize + inverse CDF97 + restore, payload and output on the device; "
                "one_call = wv_decompress from a host file image to a host volume (512 MB D2H)"}


def make_stream(args, wv, torch, dev):
    """The mixed stream's cubelets, generated on the device."""
    from wavelets_amd import synth
    stream = synth.mixed_stream(args.count, seed=4)
    t0 = time.time()
    cubes, in_bytes = [], 0
    for i, (size, dt, wave) in enumerate(stream):
        vol = synth.cubelet_device(size, dt, 1000 + i, device=dev)
        in_bytes += vol.numel() * ITEMSIZE[dt]
        cubes.append(dict(vol=vol, wavelet=wave, steps=synth.stream_steps(size),
                          alg=wv.LOG, bins=256, frac=0.9))
    torch.cuda.synchronize()
    log("generated %d cubelets (%.2f GB) in %.1f s" % (len(cubes), in_bytes / 1e9,
                                                      time.time() - t0))
    return cubes, in_bytes


def stream_stage_bytes(cubes):
    """Algorithmic bytes of one submit of the stream, per stage: transform
    s_in + 4 + 8 (npass - 1) per sample (pass 1 reads the input and writes
    4 B, each further pass reads and writes 4 B), select 4 (one read),
    quantize + histogram 8."""
    t = s = q = 0
    for c in cubes:
        n = c["vol"].numel()
        npass = sum(1 for v in c["steps"] if v > 0)
        t += n * (c["vol"].element_size() + 4 + 8 * (npass - 1))
        s += n * 4
        q += n * 8
    return {"transform": t, "select": s, "quantize": q}


def mixed_stream(args, wv, torch, dev, steps, warmup):
    """The batched launch on the mixed stream, and the plain path beside it."""
    L = wv.lib()
    cubes, in_bytes = make_stream(args, wv, torch, dev)
    t0 = time.time()
    flags = 0 if args.no_autotune else wv.BATCH_AUTOTUNE
    b = wv.Batch(cubes, flags=flags)
    log("planned batch (%d groups%s) in %.1f s" % (
        len(b.groups()), "" if args.no_autotune else ", autotuned", time.time() - t0))
    for _ in range(max(1, warmup)):
        b.submit()
        b.wait()
    torch.cuda.synchronize()
    t0 = time.time()
    for _ in range(steps):
        b.submit()
    b.wait()
    torch.cuda.synchronize()
    ms_batch = (time.time() - t0) * 1e3 / steps
    batched_gbps = in_bytes / (ms_batch * 1e-3) / 1e9
    log("batched: %.3f ms per stream, %.1f GB/s" % (ms_batch, batched_gbps))

    # plain path: one cubelet at a time through the drop-in device stages
    P = min(args.plain_count, len(cubes))
    opt_of = {}
    work_bytes = 0
    for c in cubes[:P]:
        v = c["vol"]
        orig = (C.c_int * 3)(v.shape[2], v.shape[1], v.shape[0])
        opt = wv.Options(c["wavelet"], c["alg"], c["bins"], (C.c_int * 3)(*c["steps"]),
                         C.c_float(c["frac"]), 1, 30, 1)
        opt_of[id(c)] = (orig, opt)
        work_bytes = max(work_bytes, L.wv_compress_work_bytes(orig, C.byref(opt)))
    work = torch.empty(work_bytes, dtype=torch.uint8, device=dev)
    sp = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    info = wv.CompressInfo()

    def plain_one(c):
        orig, opt = opt_of[id(c)]
        wv.check(L.wv_compress_device(C.c_void_p(c["vol"].data_ptr()), wv.dtype_id(c["vol"]),
                                      orig, C.byref(opt), C.c_void_p(work.data_ptr()),
                                      work_bytes, None, C.byref(info), sp))

    for c in cubes[:min(P, 20)]:
        plain_one(c)
    torch.cuda.synchronize()
    t0 = time.time()
    for c in cubes[:P]:
        plain_one(c)
    torch.cuda.synchronize()
    ms_plain = (time.time() - t0) * 1e3
    plain_bytes = sum(c["vol"].numel() * c["vol"].element_size() for c in cubes[:P])
    plain_gbps = plain_bytes / (ms_plain * 1e-3) / 1e9
    log("plain: %d cubelets in %.1f ms, %.1f GB/s" % (P, ms_plain, plain_gbps))
    # identical results: the plain bins of a spread of cubelets against the batch's
    same = True
    for i in list(range(0, P, max(1, P // 25))):
        plain_one(cubes[i])
        torch.cuda.synchronize()
        n = b.bins[i].numel()
        pb = work[:n * 4].view(torch.int32)
        if not torch.equal(pb, b.bins[i]):
            same = False
    # the whole stream into one .cube (coding on the device, payload D2H,
    # container on the host), after a batch submit
    b.submit()
    f = b.write_stream()
    t0 = time.time()
    b.submit()
    b.wait()
    t1 = time.time()
    f = b.write_stream()
    t2 = time.time()
    ms_cube = (t2 - t1) * 1e3
    log("batched to .cube: submit %.1f ms + coding/stream %.1f ms, %d bytes"
        % ((t1 - t0) * 1e3, ms_cube, len(f)))
    # per-stage times: eager runs serialised on one stream with events
    # between the stages (median of 3), against each stage's algorithmic
    # bytes: transform s_in + 4 + 8 (npass - 1), select 4, quantize 8 per sample
    runs = [b.stage_times() for _ in range(3)]
    st = {k: float(np.median([r[k] for r in runs])) for k in runs[0]}
    samples = int(sum(c["vol"].numel() for c in cubes))
    stage_bytes = stream_stage_bytes(cubes)
    stages = {}
    for k in ("transform", "select", "quantize"):
        ms = st[k + "_ms"]
        gbps = stage_bytes[k] / (ms * 1e-3) / 1e9
        stages[k] = {"ms": round(ms, 3), "algorithmic_bytes": int(stage_bytes[k]),
                     "gbps": round(gbps, 1), "frac": round(gbps / HBM_PEAK_GBS, 4)}
    stages["transform"]["launches"] = int(st["transform_launches"])
    stream_bytes = in_bytes + 32 * samples
    log("stages (serialised): transform %.2f ms (%.2f), select %.2f ms (%.2f), quantize %.2f "
        "ms (%.2f)" % (stages["transform"]["ms"], stages["transform"]["frac"],
                       stages["select"]["ms"], stages["select"]["frac"],
                       stages["quantize"]["ms"], stages["quantize"]["frac"]))
    # select statistics: candidates examined, exact fallbacks (top bit)
    nc = [b.result(i)[0].n_candidates for i in range(len(cubes))]
    fallbacks = sum(1 for v in nc if v < 0)
    out = {
        "workload": "mixed stream: %d cubelets {32^3, 64^3, 128^3, 256x256x64, 512x512x16}, "
                    "u8/i16/f32, CDF97|Haar, min(3,log2 dim) steps, log 256 bins, "
                    "thresh 0.9" % len(cubes),
        "cubelets": len(cubes), "input_bytes": int(in_bytes),
        "samples": int(sum(c["vol"].numel() for c in cubes)),
        "batched_ms_per_stream": round(ms_batch, 3),
        "batched_gbps": round(batched_gbps, 3),
        "plain_gbps": round(plain_gbps, 3),
        "plain_cubelets_timed": P,
        "batched_over_plain": round(batched_gbps / plain_gbps, 3),
        "plain_bins_equal_batched": same,
        "groups": len(b.groups()),
        "to_cube_ms": round(ms_cube, 3),
        "to_cube_file_bytes": len(f),
        "to_cube_gbps": round(in_bytes / ((ms_batch + ms_cube) * 1e-3) / 1e9, 3),
        "to_cube_note": "batched bins (ms_per_stream) plus coding every cubelet on the device, "
                        "payload D2H and the one .cube stream on the host",
        "select_fallbacks": fallbacks,
        "select_max_candidates": max(v & 0x7fffffff for v in nc),
        "stream_algorithmic_bytes": int(stream_bytes),
        "stream_frac": round(stream_bytes / (ms_batch * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "stream_note": "(s_in + 32) B per sample over the whole submit: transform s_in + 20, "
                       "select 4, quantize + histogram 8",
        "stages": stages,
        "stages_note": "one eager run serialised on one stream (no stream overlap, no graph) "
                       "with HIP events between stages, median of 3; transform = sum of the "
                       "groups' pass-launch spans",
    }
    b.close()
    del work
    torch.cuda.empty_cache()
    return out, cubes


def split_leg(args, wv, torch, dev, vol):
    """cube split with overlap feeding the batched path: the single
    cubelet's volume (device-resident) cut into cubelets, every box read in
    place by the batched transform, coded on the device, written as one
    .cube stream."""
    vd = torch.from_numpy(np.ascontiguousarray(vol)).to(dev)
    cub = [args.split_cube] * 3
    ov = args.split_overlap
    f, info = wv.split_compress(vd, cub, ov, wv.CDF97, [3, 3, 3], wv.UNIFORM, 1024, 0.95)
    torch.cuda.synchronize()
    runs = []
    for _ in range(3):
        t0 = time.time()
        f, info = wv.split_compress(vd, cub, ov, wv.CDF97, [3, 3, 3], wv.UNIFORM, 1024, 0.95)
        runs.append(((time.time() - t0) * 1e3, info.ms_plan, info.ms_batch, info.ms_encode))
    runs.sort()
    ms, ms_plan, ms_batch, ms_enc = runs[len(runs) // 2]
    nbytes = vd.numel() * 4
    log("split: %d cubelets of %d^3 (overlap %d) in %.1f ms (batch %.1f, coding %.1f), %.1f GB/s"
        % (info.cubelets, args.split_cube, ov, ms, ms_batch, ms_enc, nbytes / ms / 1e6))
    out = {
        "workload": "%d^3 float32 volume split into %d^3 cubelets with overlap %d, CDF97 3,3,3, "
                    "thresh 0.95, uniform 1024, one .cube stream" % (
                        vol.shape[0], args.split_cube, ov),
        "cubelets": info.cubelets,
        "ms_total": round(ms, 3), "ms_batch": round(ms_batch, 3),
        "ms_coding": round(ms_enc, 3), "ms_plan": round(ms_plan, 3),
        "gbps_input": round(nbytes / ms / 1e6, 3),
        "gbps_input_batch_only": round(nbytes / ms_batch / 1e6, 3),
        "file_bytes": len(f), "ratio": round(nbytes / len(f), 2),
    }
    del vd
    torch.cuda.empty_cache()
    return out


def opt_leg(args, wv, torch, dev):
    """-opt on the 2048 x 2048 u8 image (CDF97 11,11,0): the optimisation
    data (transform + device sort) once, then the reference's default
    plotAll sweep of 50 threshold fractions x 32 log bin counts."""
    from wavelets_amd import synth
    img = synth.smooth_field_u8(2048, 2048, seed=1)      # [1, 2048, 2048]
    vd = torch.from_numpy(np.ascontiguousarray(img)).to(dev)
    o = wv.Opt(vd, [11, 11, 0], wv.CDF97)
    o.test(o.sorted(o.count() // 2), 256, wv.LOG)      # warm
    torch.cuda.synchronize()
    t0 = time.time()
    o2 = wv.Opt(vd, [11, 11, 0], wv.CDF97)
    torch.cuda.synchronize()
    ms_create = (time.time() - t0) * 1e3
    t0 = time.time()
    rows = wv.opt_plot_all(o2)
    ms_sweep = (time.time() - t0) * 1e3
    best = max(rows, key=lambda r: r[3] - 1e-6 * r[2])
    log("opt: create %.2f ms, %d points in %.1f ms (%.3f ms each)"
        % (ms_create, len(rows), ms_sweep, ms_sweep / len(rows)))
    o.close()
    o2.close()
    return {
        "workload": "2048x2048 u8 image, CDF97 11,11,0, plotAll: 50 threshold fractions x 32 "
                    "log bin counts (test point = quantize, coder size, dequantize, inverse, "
                    "error sums)",
        "ms_create": round(ms_create, 3), "points": len(rows),
        "ms_sweep": round(ms_sweep, 2), "ms_per_point": round(ms_sweep / len(rows), 4),
        "best_psnr_row": list(best),
    }


def serving(args, wv, torch, dev):
    """The serving loop on a Poisson request stream at a fixed offered rate.
    Requests: edges lognormal (median 128, sigma 0.6) clipped to [32, 512],
    depth lognormal (median 48, sigma 1.0) clipped to [1, 512], dtypes
    u8 / i16 / f32, Haar at 4,4,4 steps, Lloyd 256 bins (30 iterations),
    kept fraction uniform on [0.02, 0.2] (thresh = 1 - kept).  Latency is
    submit -> bins, counts and .cube image on the host."""
    from wavelets_amd import synth
    rng = np.random.default_rng(6)          # arrival times
    N = args.serve_count
    reqs, in_bytes = [], 0
    t0 = time.time()
    if args.serve_edge:
        # equal requests: int16 cubes of one edge, keep 10%
        e = args.serve_edge
        for i in range(N):
            vol = synth.cubelet_device((e, e, e), "int16", 7000 + i, device=dev).cpu().numpy()
            in_bytes += vol.nbytes
            reqs.append((vol, 0.9))
        mix = "%d requests, int16 %d^3 each, keep 10%%" % (N, e)
    else:
        for i, (size, dt, keep) in enumerate(synth.serving_requests(N, seed=5)):
            vol = synth.cubelet_device(size, dt, 5000 + i, device=dev).cpu().numpy()
            in_bytes += vol.nbytes
            reqs.append((vol, 1.0 - keep))
        mix = ("%d requests, edges lognormal [32,512], depth [1,512], u8/i16/f32, "
               "keep U[0.02,0.2]" % N)
    log("generated %d serving requests (%.2f GB) in %.1f s" % (N, in_bytes / 1e9, time.time() - t0))

    def run(rate, lloyd_batch=False):
        srv = wv.Server(slots=args.serve_slots, hbm_budget=200 << 30, encode=True,
                        lloyd_batch=lloyd_batch)
        arrivals = (np.cumsum(rng.exponential(1.0 / rate, N)) if rate else np.zeros(N))
        done = []
        start = time.time()
        for i, (vol, thr) in enumerate(reqs):
            wait = start + arrivals[i] - time.time()
            if wait > 0:
                time.sleep(wait)
            srv.submit(vol, wv.HAAR, [4, 4, 4], wv.LLOYD, 256, thr, lloyd_iters=30)
            if i % 16 == 0:
                done += srv.poll()
        done += srv.drain()
        st = srv.stats()
        srv.close()
        assert len(done) == N and all(d["status"] == 0 for d in done)
        lat = np.array([d["t_done"] - d["t_submit"] for d in done]) * 1e3
        span = max(d["t_done"] for d in done) - min(d["t_submit"] for d in done)
        out = dict(p50_ms=round(float(np.percentile(lat, 50)), 3),
                   p99_ms=round(float(np.percentile(lat, 99)), 3),
                   gbps=round(in_bytes / span / 1e9, 3), rps=round(N / span, 2),
                   peak_inflight=st["peak_inflight"],
                   peak_device_gb=round(st["peak_bytes"] / 1e9, 3))
        if lloyd_batch and st["lloyd_windows"]:
            out["lloyd_runs_per_window"] = round(st["lloyd_run_windows"] / st["lloyd_windows"], 2)
        return out

    sat = run(None)                    # all requests at once: capacity
    rate = args.serve_load * sat["rps"]
    fixed = run(rate)
    log("serving: saturation %.1f req/s %.2f GB/s; at %.1f req/s p50 %.2f ms p99 %.2f ms" % (
        sat["rps"], sat["gbps"], rate, fixed["p50_ms"], fixed["p99_ms"]))
    # the same stream with the Lloyd windows of the requests in flight
    # sharing launches (lloyd_batch)
    sat_b = run(None, lloyd_batch=True)
    rate_b = args.serve_load * sat_b["rps"]
    fixed_b = run(rate_b, lloyd_batch=True)
    log("serving, batched Lloyd windows: saturation %.1f req/s; at %.1f req/s p99 %.2f ms" % (
        sat_b["rps"], rate_b, fixed_b["p99_ms"]))
    return {"workload": "%s; Haar 4,4,4, Lloyd 256 (30 it); %d device workers, 200 GB HBM "
                        "budget; latency = submit -> bins+counts+.cube on host; each worker "
                        "runs its own Lloyd windows" % (mix, args.serve_slots),
            "input_bytes": int(in_bytes),
            "saturation": sat,
            "offered_rps": round(rate, 2),
            "at_offered_rate": fixed,
            "lloyd_batch": {"saturation": sat_b, "offered_rps": round(rate_b, 2),
                            "at_offered_rate": fixed_b}}


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(cubelets shard across GPUs with no exchange)",
              file=sys.stderr)
        sys.exit(2)
    import torch
    import wavelets_amd as wv
    wv.lib()
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)

    single, roof, vol = single_cubelet(args, wv, torch, dev,
                                       args.steps if args.workload == "single" else 20,
                                       args.warmup)
    copy = copy_ceiling(wv, torch, dev)
    single["roofline"] = roof
    batched = cubes = None
    if args.workload == "batched":
        batched, cubes = mixed_stream(args, wv, torch, dev, args.steps, args.warmup)

    serve = None
    if "serving" in args.extras and args.workload == "batched":
        serve = serving(args, wv, torch, dev)
    split = None
    if "split" in args.extras:
        split = split_leg(args, wv, torch, dev, vol)
    opt = None
    if "opt" in args.extras:
        opt = opt_leg(args, wv, torch, dev)

    cpu = None
    if not args.no_cpu:
        q = max(32, args.size // 2)
        sample = np.ascontiguousarray(vol[:q, :q, :])
        single["cpu_baseline"] = cpu_baseline(sample)
        cpu = cpu_baseline_stream(cubes) if cubes is not None else single["cpu_baseline"]
    del cubes
    torch.cuda.empty_cache()

    if batched is not None:
        value, ms_step = batched["batched_gbps"], batched["batched_ms_per_stream"]
        workload = batched["workload"]
        dtype = "f32"
        # the headline's dominant kernel: the fused DWT pass over every
        # group of the stream (serialised launches, HIP events on their stream)
        t = batched["stages"]["transform"]
        roof = {"bound": "hbm", "achieved": t["gbps"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": t["frac"], "traffic": None,
                "traffic_note": "not measured on the stream; the fused pass's own HBM traffic "
                                "per launch, against its algorithmic bytes, is measured on the "
                                "512^3 cubelet (single_cubelet.roofline)",
                "kernel": "dwt_pass_kernel: every transform pass of the 2000-cubelet stream "
                          "(%d launches over 30 groups, run serialised)" % t["launches"],
                "algorithmic_bytes": t["algorithmic_bytes"], "ms": t["ms"],
                "per_stage": {k: batched["stages"][k]["frac"]
                              for k in ("transform", "select", "quantize")},
                "stream_frac": batched["stream_frac"],
                "copy_ceiling_gbps": copy["gbps"],
                "frac_of_copy_ceiling": round(t["gbps"] / copy["gbps"], 4)}
    else:
        value, ms_step = single["gbps_input"], single["ms_per_step"]
        workload = single["workload"]
        dtype = "f32"
    line = {
        "metric": "input GB/s compressed, batched vs plain; p50/p99 latency at fixed req/s",
        "value": value,
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype,
        "data": "synthetic (seeded; generated on the host / device, inputs resident in HBM)",
        "config": {"workload": workload},
        "roofline": roof,
        "cpu_baseline": cpu,
        "copy_ceiling": copy,
        "single_cubelet": single,
    }
    if batched is not None:
        line["batched"] = batched
    if serve is not None:
        line["serving"] = serve
    if split is not None:
        line["split"] = split
    if opt is not None:
        line["opt"] = opt
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
