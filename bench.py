"""Benchmark of the MI355X Wavelets compression hot path.

Default workload (one GPU): the 512^3 float32 cubelet (40 ellipsoids +
N(0,20^2) noise), CDF 9/7 at 3,3,3 steps, keep the largest 5% of |coef|
(-thresh 0.95), uniform quantizer with 1024 bins.  One step is the whole
device-resident pipeline on that cubelet with the input already in HBM:
fused-axis-pass DWT (threshold statistics fused into the last pass), radix
select of the threshold, quantization fused with the bin histogram.

Prints one JSON line: input GB/s, the DWT pass roofline (algorithmic bytes /
average pass duration from HIP events recorded on the pipeline's own stream)
and the CPU oracle timed on a bounded sample on this host.
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


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--size", type=int, default=512)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--extras", default="e2e",
                    help="comma list of extra measurements: e2e,batched,serving")
    return ap.parse_args()


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
            "sample": "%dx%dx%d float32 sub-volume of the same cubelet, CDF97 3,3,3, "
                      "thresh 0.95, uniform 1024; DWT+threshold(sort)+quantize "
                      "= %.2f s single-threaded (Huffman %.2f s not counted)" % (
                          vol_sample.shape[2], vol_sample.shape[1], vol_sample.shape[0],
                          t, r.t_huff)}


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(cubelets shard across GPUs with no exchange; see DESIGN.md)",
              file=sys.stderr)
        sys.exit(2)
    import torch
    import wavelets_amd as wv
    from wavelets_amd import synth
    L = wv.lib()
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)

    n = args.size
    t0 = time.time()
    vol = synth.ellipsoids_f32(n=n, seed=2)
    gen_s = time.time() - t0
    print("generated %d^3 float32 in %.1f s" % (n, gen_s), file=sys.stderr, flush=True)
    vol_d = torch.from_numpy(vol).to(dev)
    N = vol.size
    in_bytes = vol.nbytes

    g = wv.geom([n, n, n], [3, 3, 3], wv.CDF97, wv.FLOAT32)
    coef = torch.empty(g.count, dtype=torch.float32, device=dev)
    ws = torch.empty(max(g.ws_floats, wv.lib().wv_threshold_ws_bytes(g.count) // 4 + 1),
                     dtype=torch.float32, device=dev)
    stats = torch.empty(wv.STATS_BYTES, dtype=torch.uint8, device=dev)
    res_d = torch.zeros(64, dtype=torch.uint8, device=dev)
    bins = torch.empty(g.count, dtype=torch.int32, device=dev)
    counts = torch.empty(1024 + 128, dtype=torch.int32, device=dev)
    stream = torch.cuda.current_stream()
    sp = C.c_void_p(stream.cuda_stream)
    npass = L.wv_dwt3d_num_passes(C.byref(g))

    def make_events():
        return [wv.Event() for _ in range(npass + 1)]

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

    for _ in range(args.warmup):
        step(make_events())
    torch.cuda.synchronize()
    evsets = [make_events() for _ in range(args.steps)]
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t_wall = time.time()
    start.record(stream)
    for k in range(args.steps):
        step(evsets[k])
    stop.record(stream)
    torch.cuda.synchronize()
    t_wall = time.time() - t_wall
    ms_total = start.elapsed_time(stop)
    ms_step = ms_total / args.steps
    pass_ms = np.zeros(npass)
    for evs in evsets:
        for i in range(npass):
            pass_ms[i] += evs[i].elapsed_ms(evs[i + 1])
    pass_ms /= args.steps
    # algorithmic bytes per DWT pass launch: read 4 B + write 4 B per element
    # (float32 input on the first pass too)
    pass_bytes = 8.0 * N
    avg_pass_ms = float(pass_ms.mean())
    achieved = pass_bytes / (avg_pass_ms * 1e-3) / 1e9
    value = in_bytes / (ms_step * 1e-3) / 1e9

    _, r = wv.threshold(coef, 0.95, stats)   # final result for the record
    extra = {}
    if "e2e" in args.extras:
        # plain drop-in end to end: wv_compress (device stages, D2H of bins
        # and counts, host Huffman + .cube image)
        opt = wv.Options(wv.CDF97, wv.UNIFORM, 1024, (C.c_int * 3)(3, 3, 3),
                         C.c_float(0.95), 1, 30, 1)
        orig = (C.c_int * 3)(n, n, n)
        wb = L.wv_compress_work_bytes(orig, C.byref(opt))
        work = torch.empty(wb, dtype=torch.uint8, device=dev)
        cap = g.count * 4 + (1 << 22)
        fbuf = np.empty(cap, np.uint8)
        info = wv.CompressInfo()
        wv.check(L.wv_compress(C.c_void_p(vol_d.data_ptr()), wv.FLOAT32, orig, C.byref(opt),
                               C.c_void_p(work.data_ptr()), wb, None,
                               fbuf.ctypes.data_as(C.c_void_p), cap, C.byref(info), sp))
        extra["e2e_plain"] = {
            "gbps_input": round(in_bytes / (info.ms_total * 1e-3) / 1e9, 3),
            "ms_total": round(info.ms_total, 3),
            "ms_device_incl_host_roundtrip": round(info.ms_device, 3),
            "file_bytes": int(info.file_bytes),
            "note": "includes bins D2H (4 B/elt) and single-threaded host Huffman"}
        del work

    cpu = None
    if not args.no_cpu:
        q = max(32, n // 2)
        sample = np.ascontiguousarray(vol[:q, :q, :])   # [z, y, x] = 512 x q x q
        cpu = cpu_baseline(sample)

    line = {
        "metric": "input GB/s compressed (device-resident pipeline, plain single cubelet)",
        "select_fallback": bool(r.n_candidates & 0x80000000 if r.n_candidates >= 0 else True),
        "value": round(value, 3),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded 40-ellipsoid volume + Gaussian noise)",
        "config": {"workload": "512^3 float32 cubelet, CDF97 3,3,3, keep 5% "
                               "(thresh 0.95), uniform 1024 bins" if n == 512 else
                   "%d^3 float32 cubelet" % n,
                   "stages": "dwt(3 fused axis passes)+sampled-window select+quantize/histogram",
                   "threshold": float(r.threshold), "max_abs": float(r.max_abs)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": None,
                     "kernel": "dwt_pass_kernel (fused axis pass)",
                     "pass_ms": [round(float(v), 4) for v in pass_ms],
                     "algorithmic_bytes_per_launch": pass_bytes},
        "cpu_baseline": cpu,
        "wall_s": round(t_wall, 3),
    }
    line.update(extra)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
