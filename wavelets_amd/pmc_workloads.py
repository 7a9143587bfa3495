"""Workloads for rocprofv3 counter runs (--pmc): inputs are generated on the
host and copied to the device, so no other device kernels run beside the
library's own.

    python3 wavelets_amd/pmc_workloads.py plane32     (or by its full path)

plane32  256 cubelets of 32^3 float32, CDF 9/7 3,3,3, log 256, keep 10%, as
         one batch submitted three times (the plane pass, then the z pass,
         per submit), then the library's streaming copy of 1 GiB three times
         (the scale for FETCH_SIZE / WRITE_SIZE).
"""
import os
import sys


def plane32():
    import torch
    import wavelets_amd as wv
    from wavelets_amd import synth
    cubes = []
    for i in range(256):
        v = synth.cubelet((32, 32, 32), "float32", 500 + i)
        cubes.append(dict(vol=torch.from_numpy(v).cuda(), wavelet=wv.CDF97, steps=[3, 3, 3],
                          alg=wv.LOG, bins=256, frac=0.9))
    b = wv.Batch(cubes)
    for _ in range(3):
        b.submit()
        b.wait()
    b.close()
    src = torch.empty(1 << 28, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        wv.copy_stream(src, dst)
    torch.cuda.synchronize()
    print("plane32 done", flush=True)


def main(argv):
    if len(argv) != 2 or argv[1] not in ("plane32",):
        print(__doc__)
        return 2
    plane32()
    return 0


if __name__ == "__main__":
    # run by path (rocprofv3 runs it from another directory): the package's
    # parent directory is the import root
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    sys.exit(main(sys.argv))
