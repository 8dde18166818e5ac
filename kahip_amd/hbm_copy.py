"""A plain device-to-device copy of known size on the GPU, used to calibrate the rocprofv3
FETCH_SIZE / WRITE_SIZE counters (the copy reads and writes exactly `bytes` bytes).

    python -m kahip_amd.hbm_copy [--mib 1024] [--repeats 3]
"""
import argparse
import sys


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--mib", type=int, default=1024)
    ap.add_argument("--repeats", type=int, default=3)
    args = ap.parse_args(argv)
    import torch
    n = args.mib * (1 << 20) // 4
    src = torch.arange(n, dtype=torch.int32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(args.repeats):
        dst.copy_(src)
    torch.cuda.synchronize()
    ok = bool(torch.equal(dst, src))
    print(f"copied {args.mib} MiB x {args.repeats}, equal={ok}", flush=True)
    return 0 if ok else 1


if __name__ == "__main__":
    sys.exit(main())
