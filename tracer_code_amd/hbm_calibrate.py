"""Device-to-device copy of a known size, for calibrating HBM byte counters.

    python -m tracer_code_amd.hbm_calibrate [MiB]

Run under `rocprofv3 --pmc FETCH_SIZE` (then `--pmc WRITE_SIZE`): the copy
kernel reads and writes exactly MiB mebibytes, so the counters' readings for
it give the factor that converts their units into bytes on this GPU.
"""
import sys


def main():
    import torch

    mib = int(sys.argv[1]) if len(sys.argv) > 1 else 256
    n = mib * (1 << 20) // 4
    src = torch.arange(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    ok = bool(torch.equal(dst[-16:], src[-16:]))
    print(f"copied {mib} MiB x3 (bytes per copy read={mib << 20} written={mib << 20}) ok={ok}", flush=True)
    return 0 if ok else 1


if __name__ == "__main__":
    sys.exit(main())
