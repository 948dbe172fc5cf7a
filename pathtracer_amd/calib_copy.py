"""Plain device-to-device copy of a known size, for calibrating rocprofv3's
FETCH_SIZE / WRITE_SIZE counters against bytes actually moved.

    python -m pathtracer_amd.calib_copy [MiB]
"""
import sys


def main():
    import torch
    mib = int(sys.argv[1]) if len(sys.argv) > 1 else 256
    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(f"copied {mib} MiB x3 (read {mib} MiB + write {mib} MiB per copy)", flush=True)


if __name__ == "__main__":
    main()
