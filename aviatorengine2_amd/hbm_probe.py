"""Plain device-to-device copy of a known size, to calibrate HBM traffic
counters (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE) against algorithmic bytes.

    python -m aviatorengine2_amd.hbm_probe [MiB]

Copies MiB mebibytes (default 1024) five times with one copy kernel each.
"""
import sys
import time


def main():
    import torch
    mib = int(sys.argv[1]) if len(sys.argv) > 1 else 1024
    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / 5
    print(f"copy {mib} MiB: read {mib} MiB + write {mib} MiB per launch, {2 * mib / 1024 / dt:.1f} GiB/s", flush=True)


if __name__ == "__main__":
    main()
