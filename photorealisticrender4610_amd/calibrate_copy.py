"""Plain device-to-device copy of a known size, for calibrating rocprofv3's
FETCH_SIZE / WRITE_SIZE counters: a copy of B bytes must fetch and write B
bytes of HBM. Usage (under rocprofv3 --pmc ...): python calibrate_copy.py [MiB]"""
import sys

import torch


def main():
    mib = int(sys.argv[1]) if len(sys.argv) > 1 else 1024
    n = mib * (1 << 20) // 4
    x = torch.rand(n, device="cuda")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print("copied %d MiB x3" % mib)


if __name__ == "__main__":
    main()
