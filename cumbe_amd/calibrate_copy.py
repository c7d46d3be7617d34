"""Plain device copy of a known size, for scaling HBM counters.

Run under `rocprofv3 --pmc FETCH_SIZE` (then WRITE_SIZE): a copy of N bytes
must fetch N and write N bytes, so the counters' readings for this kernel give
the factor that converts them to bytes for the enumeration kernel.
"""
import sys

import torch


def main(mib: int = 256) -> None:
    n = mib * (1 << 20) // 4
    x = torch.arange(n, dtype=torch.int32, device="cuda")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    y.copy_(x)
    torch.cuda.synchronize()
    print(f"copied {mib} MiB: {int(y[-1])}")


if __name__ == "__main__":
    main(int(sys.argv[1]) if len(sys.argv) > 1 else 256)
