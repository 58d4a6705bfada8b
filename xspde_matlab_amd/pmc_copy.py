"""Plain device copy of known size, the reference kernel for scaling the
FETCH_SIZE / WRITE_SIZE counters of rocprofv3 --pmc to bytes.

    rocprofv3 --pmc FETCH_SIZE -d <dir> -o copy -- python3 -m xspde_matlab_amd.pmc_copy

Copies 256 MiB (64 Mi float32) three times with torch (one copy kernel per
call): each call reads 262,144 KiB and writes 262,144 KiB.
"""
import torch


def main():
    n = 64 * 1024 * 1024
    x = torch.ones(n, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print("copied %d bytes x3" % (4 * n), flush=True)


if __name__ == "__main__":
    main()
