"""A plain HBM copy of known size: the calibration run for rocprofv3's FETCH_SIZE
and WRITE_SIZE counters.  Each copy reads and writes exactly 1 GiB; the ratio of
that size to what a counter reports for the copy kernel scales the same counter's
readings for the join kernels.

    rocprofv3 --pmc FETCH_SIZE -f csv -d <dir> -- python3 new_fgssjoin_amd/copy_calibration.py
"""
import torch


def main(n_bytes: int = 1 << 30, reps: int = 4):
    src = torch.ones(n_bytes // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(f"copied {n_bytes} bytes x {reps}", flush=True)


if __name__ == "__main__":
    main()
