"""Reference copy kernel for HBM traffic counters.

Copies a 1 GiB device buffer into another (torch on the HIP device) a few times, so that
rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE of this plain copy can be compared with the
known 1 GiB per dispatch: the ratio calibrates the counters read for the step kernel.

    rocprofv3 --pmc FETCH_SIZE -f csv -d <dir> -- python3 peridynamics.jl_amd/copy_calibration.py
"""
import torch

N = (1 << 30) // 8
src = torch.ones(N, dtype=torch.float64, device="cuda")
dst = torch.empty_like(src)
for _ in range(3):
    dst.copy_(src)
torch.cuda.synchronize()
print(f"copied {N * 8} bytes x 3")
