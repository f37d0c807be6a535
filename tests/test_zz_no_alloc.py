"""The allocation-free map seeding entry point (sdfslam_hip_init_from_scans_noalloc)
against the allocating one (sdfslam_hip_init_from_scans), through the C ABI.

Device memory for inputs and outputs comes from torch (the caller's); the
entry point itself allocates nothing. Two calls in a row on one stream, the
second with fewer points and a different grid, each bitwise equal to the
allocating entry point (the distances are exact nearest-point searches).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _case(sdfslam, seed, n, nx, ny, lo, hi):
    rng = np.random.default_rng(seed)
    grid = sdfslam.Grid(nx, ny, lo[0], lo[1], (hi[0] - lo[0]) / (nx - 1), (hi[1] - lo[1]) / (ny - 1))
    pts = np.stack([rng.uniform(lo[0], hi[0], n), rng.uniform(lo[1], hi[1], n)], axis=1)
    pts[n // 2: n // 2 + 50] = pts[:50]  # exact duplicates: ties go to the lower index
    origins = pts + rng.normal(0, 2.0, size=pts.shape)
    init = rng.normal(0, 1, size=nx * ny)
    return grid, pts, origins, init


def test_noalloc_matches_allocating_entry_point(sdfslam):
    import torch
    print("case no-alloc init_from_scans: 300000 points 150x120, then 3000 points 60x80", flush=True)
    dev = torch.device("cuda:0")
    cases = [_case(sdfslam, 1, 300_000, 150, 120, (-1.0, -2.0), (16.0, 12.0)),
             _case(sdfslam, 2, 3_000, 60, 80, (0.5, 0.5), (7.0, 9.0))]
    assert all(len(c[1]) <= sdfslam.NO_ALLOC_MAX_POINTS for c in cases)
    stream = torch.cuda.current_stream(dev)
    keep, outs = [], []
    for signed in (True, False):
        for grid, pts, origins, init in cases:
            tp = torch.from_numpy(pts.reshape(-1).copy()).to(dev)
            to = torch.from_numpy(origins.reshape(-1).copy()).to(dev)
            tm = torch.from_numpy(init.copy()).to(dev)
            keep += [tp, to]
            sdfslam.init_from_scans_noalloc(grid, tp.data_ptr(), to.data_ptr(), len(pts), tm.data_ptr(),
                                            signed, stream.cuda_stream)
            outs.append((grid, pts, origins, init, signed, tm))
    torch.cuda.synchronize()
    for grid, pts, origins, init, signed, tm in outs:
        want = sdfslam.init_from_scans(grid, init, (pts, origins), None, signed_distance=signed, device=False)
        got = tm.cpu().numpy()
        assert np.array_equal(got, want), (len(pts), signed)


def test_noalloc_empty_and_oversize(sdfslam):
    import torch
    print("case no-alloc empty and oversize", flush=True)
    dev = torch.device("cuda:0")
    grid, pts, origins, init = _case(sdfslam, 3, 100, 20, 20, (0.0, 0.0), (5.0, 5.0))
    tp = torch.from_numpy(pts.reshape(-1).copy()).to(dev)
    tm = torch.from_numpy(init.copy()).to(dev)
    # No points: the map stays as it is.
    sdfslam.init_from_scans_noalloc(grid, tp.data_ptr(), tp.data_ptr(), 0, tm.data_ptr(), True,
                                    torch.cuda.current_stream(dev).cuda_stream)
    torch.cuda.synchronize()
    assert np.array_equal(tm.cpu().numpy(), init)
    # Beyond the maximum: refused before anything is enqueued.
    with pytest.raises(ValueError, match="allocation-free maximum"):
        sdfslam.init_from_scans_noalloc(grid, tp.data_ptr(), tp.data_ptr(), sdfslam.NO_ALLOC_MAX_POINTS + 1,
                                        tm.data_ptr(), True, torch.cuda.current_stream(dev).cuda_stream)
    torch.cuda.synchronize()
    assert np.array_equal(tm.cpu().numpy(), init)
