"""Node normals of the problem build on the device equal the host builder's
(kNN PCA over all scan points, nearest-point normal per node) bit for bit."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _world_points(sdfslam, name, seed, frames):
    d = sdfslam.Dataset(name, seed)
    pts, org = [], []
    for k in range(frames):
        xy, pose, _ = d.frame(k)
        c, s = np.cos(pose[2]), np.sin(pose[2])
        pts.append(np.stack([c * xy[:, 0] - s * xy[:, 1] + pose[0], s * xy[:, 0] + c * xy[:, 1] + pose[1]], 1))
        org.append(np.tile(pose[:2], (len(xy), 1)))
    return d, np.concatenate(pts), np.concatenate(org)


@pytest.mark.parametrize("name,seed,frames,k", [("served0", 2000, 40, 7), ("served3", 2000, 12, 7),
                                                ("office150", 1, 150, 10)])
@pytest.mark.parametrize("method", ["pca", "hybrid", "weighted"])
def test_device_node_normals_match_host(sdfslam, name, seed, frames, k, method):
    print(f"case node normals {name} frames={frames} k={k} {method}", flush=True)
    d, pts, org = _world_points(sdfslam, name, seed, frames)
    dev = sdfslam.node_normals(pts, org, k, d.grid, method, 0.2, device=True)
    ref = sdfslam.node_normals(pts, org, k, d.grid, method, 0.2, device=False)
    assert np.array_equal(dev[0], ref[0]) and np.array_equal(dev[1], ref[1])


def test_device_node_normals_edge_cases(sdfslam):
    print("case node normals edge cases", flush=True)
    g = sdfslam.Grid(9, 7, -4.0, -3.0, 1.0, 1.0)
    rng = np.random.default_rng(5)
    cases = [np.zeros((0, 2)),                      # no points
             np.array([[0.3, 0.2]]),                # fewer points than k
             np.array([[0.5, 0.5]] * 12),           # duplicates: ties broken by index
             np.round(rng.uniform(-3, 3, (200, 2)), 1),  # grid-aligned coordinates, many equal distances
             rng.uniform(30.0, 31.0, (50, 2))]      # every node far from every point
    for pts in cases:
        org = np.zeros_like(pts) + [0.0, -10.0]
        for method in ("pca", "hybrid", "weighted"):
            dev = sdfslam.node_normals(pts, org, 7, g, method, 0.2, device=True)
            ref = sdfslam.node_normals(pts, org, 7, g, method, 0.2, device=False)
            assert np.array_equal(dev[0], ref[0]) and np.array_equal(dev[1], ref[1]), (len(pts), method)
