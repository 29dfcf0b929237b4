"""FP32 force error against FP64 at the 1M-node workload, on the CPU.

The oracle's f32 force pass equals the device's bit for bit at this size
(test_fa2_gpu.py checks both the forces and the whole trajectories), so its
error against the f64 walk of the same tree is the device's error.  The f64
walk re-sums every centre of mass in f64 and follows the f32 acceptance
decisions (fa2_barnes_hut.wgsl:108-156); the cells it would have decided
differently are counted as flips."""
import numpy as np

from jumpcannon import workloads
from oracle import oracle as O


def test_f32_forces_against_f64_ba1m():
    off, nb, xyz = workloads.ba_1m()
    n = len(off) - 1
    deg, aoff, aadj = O.fa2_prepare(n, off, nb)
    t = O.octree_morton(xyz, deg.astype(np.float32) + np.float32(1.0))
    s = O.Fa2Settings.make(theta=0.8, scaling_ratio=2.0)
    f32, _ = O.fa2_forces(xyz, None, deg, aoff, aadj, t, s, threads=8)
    f64, flips = O.fa2_forces_f64(xyz, deg, aoff, aadj, t, s, threads=8)
    a = f32.reshape(-1, 3).astype(np.float64)
    b = f64.reshape(-1, 3)
    nb_ = np.linalg.norm(b, axis=1)
    floor = 1e-3 * np.sqrt(np.mean(nb_ ** 2))
    rel = np.linalg.norm(a - b, axis=1) / np.maximum(nb_, floor)
    print(f"max {rel.max():.3e} p99 {np.quantile(rel, 0.99):.3e} flips {flips} below floor {(nb_ < floor).sum()}")
    assert rel.max() <= 1e-4
    assert np.quantile(rel, 0.99) <= 2e-6
    assert flips <= 100 and (nb_ < floor).sum() == 0
