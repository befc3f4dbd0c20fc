"""CPU checks that the synthetic workload generators restate the reference's
benchmark programs (their route/weight/keep functions), with a literal
per-point loop over the program text as the second implementation."""
import numpy as np

import workloads


def imod(x, m):
    return x - (x // m) * m


def test_voxel_rasterization_generator():
    # bench/voxel_rasterization/voxel_rasterization.hyd:22-41
    n, nx, ny, nz, kp = 1000, 64, 64, 64, 3
    idx, vals, guard, nvox = workloads.voxel_rasterization(n, nx, ny, nz, kp)
    assert nvox == nx * ny * nz
    for p in range(n):
        x, y, z = imod(p * 17 + 3, nx), imod(p * 29 + 5, ny), imod(p * 43 + 7, nz)
        assert idx[p] == ((z * ny) + y) * nx + x
        assert vals[p] == ((x + 1) * (y + 2)) + (z * 3) + 1
        assert guard[p] == (1 if (p // kp) * kp == p else 0)


def test_voxel_trilinear_splat_generator():
    # bench/voxel_trilinear_splat/voxel_trilinear_splat.hyd:19-57
    n, nx, ny, nz, kp = 200, 64, 64, 64, 3
    idx, vals, guard, nvox = workloads.voxel_trilinear_splat(n, nx, ny, nz, kp)
    assert idx.size == 8 * n
    for i in range(8 * n):
        p, c = i // 8, imod(i, 8)
        bx, by, bz = imod(c, 2), imod(c // 2, 2), imod(c // 4, 2)
        x, y, z = imod(p * 17 + 3 + bx, nx), imod(p * 29 + 5 + by, ny), imod(p * 43 + 7 + bz, nz)
        assert idx[i] == ((z * ny) + y) * nx + x
        fx, fy, fz = imod(p * 5 + 1, 4), imod(p * 7 + 2, 4), imod(p * 11 + 3, 4)
        wx = (1 - bx) * (4 - fx) + bx * fx
        wy = (1 - by) * (4 - fy) + by * fy
        wz = (1 - bz) * (4 - fz) + bz * fz
        assert vals[i] == float((p * 3 + 1) * wx * wy * wz) / 64.0
        assert guard[i] == (1 if (p // kp) * kp == p else 0)
