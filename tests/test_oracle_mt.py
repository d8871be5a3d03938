"""The all-cores CPU build (oracle/dmsh_cpu_mt.c, the benchmark's all-cores
baseline) gives the serial del-msh-cpu pipeline's tree and boxes bit for bit,
for any thread count, also with repeated codes and tiny meshes."""

import numpy as np
import pytest

from del_msh_dlpack import primitives


def _bits_equal(a, b):
    return a.shape == b.shape and np.array_equal(a.view(np.uint32), b.view(np.uint32))


@pytest.mark.parametrize("threads", [1, 3, 8])
def test_mt_build_equals_serial(oracle, threads):
    for tri, vtx in (primitives.octahedral_sphere(5, noise_sigma=0.01, seed=2),
                     primitives.torus_zup(1.3, 0.4, 64, 32),
                     primitives.cylinder_closed_end_yup(0.5, 2.0, 40, 30)):
        wbn, waabb = oracle.build_trimesh3(tri, vtx, morton_bits=63)
        bn, aabb = oracle.build_trimesh3_mt(tri, vtx, threads)
        assert np.array_equal(bn, wbn)
        assert _bits_equal(aabb, waabb)


def test_mt_build_duplicates_and_tiny(oracle):
    # every triangle twice: each centroid repeats, so runs of equal codes
    tri, vtx = primitives.torus_zup(1.0, 0.3, 16, 8)
    tri2 = np.concatenate([tri, tri]).astype(np.uint32)
    wbn, waabb = oracle.build_trimesh3(tri2, vtx, morton_bits=63)
    bn, aabb = oracle.build_trimesh3_mt(tri2, vtx, 4)
    assert np.array_equal(bn, wbn) and _bits_equal(aabb, waabb)
    for n in (1, 2, 3):
        t = tri[:n].copy()
        wbn, waabb = oracle.build_trimesh3(t, vtx, morton_bits=63)
        bn, aabb = oracle.build_trimesh3_mt(t, vtx, 2)
        assert np.array_equal(bn, wbn) and _bits_equal(aabb, waabb)
