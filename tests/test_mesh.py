"""Mesh generation (source/generateMesh.m) restated in gatt_dns_amd/prep/mesh.py.

MATLAB is not available, so the meshes are pinned by what each construction fixes:
end points, monotonicity, symmetry, where the refinement goes, fixed points landing
exactly on nodes, buffer growth, the periodic end node, kept nodes under extra
refinement.  Bit-level agreement with MATLAB's own rounding is parity unpinned.
"""
import numpy as np
import pytest

from gatt_dns_amd.prep.mesh import BufferSpec, MeshSpec, calc_buffer_zone, generate_mesh


def test_uniform_and_power():
    np.testing.assert_array_equal(generate_mesh(0.0, 2.0, MeshSpec(n=11)), np.linspace(0.0, 2.0, 11))
    X = generate_mesh(1.0, 3.0, MeshSpec(n=21, type="power", power=2.0))
    assert X[0] == 1.0 and X[-1] == 3.0 and np.all(np.diff(X) > 0)
    assert np.diff(X)[0] < np.diff(X)[-1]          # x^2: refined at the start


@pytest.mark.parametrize("local", ["b", "i", "f"])
def test_tanh(local):
    X = generate_mesh(-0.5, 0.5, MeshSpec(n=41, type="tanh", local=local, par=2.0))
    d = np.diff(X)
    assert abs(X[0] + 0.5) < 1e-15 and abs(X[-1] - 0.5) < 1e-15 and np.all(d > 0)
    if local == "b":
        np.testing.assert_allclose(X + X[::-1], 0.0, atol=1e-15)
        assert d[0] < d[20] and d[-1] < d[20]
    elif local == "i":
        assert d[0] < d[-1]
    else:
        assert d[-1] < d[0]


def test_attractors_refine_near_the_points():
    m = MeshSpec(n=101, type="attractors", attractor_points=[0.3], attractor_strength=[5.0],
                 attractor_size=[0.05])
    X = generate_mesh(0.0, 1.0, m)
    d = np.diff(X)
    assert abs(X[0]) < 1e-12 and abs(X[-1] - 1.0) < 1e-12 and np.all(d > 0)
    mid = 0.5 * (X[1:] + X[:-1])
    assert abs(mid[np.argmin(d)] - 0.3) < 0.02
    assert d.max() / d.min() > 3


@pytest.mark.parametrize("kind", [1, 2])
def test_match_fixed_points(kind):
    m = MeshSpec(n=51, type="tanh", local="b", par=1.5, match_fixed=kind, fix_points=[0.237, 0.61])
    X = generate_mesh(0.0, 1.0, m)
    for p in (0.237, 0.61):
        assert np.min(np.abs(X - p)) == 0.0
    assert np.all(np.diff(X) > 0) and X[0] == 0.0 and X[-1] == 1.0


@pytest.mark.parametrize("btype", ["sigmoid", "exponential"])
def test_buffers(btype):
    b = BufferSpec(n=10, type=btype, stretching=0.1 if btype == "exponential" else 4.0)
    m = MeshSpec(n=31, buffer_i=b, buffer_f=b)
    X = generate_mesh(0.0, 3.0, m)
    assert len(X) == 51 and np.all(np.diff(X) > 0)
    np.testing.assert_allclose(X[10:41], np.linspace(0.0, 3.0, 31), rtol=0, atol=1e-15)
    dx = 0.1
    np.testing.assert_allclose(X[41:] - 3.0, calc_buffer_zone(b) * dx, rtol=1e-12)
    np.testing.assert_allclose(0.0 - X[:10][::-1], calc_buffer_zone(b) * dx, rtol=1e-12)
    assert np.all(np.diff(np.diff(X[40:])) > 0)            # spacing grows outward


def test_buffer_length_target():
    m = MeshSpec(n=21, buffer_f=BufferSpec(n=None, l=2.0, type="sigmoid", stretching=3.0))
    X = generate_mesh(0.0, 1.0, m)
    assert X[-1] - 1.0 >= 2.0                               # the smallest count that reaches l
    b = BufferSpec(n=len(X) - 21 - 1, type="sigmoid", stretching=3.0)
    assert calc_buffer_zone(b)[-1] * 0.05 < 2.0


def test_periodic_fixed_domain_size_and_refinement():
    X = generate_mesh(0.0, 1.0, MeshSpec(n=8, periodic=True, fix_periodic_domain_size=True))
    np.testing.assert_allclose(X, np.arange(8) / 8, atol=1e-15)
    base = generate_mesh(0.0, 1.0, MeshSpec(n=11, type="tanh", local="b", par=1.0))
    Xr = generate_mesh(0.0, 1.0, MeshSpec(n=11, type="tanh", local="b", par=1.0, extra_refinement=2))
    assert len(Xr) == 31 and np.all(np.diff(Xr) > 0)
    np.testing.assert_array_equal(Xr[::3], base)


def test_node_count_from_spacing():
    X = generate_mesh(0.0, 4.0, MeshSpec(d0=0.1))
    assert np.max(np.diff(X)) <= 0.1 + 1e-15


def test_file_mesh(tmp_path):
    """'file' meshes (generateMesh.m:162-180): read as they are when the file holds every node,
    the buffers added to the file's physical nodes with fileCalcBuffer, the node count checked
    either way; MAT files give the variable named after the direction, text a column or row."""
    from scipy.io import savemat
    nodes = np.cumsum(np.linspace(0.1, 0.3, 17)) - 0.1
    txt = tmp_path / "y.txt"
    np.savetxt(txt, nodes[:, None])
    mat = tmp_path / "mesh.mat"
    savemat(mat, {"x": nodes[None, :], "y": 2 * nodes[None, :]})
    np.testing.assert_array_equal(generate_mesh(0.0, 1.0, MeshSpec(n=17, type="file", file=str(txt)), "y"), nodes)
    np.testing.assert_array_equal(generate_mesh(0.0, 1.0, MeshSpec(n=17, type="file", file=str(mat)), "y"), 2 * nodes)
    b = BufferSpec(n=6, type="sigmoid", stretching=4.0)
    with pytest.raises(ValueError, match="contains 17 nodes instead of 29"):
        generate_mesh(0.0, 1.0, MeshSpec(n=17, type="file", file=str(mat), buffer_i=b, buffer_f=b), "x")
    X = generate_mesh(0.0, 1.0, MeshSpec(n=17, type="file", file=str(mat), buffer_i=b, buffer_f=b,
                                         file_calc_buffer=True), "x")
    np.testing.assert_array_equal(X[6:23], nodes)
    # the buffers are those a built-in mesh with the same edge spacings gets
    ref = generate_mesh(0.0, nodes[-1], MeshSpec(n=17, buffer_i=b, buffer_f=b))
    dref, dfile = ref[6] - ref[5], X[6] - X[5]
    np.testing.assert_allclose((X[6] - X[:6]) / dfile, (ref[6] - ref[:6]) / dref, rtol=1e-12)
    with pytest.raises(ValueError, match="contains 17 nodes instead of 18"):
        generate_mesh(0.0, 1.0, MeshSpec(n=18, type="file", file=str(txt), file_calc_buffer=True), "y")
    # a periodic file mesh takes no temporary end node
    Xp = generate_mesh(0.0, 1.0, MeshSpec(n=17, type="file", file=str(txt), periodic=True,
                                          fix_periodic_domain_size=True), "y")
    assert len(Xp) == 17
