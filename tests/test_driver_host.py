"""Host driver (include/ptycho.h): parameters, scan geometry, CSV format, Python argument
handling.  No GPU calls."""
import numpy as np
import pytest

from ptychopy_amd import _driver as D


def f32(x):
    return np.float32(x)


# ---------------------------------------------------------------- a1: sample-plane pixel size
def test_dx_s_is_rounded_to_float():
    v = D.sample_pixel_size(0.248e-9, 1.0, 512, 172e-6)
    assert v == float(np.float32(0.248e-9 * 1.0 / (512 * 172e-6)))
    assert abs(v - 2.81613e-9) < 1e-14  # 0.248 nm * 1 m / (512 * 172 um)
    assert v != 0.248e-9 / (512 * 172e-6)  # the double value is not kept


# ---------------------------------------------------------------- a2/a3: Cartesian mesh
def _cartesian_closed_form(nx, ny, step, dx_s, m):
    """Float32 arithmetic of d_generateMesh<1> with no jitter or drift, extents with the
    pitched buffer's zero padding (rows padded to 32), canvas ceil(max - min + M)."""
    o = f32(1.0 / dx_s)
    s = f32(step)
    px = np.empty((nx, ny), np.float32)
    py = np.empty((nx, ny), np.float32)
    for i in range(nx):
        for j in range(ny):
            vx = (f32(i) - f32(nx - 1) * f32(0.5)) * s
            vy = (f32(j) - f32(ny - 1) * f32(0.5)) * s
            px[i, j] = vx * o
            py[i, j] = vy * o
    mx = [px.max(), py.max()]
    mn = [px.min(), py.min()]
    if ny % 32:
        mx = [max(mx[0], f32(0)), max(mx[1], f32(0))]
        mn = [min(mn[0], f32(0)), min(mn[1], f32(0))]
    canvas = (int(np.ceil(f32(mx[0] - mn[0]) + f32(m))), int(np.ceil(f32(mx[1] - mn[1]) + f32(m))))
    pos = np.stack([(px - mn[0]).ravel(), (py - mn[1]).ravel()], axis=1)
    return pos, canvas


@pytest.mark.parametrize("nx,ny,step,size", [(30, 30, 50e-9, 512), (12, 12, 40e-9, 256), (51, 51, 100e-9, 256),
                                             (5, 7, 33e-9, 128), (32, 32, 40e-9, 64)])
def test_cartesian_mesh_bit_exact(nx, ny, step, size):
    dx_s = D.sample_pixel_size(0.248e-9, 1.0, size, 172e-6)
    pos, canvas, ext = D.cartesian_mesh(nx, ny, step, step, dx_s, size)
    ref, rcanvas = _cartesian_closed_form(nx, ny, step, dx_s, size)
    assert canvas == rcanvas
    assert np.array_equal(pos.view(np.uint32), ref.astype(np.float32).view(np.uint32))
    # raster order: row index i is the first axis (x), j the contiguous one (y)
    assert pos[1, 1] > pos[0, 1] and pos[ny, 0] > pos[0, 0]


def test_cartesian_c1_geometry():
    dx_s = D.sample_pixel_size(0.248e-9, 1.0, 512, 172e-6)
    pos, canvas, _ = D.cartesian_mesh(30, 30, 50e-9, 50e-9, dx_s, 512)
    step_px = np.diff(pos[:30, 1]).mean()
    assert abs(step_px - 50e-9 / 2.816134e-9) < 1e-3  # 17.7548 px
    assert canvas == (1027, 1027)
    assert pos.min() == 0.0


def test_cartesian_mirror_flip_and_single_position():
    dx_s = 1e-8
    a, _, _ = D.cartesian_mesh(4, 6, 1e-7, 1e-7, dx_s, 64)
    b, _, _ = D.cartesian_mesh(4, 6, 1e-7, 1e-7, dx_s, 64, mirror_x=True)
    assert np.allclose(a[:, 0].reshape(4, 6)[::-1], b[:, 0].reshape(4, 6))
    c, _, _ = D.cartesian_mesh(4, 6, 1e-7, 1e-7, dx_s, 64, flip=True)
    assert np.array_equal(c[:, 0], a[:, 1]) and np.array_equal(c[:, 1], a[:, 0])
    p, canvas, _ = D.cartesian_mesh(1, 1, 1e-7, 1e-7, dx_s, 64)
    assert canvas == (64, 64) and p.tolist() == [[0.0, 0.0]]


def test_cartesian_jitter_bounds_and_seed():
    dx_s = 1e-8
    a, _, _ = D.cartesian_mesh(12, 12, 16e-8, 16e-8, dx_s, 128)
    j1, _, e1 = D.cartesian_mesh(12, 12, 16e-8, 16e-8, dx_s, 128, jitter_radius=1, seed=5)
    j2, _, e2 = D.cartesian_mesh(12, 12, 16e-8, 16e-8, dx_s, 128, jitter_radius=1, seed=5)
    j3, _, _ = D.cartesian_mesh(12, 12, 16e-8, 16e-8, dx_s, 128, jitter_radius=1, seed=6)
    assert np.array_equal(j1, j2) and not np.array_equal(j1, j3)
    raw_a = a + 0  # unshifted difference: jitter in [-1, 1)
    d = (j1 + e1[:2]) - (raw_a + np.array([-88.0, -88.0], np.float32))
    assert np.all(np.abs(d) <= 1.0 + 1e-4)


# ---------------------------------------------------------------- list / spiral meshes
def test_list_mesh_extents_include_zero_padding():
    # all-positive positions: the reference's min over the padded buffer is 0, not min(pos)
    dx_s = 1e-8
    metres = np.array([[1e-6, 2e-6], [1.5e-6, 3e-6], [2e-6, 2.5e-6]])
    pos, canvas, ext = D.list_mesh(metres, dx_s, 64)
    o = f32(1 / dx_s)
    raw = np.array([[f32(x) * o, f32(y) * o] for x, y in metres], np.float32)
    assert ext[0] == 0.0 and ext[1] == 0.0
    assert np.array_equal(pos, raw)                       # shifted by the padded minimum, 0
    assert canvas == (int(np.ceil(raw[:, 0].max() + 64)), int(np.ceil(raw[:, 1].max() + 64)))


def test_spiral_mesh_closed_form():
    dx_s = 1e-8
    pos, canvas, ext = D.spiral_mesh(5, 5, 1e-7, 1e-7, dx_s, 64)
    o = f32(1 / dx_s)
    q = np.arange(1024, dtype=np.float32)
    s = np.sqrt(q)
    x = (s * np.cos(f32(4) * s).astype(np.float32) * f32(1e-7)).astype(np.float32) * o
    y = (s * np.sin(f32(4) * s).astype(np.float32) * f32(1e-7)).astype(np.float32) * o
    # extents cover all 1024 entries of the padded buffer (only 25 are scan points)
    assert abs(ext[2] - x.max()) <= 1e-3 * abs(x.max()) and abs(ext[0] - x.min()) <= 1e-3 * abs(x.min())
    assert np.allclose(pos[:, 0], x[:25] - ext[0], rtol=1e-5, atol=1e-3)
    assert np.allclose(pos[:, 1], y[:25] - ext[1], rtol=1e-5, atol=1e-3)


# ---------------------------------------------------------------- CSV
def test_complex_csv_format_and_round_trip(tmp_path):
    a = np.array([[1.5 + 2j, -0.25 - 1e-7j], [3.14159265 + 0j, 1e-30 - 2.5e10j]], np.complex64)
    f = tmp_path / "x.csv"
    assert D.save_complex(f, a)
    lines = f.read_text().splitlines()
    assert lines[0] == "1.5+2j, -0.25-1e-07j"
    assert lines[1] == "3.14159+0j, 1e-30-2.5e+10j"
    b = D.load_complex_csv(f)
    assert b.shape == (2, 2)
    assert np.allclose(b, a, rtol=1e-5)
    # exponent only on the imaginary part (misread by the reference's parser)
    h = tmp_path / "z.csv"
    h.write_text("0.5+2.3e-06j, (1-2j), 3, -4.5e-03-1e+02j\n")
    assert np.allclose(D.load_complex_csv(h), [[0.5 + 2.3e-6j, 1 - 2j, 3, -4.5e-3 - 100j]], rtol=1e-6)
    g = tmp_path / "y.bin"
    assert D.save_complex(g, a, binary=True)
    assert np.array_equal(np.fromfile(g, np.complex64).reshape(2, 2), a)


# ---------------------------------------------------------------- parameters / Python surface
def test_parameter_parsers():
    from ptychopy_amd.ptychopy import Phaser
    ph = Phaser()
    ph.parse_cmdline("-jobID=abc -algorithm=DM -scanDims=30,20 -step=50e-9 -size=512 -lambda=0.248e-9 "
                     "-z=1 -dx_d=172e-6 -i=7 -updateProbe=3")
    assert ph.get_param("jobID") == "abc" and ph.get_param("algorithm") == "DM"
    assert ph.get_param("scanDimsx") == "30" and ph.get_param("scanDimsy") == "20"
    assert ph.get_param("iterations") == "7" and ph.get_param("updateProbe") == "3"
    assert float(ph.get_param("dx_s")) == D.sample_pixel_size(0.248e-9, 1.0, 512, 172e-6)
    with pytest.raises(RuntimeError):
        ph.parse_cmdline("-notAnOption=1")
    ph.set_param("T", 0)          # T > 0 only (Parameters.cpp:486-489)
    assert float(ph.get_param("time")) == -1.0
    ph.set_param("iter", 5)
    assert ph.get_param("iterations") == "5"
    with pytest.raises(RuntimeError):
        ph.set_param("size", 16)  # "the size must be large than 16"
    ph.close()


def test_python_surface_argument_handling(capsys):
    from ptychopy_amd import ptychopy
    with pytest.raises(TypeError):
        ptychopy.epie(notakeyword=1)
    with pytest.raises(TypeError):
        ptychopy.dm("job", iter="many")
    # no diffractionNP: the message, and True as every whole-mode entry point returns
    # (ptychopy.c:366-367, 523-540)
    assert ptychopy.epienp(jobID="x") is True
    assert "Please input a diffraction pattern numpy array" in capsys.readouterr().out
    with pytest.raises(ValueError, match="2 dimensional object"):
        ptychopy.epienp(diffractionNP=np.zeros((1, 32, 32)), objectNP=np.zeros(3, complex))
    names = set(ptychopy.__all__)
    assert {"epie", "dm", "mls", "epienp", "dmnp", "mlsnp", "epiecmdstr", "dmcmdstr", "mlscmdstr", "epieinit",
            "epiepost", "epiestep", "epieresobj", "epieresprobe", "epienpinit", "mlsinit", "mlspost", "mlsstep",
            "mlsresobj", "mlsresprobe", "helloworld"} <= names


def test_driver_exports():
    import ctypes
    import re
    from pathlib import Path
    from ptychopy_amd import _engine
    text = (Path(__file__).resolve().parent.parent / "include" / "ptycho.h").read_text()
    declared = set(re.findall(r"\b(ptycho_[a-z0-9_]+)\s*\(", text))
    assert declared == set(D.DRIVER_EXPORTS)
    lib = ctypes.CDLL(str(_engine.LIB_PATH))
    assert not [s for s in declared if not hasattr(lib, s)]


# ---------------------------------------------------------------- MLs geometry
def _padded_extent(raw, ny):
    mn, mx = raw.min(axis=0), raw.max(axis=0)
    if ny % 32:
        mn, mx = np.minimum(mn, 0), np.maximum(mx, 0)
    return np.array([mn[0], mn[1], mx[0], mx[1]], np.float32)


@pytest.mark.parametrize("nx,ny", [(51, 51), (3, 40), (40, 3)])
def test_mls_canvas_holds_every_roi(nx, ny):
    import oracle
    m = 256
    dx_s = D.sample_pixel_size(0.248e-9, 1.0, m, 172e-6)
    raw = D.cartesian_mesh_raw(nx, ny, 100e-9, 100e-9, dx_s, m)
    # unshifted and centred (CartesianScanMesh::generateMeshML keeps the kernel output)
    assert np.allclose(raw.mean(axis=0), 0, atol=1e-3)
    (rows, cols), box = D.mls_canvas(raw, _padded_extent(raw, ny), m)
    roi, sub = oracle.mls_roi(raw, m, rows, cols)
    assert roi.min() >= 0 and roi[:, 0].max() + m <= rows and roi[:, 1].max() + m <= cols
    # 20 % margin around the ROI span, per axis
    span = roi.max(axis=0) - roi.min(axis=0) + m - 1
    assert abs(rows - 1.2 * span[0]) <= 2 and abs(cols - 1.2 * span[1]) <= 2
    # calc_object_norm box: 1-based, inside the canvas, centred on the scan
    assert 1 <= box[0] <= box[1] <= rows and 1 <= box[2] <= box[3] <= cols
    assert abs((box[0] + box[1]) / 2 - rows / 2) <= 1.5 and abs((box[2] + box[3]) / 2 - cols / 2) <= 1.5


def test_mls_partitions():
    P = D.mls_partitions(2601, 100, seed=3)
    assert len(P) == 30  # min(Niter, 30)
    for part in P:
        assert [len(g) for g in part] == [26] * 100 + [1]
        assert np.array_equal(np.sort(np.concatenate(part)), np.arange(2601))
    assert not np.array_equal(P[0][0], P[1][0])
    Q = D.mls_partitions(2601, 100, seed=3)
    assert all(np.array_equal(a, b) for a, b in zip(P[5], Q[5]))
    small = D.mls_partitions(50, 5, seed=1)  # fewer than 100 frames: groups of one
    assert len(small) == 5 and all(len(g) == 1 for g in small[0]) and len(small[0]) == 50


# ---------------------------------------------------------------- compact (MLC) groups
def _raster(side, step=14.0, jitter=0.0, seed=0):
    rng = np.random.default_rng(seed)
    r, c = np.meshgrid(np.arange(side), np.arange(side), indexing="ij")
    pos = np.stack([r.ravel(), c.ravel()], axis=1).astype(np.float64) * step
    return (pos + rng.uniform(-jitter, jitter, pos.shape)).astype(np.float32)


@pytest.mark.parametrize("side", [20, 51])
def test_compact_groups_partition_and_sizes(side):
    """Every frame once; ceil(N / max(1, N/100)) groups by decreasing size, at most two sizes."""
    from ptychopy_amd import _driver as D
    pos = _raster(side, jitter=1.0)
    n = len(pos)
    groups = D.mls_compact_groups(pos, seed=3)
    grouping = max(1, n // 100)
    assert len(groups) == -(-n // grouping)
    allf = np.concatenate(groups)
    assert np.array_equal(np.sort(allf), np.arange(n))
    sizes = [len(g) for g in groups]
    assert sizes == sorted(sizes, reverse=True) and len(set(sizes)) <= 2
    assert all(np.array_equal(g, np.sort(g)) for g in groups)


@pytest.mark.parametrize("side,typical_steps", [(20, 1.5), (51, 4.0)])
def test_compact_groups_are_spatially_compact_and_seeded(side, typical_steps):
    """k-means groups of a raster: the typical (median) group extends a few steps from its centroid
    (groups of 4 on 20 x 20, of 26 on 51 x 51), the mean spread is well below that of the random
    partitions (balancing moves a few frames into far groups), and the result depends only on
    the seed."""
    from ptychopy_amd import _driver as D
    step = 14.0
    pos = _raster(side, step=step, jitter=1.0)
    groups = D.mls_compact_groups(pos, seed=5)
    spread = np.array([np.abs(pos[g] - pos[g].mean(axis=0)).max() for g in groups])
    parts = D.mls_partitions(len(pos), 1, 5)[0]
    rand_spread = np.mean([np.abs(pos[g] - pos[g].mean(axis=0)).max() for g in parts])
    assert np.median(spread) <= typical_steps * step
    assert spread.mean() < 0.5 * rand_spread
    again = D.mls_compact_groups(pos, seed=5)
    assert all(np.array_equal(a, b) for a, b in zip(groups, again))


def test_compact_groups_small_scan_is_one_frame_per_group():
    from ptychopy_amd import _driver as D
    pos = _raster(7)
    groups = D.mls_compact_groups(pos, seed=0)
    assert len(groups) == 49 and all(len(g) == 1 for g in groups)
