"""End to end through the Python `ptychopy` surface and the C++ IPhaser (GPU)."""
import os

import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

LAMBD, Z, DX_D = 0.248e-9, 1.0, 172e-6


def _dataset(m=64, n=5, step=6e-8, seed=0):
    """Frames simulated on the driver's own Cartesian mesh (same arithmetic as the run)."""
    import oracle
    from ptychopy_amd import _driver as D, sim
    dx_s = D.sample_pixel_size(LAMBD, Z, m, DX_D)
    pos, canvas, _ = D.cartesian_mesh(n, n, step, step, dx_s, m)
    rng = np.random.default_rng(seed)
    obj = sim.random_field_object(canvas, rng)
    probe = sim.gaussian_probe(m, fwhm=m / 3)
    amp = sim.poisson_amplitudes(oracle.simulate(obj, probe, pos, squared=True), 1e6, rng)
    kw = dict(size=m, scanDimsx=n, scanDimsy=n, stepx=step, stepy=step, lambd=LAMBD, z=Z, dx_d=DX_D,
              beamSize=1e-7)
    return amp, pos, canvas, kw, dx_s


def test_dmnp_matches_engine_and_writes_results(gpu, tmp_path, monkeypatch):
    from ptychopy_amd import _driver as D, ptychopy
    from ptychopy_amd._engine import Engine, FLAG_PHASE_CONSTRAINT, FLAG_UPDATE_PROBE
    monkeypatch.chdir(tmp_path)
    amp, pos, canvas, kw, dx_s = _dataset()
    assert ptychopy.dmnp(jobID="t", diffractionNP=amp.astype(np.float64), iter=12, **kw) is True
    obj = D.load_complex_csv("data/t_object_0.csv")
    prb = D.load_complex_csv("data/t_probes_0.csv")
    assert obj.shape == canvas and prb.shape == (64, 64)
    with Engine(len(pos), 64, canvas) as e:
        e.set_frames(amp)
        e.set_positions(pos)
        e.clear_object()
        e.init_probe(1e-7, dx_s)
        for i in range(1, 13):
            e.dm_iteration((FLAG_PHASE_CONSTRAINT if i <= 1 else 0) | (FLAG_UPDATE_PROBE if i >= 10 else 0))
        ref = e.get_object()
    assert rel_err(obj, ref) < 2e-6  # CSV keeps 6 significant digits


def test_epie_step_mode_equals_whole_run(gpu, tmp_path, monkeypatch):
    from ptychopy_amd import ptychopy
    monkeypatch.chdir(tmp_path)
    amp, pos, canvas, kw, _ = _dataset(seed=1)
    assert ptychopy.epienpinit(jobID="s", diffractionNP=amp, iter=4, seed=3, **kw) is True
    for _ in range(4):
        assert ptychopy.epiestep() is True
    o_step = ptychopy.epieresobj()
    p_step = ptychopy.epieresprobe()
    assert ptychopy.epiepost() is True
    ph = ptychopy._make_phaser("ePIE", *ptychopy._collect((), dict(diffractionNP=amp, iter=4, seed=3, **kw), True))
    ph.init()
    ph.add_method("ePIE", 4)
    ph.phase()
    assert np.array_equal(ph.get_object(), o_step)
    assert np.array_equal(ph.get_probe(), p_step)
    assert p_step.shape == (64, 64) and np.isfinite(o_step).all()
    ph.close()


def test_cmdstr_simulated_run_returns_the_object(gpu, tmp_path, monkeypatch):
    from ptychopy_amd import ptychopy
    monkeypatch.chdir(tmp_path)
    o = ptychopy.epiecmdstr("-jobID=c -simulate=1 -scanDims=4,4 -step=8e-8 -size=64 -lambda=0.248e-9 -z=1 "
                            "-dx_d=172e-6 -beamSize=1e-7 -i=3")
    assert o.dtype == np.complex64 and o.ndim == 2 and np.isfinite(o).all()
    assert os.path.exists("data/c_object_0.csv") and os.path.exists("data/c_probes_0.csv")


def test_dm_whole_mode_with_position_list(gpu, tmp_path, monkeypatch):
    from ptychopy_amd import _driver as D, ptychopy
    monkeypatch.chdir(tmp_path)
    amp, pos, canvas, kw, dx_s = _dataset(seed=2)
    # the same scan given as a list in metres (positionNP): list-mesh geometry
    metres = (pos.astype(np.float64) - pos.mean(axis=0)) * dx_s
    lpos, lcanvas, _ = D.list_mesh(metres, dx_s, 64)
    ph = ptychopy._make_phaser("DM", *ptychopy._collect((), dict(diffractionNP=amp, positionNP=metres, iter=3,
                                                                  **kw), True))
    ph.init()
    assert np.array_equal(ph.get_positions(), lpos)
    ph.add_method("DM", 3)
    ph.phase()
    assert ph.get_object().shape == lcanvas
    ph.close()


# ---------------------------------------------------------------- MLs through the driver
MLS_CMD = ("-jobID=m -simulate=1 -scanDims=6,6 -step=5e-8 -size=64 -lambda=0.248e-9 -z=1 -dx_d=172e-6 "
           "-beamSize=1e-7 -probeModes=2 -PPS=5")


def test_mls_simulated_run_lowers_the_fourier_error(gpu, tmp_path, monkeypatch):
    from ptychopy_amd import ptychopy
    monkeypatch.chdir(tmp_path)
    ph = ptychopy.Phaser()
    ph.parse_cmdline(MLS_CMD + " -i=9")
    ph.set_param("algorithm", "MLs")
    ph.init()
    ph.add_method("MLs", 9)
    ph.phase()
    e1, e9 = ph.get_fourier_errors(1), ph.get_fourier_errors(9)
    assert np.isfinite(ph.get_object()).all() and np.isfinite(e9).all()
    assert e9.mean() < 0.7 * e1.mean()
    assert np.array_equal(ph.get_errors(9), np.ones(9))  # MLS::iteration returns 1 (MLS.cpp:489)
    ph.close()


def test_mls_initial_modes_are_orthogonal_by_power(gpu, tmp_path, monkeypatch):
    from ptychopy_amd import ptychopy
    monkeypatch.chdir(tmp_path)
    assert ptychopy.mlsinit(MLS_CMD.replace("-probeModes=2", "-probeModes=3") + " -i=2") is True
    P = ptychopy.mlsresprobe().reshape(3, -1).astype(np.complex128)
    G = P.conj() @ P.T
    pw = np.real(np.diag(G))
    assert np.all(np.diff(pw) < 0)
    off = np.abs(G - np.diag(np.diag(G)))
    assert off.max() < 1e-5 * pw.max()
    assert ptychopy.mlspost() is True


def test_mls_step_mode_equals_whole_run(gpu, tmp_path, monkeypatch):
    from ptychopy_amd import ptychopy
    monkeypatch.chdir(tmp_path)
    cmd = MLS_CMD + " -i=6"
    assert ptychopy.mlsinit(cmd) is True
    for _ in range(6):
        assert ptychopy.mlsstep() is True
    o_step = ptychopy.mlsresobj()
    p_step = ptychopy.mlsresprobe()
    assert ptychopy.mlspost() is True
    o_whole = ptychopy.mlscmdstr(cmd)
    assert np.array_equal(o_step, o_whole)
    assert p_step.shape == (128, 64)


def test_mlsnp_with_all_probe_modes_given(gpu, tmp_path, monkeypatch):
    import oracle
    from problems import mls_modes
    from ptychopy_amd import _driver as D, ptychopy, sim
    monkeypatch.chdir(tmp_path)
    m, n, step = 64, 6, 5e-8
    dx_s = D.sample_pixel_size(LAMBD, Z, m, DX_D)
    raw = D.cartesian_mesh_raw(n, n, step, step, dx_s, m)
    ext = np.array([min(raw[:, 0].min(), 0), min(raw[:, 1].min(), 0), max(raw[:, 0].max(), 0),
                    max(raw[:, 1].max(), 0)], np.float32)
    (rows, cols), _ = D.mls_canvas(raw, ext, m)
    rng = np.random.default_rng(5)
    obj = sim.random_field_object((rows, cols), rng)
    modes = mls_modes(m, 2, rng)
    sim_pos = raw + np.float32([np.ceil(rows / 2 - m / 2) - 1, np.ceil(cols / 2 - m / 2) - 1])
    inten = sum(oracle.simulate(obj, modes[k], sim_pos, squared=True) for k in range(2))
    inten = rng.poisson(inten / inten.sum(axis=(1, 2), keepdims=True) * 1e6).astype(np.float64)
    ph = ptychopy._make_phaser("MLs", *ptychopy._collect(
        (), dict(diffractionNP=inten, probeNP=modes.reshape(2 * m, m), iter=8, size=m, scanDimsx=n, scanDimsy=n,
                 stepx=step, stepy=step, lambd=LAMBD, z=Z, dx_d=DX_D, probeModes=2, PPS=5), True))
    ph.init()
    ph.add_method("MLs", 8)
    ph.phase()
    assert ph.get_object().shape == (rows, cols)
    assert ph.get_fourier_errors(8).mean() < 0.5 * ph.get_fourier_errors(1).mean()
    ph.close()


def test_mls_compact_groups_run_lowers_the_fourier_error(gpu, tmp_path, monkeypatch):
    """MLs over the compact (k-means) groups of get_close_indices (-MLC=1): the solver runs on the
    one compact partition and reduces the Fourier error as with the random partitions."""
    from ptychopy_amd import ptychopy
    monkeypatch.chdir(tmp_path)
    ph = ptychopy.Phaser()
    ph.parse_cmdline(MLS_CMD + " -i=9 -MLC=1")
    ph.set_param("algorithm", "MLs")
    ph.init()
    ph.add_method("MLs", 9)
    ph.phase()
    e1, e9 = ph.get_fourier_errors(1), ph.get_fourier_errors(9)
    assert np.isfinite(ph.get_object()).all() and np.isfinite(e9).all()
    assert e9.mean() < 0.7 * e1.mean()
    ph.close()


# ---------------------------------------------------------------- input files (lf=, objectGuess, probeGuess)
def test_positions_file_equals_position_list(gpu, tmp_path, monkeypatch):
    """`lf=` (ListScanMesh::loadList, ListScanMesh.cpp:64-89): `x,y` lines in metres, parsed into
    float, give the same list-mesh geometry as the same values passed as positionNP."""
    from ptychopy_amd import _driver as D, ptychopy
    monkeypatch.chdir(tmp_path)
    amp, pos, canvas, kw, dx_s = _dataset(seed=4)
    metres = ((pos.astype(np.float64) - pos.mean(axis=0)) * dx_s).astype(np.float32)
    with open("pos.txt", "w") as f:
        for x, y in metres:
            f.write(f"{float(x):.9g},{float(y):.9g}\n")
    ph_file = ptychopy._make_phaser("DM", *ptychopy._collect((), dict(diffractionNP=amp, lf="pos.txt", iter=2, **kw),
                                                               True))
    ph_file.init()
    ph_list = ptychopy._make_phaser("DM", *ptychopy._collect(
        (), dict(diffractionNP=amp, positionNP=metres.astype(np.float64), iter=2, **kw), True))
    ph_list.init()
    lpos, lcanvas, _ = D.list_mesh(metres.astype(np.float64), dx_s, 64)
    assert np.array_equal(ph_file.get_positions(), ph_list.get_positions())
    assert np.array_equal(ph_file.get_positions(), lpos)
    for ph in (ph_file, ph_list):
        ph.add_method("DM", 2)
        ph.phase()
    assert np.array_equal(ph_file.get_object(), ph_list.get_object())
    ph_file.close()
    ph_list.close()
    # a malformed line is an error, not a silent partial mesh
    with open("bad.txt", "w") as f:
        f.write("1e-7;2e-7\n")
    ph = ptychopy._make_phaser("DM", *ptychopy._collect((), dict(diffractionNP=amp, lf="bad.txt", iter=1, **kw), True))
    with pytest.raises(RuntimeError, match="positions file"):
        ph.init()
    ph.close()


@pytest.mark.parametrize("algorithm", ["DM", "MLs"])
def test_object_and_probe_guess_files(gpu, tmp_path, monkeypatch, algorithm):
    """`objectGuess=` is centred in the canvas (Sample::loadGuess -> h_setObjectArray,
    SampleKernels.cu:626-640: offset rows/2 - r/2, cols/2 - c/2); `probeGuess=` becomes mode 0
    (Probe::init, Probe.cpp:342-352).  MLs then rescales the object by one real factor
    (Probe::initEvo's normalisation), so its placement is compared up to that factor."""
    from ptychopy_amd import _driver as D, ptychopy
    monkeypatch.chdir(tmp_path)
    amp, pos, canvas, kw, dx_s = _dataset(seed=5)
    rng = np.random.default_rng(9)
    gr, gc = canvas[0] - 9, canvas[1] - 6
    guess = (rng.uniform(0.5, 1, (gr, gc)) * np.exp(1j * rng.uniform(-1, 1, (gr, gc)))).astype(np.complex64)
    yy, xx = np.mgrid[0:64, 0:64] - 32
    pguess = (np.exp(-(xx ** 2 + yy ** 2) / 200.0) * np.exp(0.2j * xx)).astype(np.complex64)
    assert D.save_complex("obj.csv", guess) and D.save_complex("probe.csv", pguess)
    gtext, ptext = D.load_complex_csv("obj.csv"), D.load_complex_csv("probe.csv")  # the values as parsed
    vals = dict(diffractionNP=amp ** 2 if algorithm == "MLs" else amp, objectGuess="obj.csv", probeGuess="probe.csv",
                iter=2, **kw)
    ph = ptychopy._make_phaser(algorithm, *ptychopy._collect((), vals, True))
    ph.init()
    ph.add_method(algorithm, 2)
    ph.phaseinit()
    o = ph.get_object().astype(np.complex128)
    want = np.zeros(o.shape, np.complex128)
    oi, oj = o.shape[0] // 2 - gr // 2, o.shape[1] // 2 - gc // 2
    want[oi:oi + gr, oj:oj + gc] = gtext
    if algorithm == "DM":
        assert np.array_equal(o, want)
        assert np.array_equal(ph.get_probe(), ptext)
    else:
        f = np.vdot(want, o) / np.vdot(want, want)
        assert abs(f.imag) < 1e-6 * abs(f) and f.real > 0
        assert rel_err(o, f * want) < 1e-6
        p0 = ph.get_probe()[:64].astype(np.complex128)
        fp = np.vdot(ptext, p0) / np.vdot(ptext, ptext)
        assert rel_err(p0, fp * ptext) < 1e-5  # mode 0 = the guess up to the power normalisation
    ph.close()
