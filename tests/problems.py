"""Small seeded ptychography problems shared by the parity tests."""
import numpy as np

from ptychopy_amd import sim


def make_problem(m=32, n_side=4, step=None, jitter=1.0, seed=0, photons=1e6, modes=1):
    """Object/probe/positions/amplitudes; amplitudes from the FP64 oracle simulator."""
    import oracle
    rng = np.random.default_rng(seed)
    step = step if step is not None else m / 4 + 0.37
    pos = sim.raster_positions(n_side, step, rng, jitter=jitter, origin=(1.0, 1.0))
    rows, cols = sim.canvas_for(pos, m)
    rows, cols = rows + 2, cols + 2
    obj = sim.random_field_object((rows, cols), rng, phase_std=0.5)
    probe = sim.gaussian_probe(m, fwhm=m / 3, defocus_edge=2.0)
    inten = oracle.simulate(obj, probe, pos, squared=True)
    amp = sim.poisson_amplitudes(inten, photons, rng)
    return dict(pos=pos, obj=obj, probe=probe, amp=amp, shape=(rows, cols), m=m, n=len(pos), k=modes)


def mls_box(pos, rows, cols):
    """calc_object_norm box of IPtychoScanMesh::precalculate_ROI (IPtychoScanMesh.cpp:179-184):
    1-based inclusive (x0, x1, y0, y1) from the scan extents including the origin."""
    mn = np.minimum(pos.min(axis=0), 0.0)
    mx = np.maximum(pos.max(axis=0), 0.0)
    return (int(np.round(rows * 0.5 + mn[0])), int(np.round(rows * 0.5 + mx[0])) - 1,
            int(np.round(cols * 0.5 + mn[1])), int(np.round(cols * 0.5 + mx[1])) - 1)


def mls_modes(m, k, rng, powers=(0.6, 0.2, 0.1, 0.06, 0.04)):
    """k mutually orthogonal probe modes, mode 0 the Gaussian test probe, powers decreasing."""
    p0 = sim.gaussian_probe(m, fwhm=m / 3, defocus_edge=2.0).astype(np.complex128)
    yy, xx = np.mgrid[0:m, 0:m] - m / 2
    modes = [p0]
    for i in range(1, k):
        poly = (xx / m) ** (i % 3) * (yy / m) ** (i // 3) + 0.3j * rng.standard_normal((m, m)) * 0.05
        v = p0 * (poly + (i + 1) * 0.2)
        for u in modes:
            v = v - np.vdot(u, v) / np.vdot(u, u) * u
        modes.append(v)
    modes = np.array(modes)
    tot = np.sum(np.abs(modes[0]) ** 2)
    pw = np.asarray(powers[:k], np.float64)
    for i in range(k):
        modes[i] *= np.sqrt(pw[i] / pw[0] * tot / np.sum(np.abs(modes[i]) ** 2))
    return modes


def mls_exit_amplitudes(st, orc):
    """Amplitudes the MLs model predicts for state st (FP64): sqrt(sum_k |FFT psi_k|^2) with
    psi_0 = O_roi * shift(evo0 P0 + evo1 V), psi_k = P_k O_roi."""
    k, m, _ = st.probe.shape
    n = st.pos.shape[0]
    out = np.zeros((n, m, m))
    for j in range(n):
        r0, c0 = st.roi[j]
        op = st.obj[r0:r0 + m, c0:c0 + m]
        vp = st.evo[0, j] * st.probe[0] + st.evo[1, j] * st.vmode
        vp = orc.mls_shift(vp, st.sub[j, 0], st.sub[j, 1])
        s = np.abs(np.fft.fft2(op * vp)) ** 2
        for kk in range(1, k):
            s += np.abs(np.fft.fft2(st.probe[kk] * op)) ** 2
        out[j] = np.sqrt(s)
    return out


def make_mls_problem(m=32, n_side=4, k=2, step=None, jitter=0.7, seed=0, phase_object=True, zero_evo01=True,
                     photons=None):
    """A small MLs problem around the origin (positions are not min-shifted for MLs,
    CartesianScanMesh.cpp:94-177), initialised with Probe::initEvo, with amplitudes consistent
    with the initialised model (photons=None) or Poisson-noisy ones."""
    import oracle
    rng = np.random.default_rng(seed)
    step = step if step is not None else m / 4 + 0.37
    c = (np.arange(n_side) - (n_side - 1) / 2) * step
    pos = np.stack(np.meshgrid(c, c, indexing="ij"), -1).reshape(-1, 2)
    pos = (pos + rng.uniform(-jitter, jitter, pos.shape)).astype(np.float32)
    ext = pos.max(axis=0) - np.minimum(pos.min(axis=0), 0)
    rows, cols = (int(np.ceil((ext[0] + m) * 1.2)), int(np.ceil((ext[1] + m) * 1.2)))
    if phase_object:
        obj = np.exp(1j * rng.uniform(-0.6, 0.6, (rows, cols)))
    else:
        obj = sim.random_field_object((rows, cols), rng, phase_std=0.5).astype(np.complex128)
    probe = mls_modes(m, k, rng)
    vmode = rng.uniform(0, 1, (m, m)) * np.exp(2j * np.pi * rng.uniform(0, 1, (m, m)))
    n = pos.shape[0]
    evo1 = 1e-6 * rng.standard_normal(n)
    if zero_evo01:
        evo1[:2] = 0.0
    st = oracle.MlsState(obj, probe, vmode, evo1, pos, np.zeros((n, m, m)))
    frame_power = np.full(n, float(np.sum(np.abs(probe) ** 2)) * m * m)
    st.init(frame_power)
    amp = mls_exit_amplitudes(st, oracle)
    if photons is not None:
        inten = amp ** 2
        inten = rng.poisson(inten / inten.sum(axis=(1, 2), keepdims=True) * photons).astype(np.float64)
        amp = np.sqrt(inten)
    st.amp = np.ascontiguousarray(amp)
    return dict(state=st, box=mls_box(pos, rows, cols), m=m, n=n, k=k, shape=(rows, cols), pos=pos,
                frame_power=frame_power, obj0=obj, probe0=probe, vmode0=vmode, evo1_0=evo1)


def mls_groups(n, g, rng):
    """Random partition of range(n) into groups of g (last one smaller), largest first."""
    idx = rng.permutation(n)
    groups = [idx[i:i + g] for i in range(0, n, g)]
    groups.sort(key=lambda a: -len(a))
    return groups
