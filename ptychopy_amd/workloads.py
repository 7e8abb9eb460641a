"""Synthetic workloads at the benchmark configurations (seeded; data formed on the GPU).

dm_c1: plain DM at the reference simulation shape -- 30 x 30 raster, 512 x 512 detector
(lambda = 0.248 nm, z = 1 m, 172 um pixels => dx_s = 2.8159 nm), 50 nm steps (17.76 px),
1088 x 1088 random-field object, 110 nm FWHM Gaussian probe with a 2 rad defocus phase at the
edge, 1e7 photons per frame with Poisson noise, one probe mode.
mls_c2: MLs with mixed modes -- 51 x 51 raster, 256 x 256 frames, 100 nm steps = 14 px
(dx_s = 100/14 nm), object on the MLs canvas (about 1146 x 1146), five orthogonalised
Hermite-Gaussian modes (48 px fundamental FWHM: the epie_c0 probe's 24 px at 128^2 scaled to 256^2) with power fractions 0.6/0.2/0.1/0.06/0.04,
1e5 photons per frame with Poisson noise, delta_p = 0.1.
batch_c3: 16 independent DM scans of uneven size -- raster sides drawn from 20..51, detectors
from {128, 256, 512}^2, 1e5..1e7 photons (log-uniform); object/probe statistics of epie_c0
scaled to the detector (probe FWHM M/5.33, steps M/8 px with +-1 px jitter).
epie_c0: plain ePIE -- 12 x 12 raster, 16 px steps with +-1 px jitter, 128 x 128 frames,
320 x 320 object, 24 px FWHM probe, 1e6 photons per frame.
"""
from __future__ import annotations

import concurrent.futures as cf

import numpy as np

from . import sim
from ._driver import cartesian_mesh, cartesian_mesh_raw, mls_canvas, sample_pixel_size
from ._engine import Engine


def _poisson_threaded(inten: np.ndarray, photons: float, seed_seq: np.random.SeedSequence, chunks: int = 16):
    n = inten.shape[0]
    bounds = np.linspace(0, n, min(chunks, n) + 1).astype(int)
    seeds = seed_seq.spawn(len(bounds) - 1)
    out = np.empty(inten.shape, np.float32)

    def work(i):
        a, b = bounds[i], bounds[i + 1]
        out[a:b] = sim.poisson_amplitudes(inten[a:b], photons, np.random.default_rng(seeds[i]))

    with cf.ThreadPoolExecutor(max_workers=min(16, len(seeds))) as ex:
        list(ex.map(work, range(len(seeds))))
    return out


def _simulate(obj, probe, pos_sim, m, device=0):
    with Engine(1, m, (m, m), device=device) as e:
        return e.simulate(obj, probe, pos_sim, squared=True)


def dm_c1(n_side: int = 30, m: int = 512, photons: float = 1e7, sim_size: int = 1088, config_id: int = 1,
          device: int = 0) -> dict:
    lambd, z, dx_d = 0.248e-9, 1.0, 172e-6
    dx_s = sample_pixel_size(lambd, z, m, dx_d)
    step = 50e-9
    pos, canvas, _ = cartesian_mesh(n_side, n_side, step, step, dx_s, m)
    rng = sim.rng_for(config_id, 0)
    obj = sim.random_field_object((sim_size, sim_size), rng, phase_std=0.5)
    beam = 110e-9
    probe = sim.gaussian_probe(m, fwhm=beam / dx_s, defocus_edge=2.0)
    off = np.array([(sim_size - canvas[0]) // 2, (sim_size - canvas[1]) // 2], np.float32)
    inten = _simulate(obj, probe, pos + off, m, device)
    amp = _poisson_threaded(inten, photons, np.random.SeedSequence([sim.SEED_BASE + config_id, 1]))
    o0 = (int(off[0]), int(off[1]))
    return dict(name="dm_c1", method="DM", pos=pos, canvas=canvas, amp=amp, m=m, n=len(pos), dx_s=dx_s,
                beam=beam, obj_true=obj[o0[0]:o0[0] + canvas[0], o0[1]:o0[1] + canvas[1]], probe_true=probe,
                iterations=100)


def mls_c2(n_side: int = 51, m: int = 256, k: int = 5, photons: float = 1e5, fwhm: float = 48.0,
           config_id: int = 2, device: int = 0, phase_corr: float = 6.0) -> dict:
    dx_s = 100e-9 / 14.0
    step = 100e-9
    raw = cartesian_mesh_raw(n_side, n_side, step, step, dx_s, m)
    lo = np.minimum(raw.min(axis=0), 0) if n_side % 32 else raw.min(axis=0)
    hi = np.maximum(raw.max(axis=0), 0) if n_side % 32 else raw.max(axis=0)
    canvas, box = mls_canvas(raw, np.array([lo[0], lo[1], hi[0], hi[1]], np.float32), m)
    rng = sim.rng_for(config_id, 0)
    obj = sim.random_field_object(canvas, rng, phase_std=0.5, phase_corr=phase_corr)
    modes = sim.hermite_gaussian_modes(m, k, fwhm)
    # frames at the MLs ROI frame: origin round(pos + ceil(Np/2 - M/2)) - 1
    off = np.array([np.ceil(canvas[0] / 2 - m / 2) - 1, np.ceil(canvas[1] / 2 - m / 2) - 1], np.float32)
    inten = sum(_simulate(obj, modes[i], raw + off, m, device) for i in range(k))
    amp = _poisson_threaded(inten, photons, np.random.SeedSequence([sim.SEED_BASE + config_id, 1]))
    return dict(name="mls_c2", method="MLs", pos=raw, canvas=canvas, box=box, amp=amp, m=m, n=len(raw), k=k,
                dx_s=dx_s, obj_true=obj, probe_true=modes, iterations=100, photons=photons)


def batch_c3(n_scans: int = 16, config_id: int = 3, device: int = 0, sides=(20, 51), sizes=(128, 256, 512),
             photons=(1e5, 1e7)) -> list:
    rng = sim.rng_for(config_id, 0)
    scans = []
    for s_i in range(n_scans):
        side = int(rng.integers(sides[0], sides[1] + 1))
        m = int(rng.choice(sizes))
        ph = float(np.exp(rng.uniform(np.log(photons[0]), np.log(photons[1]))))
        srng = sim.rng_for(config_id, 100 + s_i)
        pos = sim.raster_positions(side, m / 8.0, srng, jitter=1.0, origin=(1.0, 1.0))
        canvas = sim.canvas_for(pos, m)
        obj = sim.random_field_object(canvas, srng, phase_std=0.5)
        probe = sim.gaussian_probe(m, fwhm=m * 24.0 / 128.0, defocus_edge=2.0)
        inten = _simulate(obj, probe, pos, m, device)
        amp = _poisson_threaded(inten, ph, np.random.SeedSequence([sim.SEED_BASE + config_id, 200 + s_i]))
        scans.append(dict(name=f"batch_c3_{s_i}", method="DM", pos=pos, canvas=canvas, amp=amp, m=m, n=len(pos),
                          photons=ph, obj_true=obj, probe_true=probe, iterations=50))
    return scans


def epie_c0(m: int = 128, photons: float = 1e6, sim_size: int = 320, config_id: int = 0, device: int = 0) -> dict:
    rng = sim.rng_for(config_id, 0)
    pos = sim.raster_positions(12, 16.0, rng, jitter=1.0, origin=(1.0, 1.0))
    canvas = sim.canvas_for(pos, m)
    obj = sim.random_field_object((sim_size, sim_size), rng, phase_std=0.5)
    probe = sim.gaussian_probe(m, fwhm=24.0, defocus_edge=2.0)
    off = np.array([(sim_size - canvas[0]) // 2, (sim_size - canvas[1]) // 2], np.float32)
    inten = _simulate(obj, probe, pos + off, m, device)
    amp = _poisson_threaded(inten, photons, np.random.SeedSequence([sim.SEED_BASE + config_id, 1]))
    o0 = (int(off[0]), int(off[1]))
    return dict(name="epie_c0", method="ePIE", pos=pos, canvas=canvas, amp=amp, m=m, n=len(pos), dx_s=1e-8,
                beam=24e-8, obj_true=obj[o0[0]:o0[0] + canvas[0], o0[1]:o0[1] + canvas[1]], probe_true=probe,
                iterations=50)


def dm_bytes_per_frame(m: int, k: int, obj_px: int, n: int) -> float:
    """Compulsory HBM bytes per frame-iteration of the psi-resident DM form:
    M^2 (2 [FP16 A] + 16 K [psi read + write]) + (16 obj_px + 16 K M^2) / N."""
    return m * m * (2 + 16 * k) + (16 * obj_px + 16 * k * m * m) / n


def mls_kind_bytes(kind: str, m: int, k: int, n: int, groups: int, refine: bool = True) -> float:
    """Algorithmic bytes of one MLs iteration's launches of a main-stream kind (all groups), at
    storage precision, for the three-pass form with probe modes 2..K deferred to their update:
    exit_row_pass reads the object patch and mode 0's illumination probe (8 + 8 B/px) and writes
    K exit-wave planes and the patch (8 K + 8), and with position refinement (refine) the patch's
    row derivative dX (8); modulus_column_pass reads and writes the K planes and reads FP16 A
    (16 K + 2); inverse_row_pass reads mode 0's plane and the patch and writes chi_0 and the
    row-shifted probe update (32; it forms the probe update's row shifts while the rows are in
    LDS); probe_modes reads planes 2..K, the patch and the illumination at the ROI
    (8 (K - 1) + 8 + 4)."""
    px = float(n) * m * m
    per_px = {"exit_row_pass": 8.0 * k + 24.0 + (8.0 if refine else 0.0), "modulus_column_pass": 16.0 * k + 2.0,
              "inverse_row_pass": 32.0, "probe_modes": 8.0 * (k - 1) + 12.0}
    if kind not in per_px:
        return 0.0
    return per_px[kind] * px + (8.0 * k * m * m * groups if kind == "exit_row_pass" else 0.0)


def mls_bytes_per_frame(m: int, obj_px: int, n: int, groups: int) -> float:
    """Compulsory HBM bytes per frame-iteration of MLs: the FP16 amplitudes (2 M^2) plus the
    canvas-sized arrays each group reads and writes (object, its direction and preconditioner,
    illumination: 36 B per canvas pixel per group), spread over the frames."""
    return 2.0 * m * m + 36.0 * obj_px * groups / n


def mls_flops_per_frame(m: int, k: int) -> float:
    """FFT work of one MLs frame-iteration, 5 n log2 n per complex n-point transform: forward
    and inverse 2D transforms per mode (2K), the probe's sub-pixel shift and its inverse (2) and
    the position solver's forward and inverse one-dimensional transforms along rows (dX) and
    along columns (dY), two 2D transforms' worth."""
    return (2 * k + 2 + 2) * 10.0 * m * m * np.log2(m)
