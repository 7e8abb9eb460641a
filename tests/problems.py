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
