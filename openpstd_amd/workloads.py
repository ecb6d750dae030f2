"""Synthetic scene configurations used by the tests and the benchmark.

Each builder returns PSTDConfiguration objects with the common settings:
dx = 0.2 m, c = 340 m/s, rho = 1.2 kg/m^3, patch error 70 (window 32),
tfactRK 0.5, 50 PML cells, ampmax 20000, openPSTD's RK coefficients,
locally-reacting flags off, calctime = (steps + 0.5) * dt.

Cell counts are exact: world coordinates are chosen so that openPSTD's grid
scaling (int)(world / dx) in single precision (kernel/PSTDKernel.cpp:73-98,
170-179) returns the intended cells.

Sizes n with n + 64 a power of two (64, 192, 448, 960, 1984, 4032) are never
generated: openPSTD's wavenumber table and FFT length disagree for them.
"""
import numpy as np

from . import abi_types as T
from .config import PSTDConfiguration, DomainConf, RK_COEFFICIENTS, _f32

DX = _f32(0.2)
EXCLUDED_SIDES = {64, 192, 448, 960, 1984, 4032, 8128}


def _world(cells):
    """Smallest float32 world coordinate w with (int)(w / dx) == cells (float arithmetic)."""
    w = np.float32(cells) * np.float32(DX)
    dx = np.float32(DX)
    for _ in range(8):
        if int(np.float32(w / dx)) >= cells:
            break
        w = np.nextafter(w, np.float32(np.inf))
    if cells >= 0:
        assert int(np.float32(w / dx)) == cells, (cells, w)
    return float(w)


def _world_pos(cell_coord):
    """World coordinate whose grid position (world / dx, float32) equals cell_coord (a float)."""
    w = np.float32(cell_coord) * np.float32(DX)
    return float(w)


def base_settings(conf, steps):
    s = conf.Settings
    s.SetSoundSpeed(340.0)
    s.SetGridSpacing(DX)
    s.SetDensityOfAir(1.2)
    s.SetPatchError(70.0)
    s.SetFactRK(0.5)
    s.SetPMLCells(50)
    s.SetAttenuationOfPMLCells(20000.0)
    q = _f32(340.0 / DX)
    s.SetBandWidth(_f32(3 * 1e-6 * q * q))
    s.SetRKCoefficients(RK_COEFFICIENTS)
    s.SetSpectralInterpolation(False)
    s.SetSaveNth(steps + 1)
    dt = s.GetTimeStep()
    s.SetRenderTime(_f32((steps + 0.5) * dt))
    assert conf.num_steps() == steps, (conf.num_steps(), steps)


def domain(x, y, nx, ny, absorption=1.0):
    """Domain at cell (x, y) of nx x ny cells; absorption per edge (dict or scalar)."""
    d = DomainConf((_world(x), _world(y)), (_world(nx), _world(ny)))
    a = absorption if isinstance(absorption, dict) else {k: absorption for k in "TLBR"}
    for k in "TLBR":
        getattr(d, k).Absorption = float(a.get(k, 0.0))
    return d


def room(steps=500):
    """Single 100x50-cell air domain, PML on all sides, centre source, receiver at (+5 m, +2 m)."""
    conf = PSTDConfiguration()
    base_settings(conf, steps)
    conf.Domains.append(domain(0, 0, 100, 50, 1.0))
    conf.Speakers.append((10.0, 5.0, 0.0))    # domain centre, grid (50, 25)
    conf.Receivers.append((15.0, 7.0, 0.0))   # (+5 m, +2 m), grid cell (75, 35)
    return conf


def lshape(seed=0, steps=2000, absorptions=(0.0, 0.5, 1.0)):
    """L-shape of three 256x128 domains A(0,0) B(256,0) C(0,128); the 8 outer edges
    draw their absorption from `absorptions`; one source in A, four receivers."""
    rng = np.random.RandomState(seed)
    conf = PSTDConfiguration()
    base_settings(conf, steps)
    A = domain(0, 0, 256, 128, 0.0)
    B = domain(256, 0, 256, 128, 0.0)
    C = domain(0, 128, 256, 128, 0.0)
    # every edge draws an absorption; shared edges only enter corner-PML coefficients
    for d in (A, B, C):
        for e in "TLBR":
            getattr(d, e).Absorption = float(absorptions[rng.randint(len(absorptions))])
    conf.Domains += [A, B, C]
    sx, sy = rng.randint(20, 236), rng.randint(20, 108)
    conf.Speakers.append((_world_pos(sx + 0.5), _world_pos(sy + 0.5), 0.0))
    boxes = [(0, 0, 256, 128), (256, 0, 256, 128), (0, 128, 256, 128)]
    for _ in range(4):
        bx, by, bw, bh = boxes[rng.randint(3)]
        rx, ry = bx + rng.randint(1, bw - 1), by + rng.randint(1, bh - 1)
        conf.Receivers.append((_world_pos(rx + 0.25), _world_pos(ry + 0.25), 0.0))
    return conf


def lshape_spectral(steps=400):
    """The L-shape of seed 0 with spectral-interpolation receivers: its four
    seeded receivers (cell + 0.25) and six more at other offsets, beside the
    A|B (x = 256) and A|C (y = 128) interfaces and an outer edge."""
    conf = lshape(0, steps)
    for x, y in [(254.6, 60.3), (257.35, 70.8), (100.45, 126.9), (140.8, 129.15), (3.3, 40.7), (128.0, 64.5)]:
        conf.Receivers.append((_world_pos(x), _world_pos(y), 0.0))
    conf.receiver_mode = T.PSTD_RECEIVER_SPECTRAL
    return conf


def square(seed=0, steps=1000, n=4096, n_receivers=16, receiver_radius=None):
    """One n x n domain, PML on all sides, centre source, seeded receivers
    (anywhere in the domain, or within `receiver_radius` cells of the source)."""
    rng = np.random.RandomState(seed)
    conf = PSTDConfiguration()
    base_settings(conf, steps)
    conf.Domains.append(domain(0, 0, n, n, 1.0))
    conf.Speakers.append((_world_pos(n / 2), _world_pos(n / 2), 0.0))
    for _ in range(n_receivers):
        if receiver_radius is None:
            rx, ry = rng.randint(1, n - 1), rng.randint(1, n - 1)
        else:
            r = receiver_radius
            rx, ry = n // 2 + rng.randint(-r, r + 1), n // 2 + rng.randint(-r, r + 1)
        conf.Receivers.append((_world_pos(rx + 0.25), _world_pos(ry + 0.25), 0.0))
    return conf


def _side(rng, lo, hi):
    while True:
        v = int(rng.randint(lo, hi + 1))
        if v not in EXCLUDED_SIDES:
            return v


def batch_scene(rng, steps, lo=64, hi=1024):
    """One scene of 1-4 conforming domains (row, column or 2x2 grid), edge absorption U[0,1]."""
    layout = rng.randint(4)  # 0 single, 1 row, 2 column, 3 grid
    conf = PSTDConfiguration()
    base_settings(conf, steps)
    if layout == 0:
        cols, rows = [_side(rng, lo, hi)], [_side(rng, lo, hi)]
    elif layout == 1:
        cols, rows = [_side(rng, lo, hi) for _ in range(rng.randint(2, 5))], [_side(rng, lo, hi)]
    elif layout == 2:
        cols, rows = [_side(rng, lo, hi)], [_side(rng, lo, hi) for _ in range(rng.randint(2, 5))]
    else:
        cols, rows = [_side(rng, lo, hi) for _ in range(2)], [_side(rng, lo, hi) for _ in range(2)]
    y = 0
    for j, h in enumerate(rows):
        x = 0
        for i, w in enumerate(cols):
            d = domain(x, y, w, h, 0.0)
            for e in "TLBR":
                getattr(d, e).Absorption = float(np.float32(rng.uniform(0, 1)))
            conf.Domains.append(d)
            x += w
        y += h
    w0, h0 = cols[0], rows[0]
    conf.Speakers.append((_world_pos(w0 / 2), _world_pos(h0 / 2), 0.0))
    conf.Receivers.append((_world_pos(rng.randint(1, w0 - 1) + 0.25), _world_pos(rng.randint(1, h0 - 1) + 0.25), 0.0))
    return conf


def scene_batch(seed=0, steps=2000, n_scenes=64, lo=64, hi=1024):
    """64 independent scenes of seeded uneven rectangular domains."""
    rng = np.random.RandomState(seed)
    return [batch_scene(rng, steps, lo, hi) for _ in range(n_scenes)]


def source_batch(seed=0, steps=3000, n_sources=256, receivers_per_source=8, n=1024):
    """One n x n geometry as two (n/2) x n domains side by side, PML everywhere;
    returns a list of configurations, one per seeded source position, each with
    its own seeded receivers."""
    rng = np.random.RandomState(seed)
    confs = []
    for _ in range(n_sources):
        conf = PSTDConfiguration()
        base_settings(conf, steps)
        conf.Domains.append(domain(0, 0, n // 2, n, 1.0))
        conf.Domains.append(domain(n // 2, 0, n // 2, n, 1.0))
        sx, sy = rng.randint(8, n - 8), rng.randint(8, n - 8)
        conf.Speakers.append((_world_pos(sx + 0.5), _world_pos(sy + 0.5), 0.0))
        for _ in range(receivers_per_source):
            conf.Receivers.append((_world_pos(rng.randint(1, n - 1) + 0.25),
                                   _world_pos(rng.randint(1, n - 1) + 0.25), 0.0))
        confs.append(conf)
    return confs


def cells_per_stage(domain_sizes_and_flags):
    """Sum of nx*ny over non-rigid domains (PML included): the cell count of the metric."""
    return sum(nx * ny for nx, ny, rigid in domain_sizes_and_flags if not rigid)
