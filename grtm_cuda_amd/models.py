"""Seeded layered-model generators and the benchmark configurations.

The random stream is splitmix64, so a seed names the same model everywhere
(the FP64 oracle and the GPU path receive the same arrays).
"""
import math

import numpy as np

_MASK = (1 << 64) - 1


class SplitMix64:
    """splitmix64 (Steele, Lea & Flood 2014); uniform doubles from the top 53 bits."""

    def __init__(self, seed):
        self.state = seed & _MASK

    def next_u64(self):
        self.state = (self.state + 0x9E3779B97F4A7C15) & _MASK
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _MASK
        return z ^ (z >> 31)

    def uniform(self, lo=0.0, hi=1.0):
        return lo + (hi - lo) * ((self.next_u64() >> 11) * (1.0 / (1 << 53)))

    def randint(self, lo, hi):
        """integer in [lo, hi] inclusive"""
        return lo + int(self.uniform() * (hi - lo + 1))


def layered_model(seed, n_layers=8, vs_top=2.8, vs_bottom=4.6):
    """n_layers finite layers over a half-space.

    Thickness U(1, 8) km; Vs rises linearly vs_top -> vs_bottom with +-5 %
    noise; Vp = Vs U(1.70, 1.80); rho = 1.74 Vp^0.25; Qs = 50 Vs^2; Qp = 2 Qs.
    Returns rows (thickness, vp, vs, rho, qp, qs); the half-space thickness is 0.
    """
    rng = SplitMix64(seed)
    n = n_layers + 1
    rows = np.zeros((n, 6))
    for i in range(n):
        frac = i / (n - 1) if n > 1 else 0.0
        h = rng.uniform(1.0, 8.0) if i < n - 1 else 0.0
        vs = (vs_top + (vs_bottom - vs_top) * frac) * (1.0 + rng.uniform(-0.05, 0.05))
        vp = vs * rng.uniform(1.70, 1.80)
        rho = 1.74 * vp ** 0.25
        qs = 50.0 * vs * vs
        rows[i] = (h, vp, vs, rho, 2.0 * qs, qs)
    return rows


# 3 layers over a half-space (thickness km, vp, vs, rho, qp, qs)
C1_LAYERS = np.array([
    [2.0, 5.0, 2.9, 2.5, 400.0, 200.0],
    [15.0, 6.1, 3.5, 2.75, 600.0, 300.0],
    [18.0, 6.6, 3.8, 2.9, 800.0, 400.0],
    [0.0, 8.0, 4.5, 3.3, 1200.0, 600.0],
])

C2_SEED = 20261203
C4_SEED = 20261204


def batch_models(seed, n_models, n_min=4, n_max=30, src_lo=0.0, src_hi=40.0):
    """n_models layered models from one seed: layer count U{n_min..n_max}
    (finite layers plus the half-space) and source depth U(src_lo, src_hi) km,
    each drawn with the layered_model generator.  Returns [(layers, src_depth)]."""
    rng = SplitMix64(seed)
    out = []
    for m in range(n_models):
        n_total = rng.randint(n_min, n_max)
        src = rng.uniform(src_lo, src_hi)
        out.append((layered_model(seed * 1000003 + m, n_total - 1), src))
    return out


def dense_grid(nw, dt, nk, vs_min, t0=0.0, taper=0.3):
    """Uniform grid: k_max = 1.5 omega_max / Vs_min, omega_I = pi / T."""
    T = 2 * nw * dt
    return dict(nw=nw, dt=dt, nk=nk, kmax=1.5 * (math.pi / dt) / vs_min, omega_i=math.pi / T, t0=t0,
                taper=taper)


def config(name, n_models=1024):
    """Workloads: c1 (3 layers, 8 receivers), c2 (8 layers, 256 receivers),
    c4 (n_models models of 4-30 layers, sources at 0-40 km, 64 receivers
    5-300 km, Nw 1024, Nk 8192, dt 0.05 s; one grid for the batch with k_max
    from the slowest shear velocity of all models)."""
    if name == "c1":
        layers = C1_LAYERS.copy()
        return dict(name="c1", layers=layers, src_depth=10.0, rcv_depth=0.0,
                    dists=np.arange(10.0, 81.0, 10.0),
                    grid=dense_grid(512, 0.1, 4096, layers[:, 2].min()))
    if name == "c2":
        layers = layered_model(C2_SEED, 8)
        return dict(name="c2", layers=layers, src_depth=12.0, rcv_depth=0.0,
                    dists=np.linspace(5.0, 500.0, 256),
                    grid=dense_grid(2048, 0.05, 16384, layers[:, 2].min()))
    if name == "c4":
        ms = batch_models(C4_SEED, n_models)
        vs_min = min(float(l[:, 2].min()) for l, _ in ms)
        return dict(name="c4", models=ms, rcv_depth=0.0, dists=np.linspace(5.0, 300.0, 64),
                    grid=dense_grid(1024, 0.05, 8192, vs_min))
    raise KeyError(name)


def analytic_kcut(layers, src_depth, rcv_depth, grid, tau=1e-7):
    """Truncation wavenumber per frequency row without the learned model:
    surface-wave poles lie below 1.15 w / Vs_min, and beyond them the integrand
    decays at least like exp(-(k - k_pole) |z_s - z_r|); keep ln(1/tau) decay
    lengths past the last pole."""
    nw, dt = grid["nw"], grid["dt"]
    w = 2 * math.pi * np.arange(nw) / (2 * nw * dt)
    vs_min = float(np.min(layers[:, 2]))
    dz = max(abs(src_depth - rcv_depth), 0.5)
    return np.minimum(1.15 * w / vs_min + math.log(1.0 / tau) / dz, grid["kmax"])


def check_window(src_depth, rcv_depth):
    """Width (km^-1) of the residual-check window beyond the cut: four decay
    lengths of the direct wave between source and receiver."""
    return 4.0 / max(abs(src_depth - rcv_depth), 0.5)
