"""Seeded synthetic inputs and the pinned workload configurations.

Every random draw comes from a splitmix64 stream (the same generator as the
C ABI's itp_init_states), so the GPU path and the CPU oracle see identical
inputs. Potentials follow the reference's formulas: harmonic 0.5*(x^2+y^2)
(potentialtypes.hpp:96-113) and Gaussian impurities A*exp(-0.5 r^2/w^2)
summed into V (noise.cpp:25-41), evaluated on the cell-centred grid
(datalayout.cpp:29-35). The hard-wall stadium uses the stadium geometry of
potentialtypes.hpp:331-342 with V = 0 inside and V_wall outside.
"""
import numpy as np

SEED_BASE = 0x20120131  # Parameters::default_rngseed (parameters.cpp:26)
_MASK = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed):
        self.s = seed & _MASK

    def next(self):
        self.s = (self.s + 0x9E3779B97F4A7C15) & _MASK
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _MASK
        return z ^ (z >> 31)

    def uniform(self, a=0.0, b=1.0):
        return a + (b - a) * ((self.next() >> 11) * (1.0 / 9007199254740992.0))

    def randint(self, lo, hi):
        """uniform integer in [lo, hi]"""
        return lo + int(self.uniform() * (hi - lo + 1))


def positions(sx, sy, dx):
    """Cell-centred positions (datalayout.cpp:29-35)."""
    px = np.array([(2 * x + 1 - sx) * 0.5 * dx for x in range(sx)])
    py = np.array([(2 * y + 1 - sy) * 0.5 * dx for y in range(sy)])
    return px, py


def harmonic(sx, sy, dx):
    px, py = positions(sx, sy, dx)
    X, Y = np.meshgrid(px, py)
    return 0.5 * (X * X + Y * Y)


def gaussian_impurities(sx, sy, dx, count, seed, box, amp=(-2.0, 2.0), width=(0.3, 0.8)):
    """Sum of A exp(-0.5 r^2 / w^2) with positions U(box), A ~ U(amp), w ~ U(width)."""
    rng = SplitMix64(seed)
    px, py = positions(sx, sy, dx)
    X, Y = np.meshgrid(px, py)
    V = np.zeros((sy, sx))
    imps = []
    for _ in range(count):
        cx = rng.uniform(box[0], box[1])
        cy = rng.uniform(box[0], box[1])
        A = rng.uniform(*amp)
        w = rng.uniform(*width)
        imps.append((cx, cy, A, w))
        r2 = (cx - X) ** 2 + (cy - Y) ** 2
        V += A * np.exp(-0.5 * (r2 / (w * w)))
    return V, imps


def hard_stadium(sx, sy, dx, radius=1.0, straight=2.0, wall=1e4):
    """V = 0 inside the stadium (straight section of length `straight`, end caps of `radius`), wall outside."""
    px, py = positions(sx, sy, dx)
    X, Y = np.meshgrid(px, py)
    half = 0.5 * straight
    inside_rect = (np.abs(X) <= half) & (np.abs(Y) <= radius)
    inside_left = np.hypot(X + half, Y) <= radius
    inside_right = np.hypot(X - half, Y) <= radius
    inside = inside_rect | inside_left | inside_right
    return np.where(inside, 0.0, wall)


def ho_levels(n):
    """Analytic 2D isotropic oscillator levels 1,2,2,3,3,3,..."""
    out, E = [], 1
    while len(out) < n:
        out.extend([float(E)] * E)
        E += 1
    return np.array(out[:n])


def config(name):
    """Pinned loop parameters and inputs of a named workload.

    Returns (cfg, V) with cfg the ITPSystem parameters (dict) and V [sy, sx].
    """
    if name == "c1_ho64":
        # 2D isotropic oscillator on a 64x64 periodic grid on [-8,8]^2, N=20,
        # 2nd-order split, eps=0.1. Converged when the energy change per step
        # is below 1e-12 relative (sigma saturates at the splitting error of
        # the 2nd-order split at eps=0.1, far above 1e-8).
        sx = sy = 64
        L = 16.0
        cfg = dict(sx=sx, sy=sy, lenx=L, boundary=0, N=20, n_wanted=16, halforder=1, B=0.0,
                   eps_values=[0.1], eps_divisor=5.0, max_steps=2000, min_time_step=1e-9,
                   seed=SEED_BASE, timestep_test=("relchange", 1e-12), final_test=("relchange", 1e-12))
        return cfg, harmonic(sx, sy, L / sx)
    if name == "c2_dirichlet256":
        # 256x256 Dirichlet grid on [-10,10]^2, N=500 (400 wanted), oscillator
        # plus 40 Gaussian impurities, B=1, 4th-order multi-product split,
        # eps=0.05. Converged when the energy change per step is below 1e-9
        # relative (energy-change tests, as recommended for Dirichlet with B).
        sx = sy = 256
        L = 20.0
        dx = L / sx
        V = harmonic(sx, sy, dx)
        imp, _ = gaussian_impurities(sx, sy, dx, 40, SEED_BASE + 1, (-10.0, 10.0))
        cfg = dict(sx=sx, sy=sy, lenx=L, boundary=1, N=500, n_wanted=400, halforder=2, B=1.0,
                   eps_values=[0.05], eps_divisor=5.0, max_steps=2000, min_time_step=1e-9,
                   seed=SEED_BASE, timestep_test=("relchange", 1e-9), final_test=("relchange", 1e-9))
        return cfg, V + imp
    if name == "c2_small":
        # the C2 physics at a size the CPU oracle runs to convergence in seconds
        sx = sy = 64
        L = 20.0
        dx = L / sx
        V = harmonic(sx, sy, dx)
        imp, _ = gaussian_impurities(sx, sy, dx, 40, SEED_BASE + 1, (-10.0, 10.0))
        cfg = dict(sx=sx, sy=sy, lenx=L, boundary=1, N=40, n_wanted=32, halforder=2, B=1.0,
                   eps_values=[0.05], eps_divisor=5.0, max_steps=2000, min_time_step=1e-9,
                   seed=SEED_BASE, timestep_test=("relchange", 1e-9), final_test=("relchange", 1e-9))
        return cfg, V + imp
    if name == "c3_ho512":
        sx = sy = 512
        L = 16.0
        cfg = dict(sx=sx, sy=sy, lenx=L, boundary=0, N=2000, n_wanted=1600, halforder=1, B=0.0,
                   eps_values=[0.1], eps_divisor=5.0, max_steps=2000, min_time_step=1e-9,
                   seed=SEED_BASE, timestep_test=("relchange", 1e-9), final_test=("relchange", 1e-9))
        return cfg, harmonic(sx, sy, L / sx)
    if name == "c3_ho128":
        # the C3 physics (oscillator, periodic, 2nd-order split, eps 0.1) at a
        # size the CPU oracle runs to convergence in minutes: 128x128 on
        # [-8,8]^2, N = 200 (160 wanted)
        sx = sy = 128
        L = 16.0
        cfg = dict(sx=sx, sy=sy, lenx=L, boundary=0, N=200, n_wanted=160, halforder=1, B=0.0,
                   eps_values=[0.1], eps_divisor=5.0, max_steps=2000, min_time_step=1e-9,
                   seed=SEED_BASE, timestep_test=("relchange", 1e-9), final_test=("relchange", 1e-9))
        return cfg, harmonic(sx, sy, L / sx)
    if name == "c5_stadium256":
        # the C5 physics (hard-wall stadium, Dirichlet, B = 2, 6th-order split,
        # eps 0.02, recovery on) at 256x256 on [-2.5,2.5]^2 with N = 100 (80 wanted)
        sx = sy = 256
        L = 5.0
        cfg = dict(sx=sx, sy=sy, lenx=L, boundary=1, N=100, n_wanted=80, halforder=3, B=2.0, recover=1,
                   eps_values=[0.02], eps_divisor=5.0, max_steps=3000, min_time_step=1e-9,
                   seed=SEED_BASE, timestep_test=("relchange", 1e-9), final_test=("relchange", 1e-9))
        return cfg, hard_stadium(sx, sy, L / sx)
    if name == "c5_stadium1024":
        # 1024x1024 Dirichlet grid on [-2.5,2.5]^2, N=4000 (3200 wanted), hard
        # stadium (V = 0 inside radius 1 / straight length 2, 1e4 outside),
        # B=2, 6th-order multi-product split, eps=0.02. Timed over a fixed
        # number of iterations (a full solve of 4000 states is not the point).
        # From noise, exp(-2 eps H) spans exp(-140) over 4000 states at
        # eps=0.02, far below FP64 resolution, so the overlap matrix has
        # non-positive eigenvalues: recovery (eps /= 5, fresh noise) is on, as
        # the reference's --recover does.
        sx = sy = 1024
        L = 5.0
        cfg = dict(sx=sx, sy=sy, lenx=L, boundary=1, N=4000, n_wanted=3200, halforder=3, B=2.0, recover=1,
                   eps_values=[0.02], eps_divisor=5.0, max_steps=2000, min_time_step=1e-9,
                   seed=SEED_BASE, timestep_test=("relchange", 1e-9), final_test=("relchange", 1e-9))
        return cfg, hard_stadium(sx, sy, L / sx)
    raise KeyError(name)


def c4_instances(count=64, seed=SEED_BASE):
    """The batched configuration: `count` independent instances, instance i
    drawn from the stream seeded base + i: sx = sy uniform in {128, 192, 256},
    N uniform in 50..400 (4/5 of them wanted), B uniform in [0, 2); V is the
    oscillator plus 40 Gaussian impurities (seed base + 1000 + i). Pinned
    where unspecified: L = 20, periodic, 4th-order multi-product split,
    eps = 0.05, energy-change tests at 1e-9 relative.
    Returns a list of (cfg, V)."""
    out = []
    for i in range(count):
        rng = SplitMix64(seed + i)
        n = (128, 192, 256)[rng.randint(0, 2)]
        N = rng.randint(50, 400)
        B = rng.uniform(0.0, 2.0)
        L = 20.0
        dx = L / n
        imp, _ = gaussian_impurities(n, n, dx, 40, seed + 1000 + i, (-10.0, 10.0))
        cfg = dict(sx=n, sy=n, lenx=L, boundary=0, N=N, n_wanted=(4 * N) // 5, halforder=2, B=B,
                   eps_values=[0.05], eps_divisor=5.0, max_steps=2000, min_time_step=1e-9,
                   seed=seed + i, timestep_test=("relchange", 1e-9), final_test=("relchange", 1e-9))
        out.append((cfg, harmonic(n, n, dx) + imp))
    return out
