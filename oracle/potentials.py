"""Oracle for the potential types and impurities: a numpy restatement.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the library's
device evaluation (itp_potential_create_type, itp_potential_add_impurities);
the product never imports this module.

Follows the reference's PotentialType::operator() formulas
(potentialtypes.hpp:96-380), the constructors' parameter counts and defaults
(potentialtypes.cpp:23-53, 116-500), the cell-centred grid
(datalayout.cpp:29-35) and the impurity formulas of ImpurityType::add_noise
(noise.cpp:25-88), evaluated in FP64.
"""
import math

import numpy as np

# name -> (defaults, accepted parameter counts): potentialtypes.cpp:23-53 and the constructors
TYPES = {
    "zero": ([], {0}),
    "harmonic": ([1.0, 0.0, 0.0], {0, 1, 3}),
    "elliptic": ([1.0, 1.6180339887498948482], {0, 2}),
    "prettyhardsquare": ([8.0], {0, 1}),
    "softpentagon": ([], {0}),
    "henonheiles": ([205.0 / 42.0, -13.0 / 3.0], {0, 2}),
    "gaussian": ([1.0, 1.0, 0.0, 0.0], {0, 2, 4}),
    "quartic": ([0.01], {0, 1}),
    "squareoscillator": ([8.0], {0, 1}),
    "power": ([4.0, 1.0], {0, 1, 2}),
    "ring": ([3.0, 1.0, 2.0, 0.0, 1.0], {0, 3, 5}),
    "cosh": ([1.0, 1.0], {0, 2}),
    "softstadium": ([1.0, 2.0, 100.0, 1.0, 10.0], {0, 3, 5}),
    "powerstadium": ([1.0, 2.0, 4.0], {0, 1, 3}),
}


def resolve(name, params):
    """Full parameter list with defaults (lengths as given, not halved)."""
    defaults, counts = TYPES[name]
    if len(params) not in counts:
        raise ValueError(f"{name} takes {sorted(counts)} parameters")
    if name == "powerstadium" and len(params) == 1:  # potentialtypes.cpp:476-480
        return [defaults[0], defaults[1], params[0]]
    return list(params) + defaults[len(params):]


def positions(sx, sy, dx):
    """Cell-centred positions (datalayout.cpp:29-35)."""
    px = np.array([(2 * x + 1 - sx) * 0.5 * dx for x in range(sx)])
    py = np.array([(2 * y + 1 - sy) * 0.5 * dx for y in range(sy)])
    return np.meshgrid(px, py)


def value(name, p, x, y):
    """V(x, y) of a resolved potential (potentialtypes.hpp:96-380)."""
    if name == "zero":
        return np.zeros_like(x)
    if name == "harmonic":
        return 0.5 * p[0] * ((x - p[1]) ** 2 + (y - p[2]) ** 2)
    if name == "elliptic":
        return 0.5 * (p[0] * x * x + p[1] * y * y)
    if name == "prettyhardsquare":
        return np.maximum(np.abs(2 * x / math.pi), np.abs(2 * y / math.pi)) ** p[0]
    if name == "softpentagon":
        return 0.5 * (x * x + y * y) * (1 + 0.5 * np.sin(5 * np.arctan2(y, x)))
    if name == "henonheiles":
        r2 = x * x + y * y
        return r2 * r2 + p[0] * r2 + p[1] * r2 ** 1.5 * np.cos(3 * np.arctan2(y, x))
    if name == "gaussian":
        return p[0] * np.exp(-((x - p[2]) ** 2 + (y - p[3]) ** 2) / (2 * p[1] * p[1]))
    if name == "quartic":
        xr, yr = (x - y) / math.sqrt(2), (x + y) / math.sqrt(2)
        return xr * xr * yr * yr / 2 + p[0] * (xr ** 4 + yr ** 4) / 4
    if name == "squareoscillator":
        return 0.5 * (np.abs(x) ** p[0] + np.abs(y) ** p[0])
    if name == "power":
        return 0.5 * p[1] * np.hypot(x, y) ** p[0]
    if name == "ring":
        z = np.abs(np.hypot(x, y) - p[0]) / p[1]
        return 0.5 * z ** p[2] + p[3] * np.exp(-((x - p[0]) ** 2 + y * y) / (2 * p[4] * p[4]))
    if name == "cosh":
        return p[0] * (np.cosh(np.hypot(x, y) / p[1]) - 1)
    if name in ("softstadium", "powerstadium"):
        R, half = p[0], p[1] / 2
        q = np.where(np.abs(x) <= half, y / R,
                     np.where(x < -half, np.hypot(x + half, y) / R, np.hypot(x - half, y) / R))
        if name == "softstadium":
            return p[2] / (1 + p[3] * np.exp(p[4] * (1 - q * q)))
        return 0.5 * np.abs(q) ** p[2]
    raise KeyError(name)


def potential(name, params, sx, sy, dx):
    X, Y = positions(sx, sy, dx)
    return value(name, resolve(name, params), X, Y)


def add_impurities(V, kind, data, kind_params, dx):
    """V plus impurities, one after another (noise.cpp:25-88, 92-110)."""
    sy, sx = V.shape
    X, Y = positions(sx, sy, dx)
    V = V.copy()
    for row in np.atleast_2d(data):
        rx, ry = row[0] - X, row[1] - Y
        r2 = rx * rx + ry * ry
        if kind == 0:
            V += row[2] * np.exp(-0.5 * (r2 / (row[3] * row[3])))
        elif kind == 1:
            V += kind_params[1] * r2 ** (-kind_params[0] / 2)
        elif kind == 2:
            R2 = kind_params[1] ** 2
            V += np.where(r2 < R2, kind_params[0] * np.sqrt(np.maximum(1 - r2 / R2, 0)), 0.0)
        else:
            def nearest(v, s):
                r = v / dx + (s - 1) / 2
                return int(math.floor(r + 0.5) if r > 0 else math.ceil(r - 0.5))
            V[nearest(row[1], sy), nearest(row[0], sx)] += row[2]
    return V
