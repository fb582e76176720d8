"""Synthetic particle sets for the FMM tick (float32 SoA, seeded).

Two families:

* the reference's own test generators, reproduced operation by operation in
  IEEE float32 (tests/accuracy.c:33-40 LCG, ``fill_uniform`` :111-121 and
  ``fill_two_clusters`` :123-137), so fixtures made from them match the C
  programs bit for bit;
* counter-based generators for the large workloads (uniform cube, Plummer
  sphere truncated at 20a, two Plummer spheres, the ragged ensemble), built
  on splitmix64 so every value depends only on (seed, stream, index).
"""
from __future__ import annotations

import numpy as np

F32 = np.float32


# ---------------------------------------------------------------------------
# glibc rand(): the input of the reference's benchmark (tests/benchmark.c)
# ---------------------------------------------------------------------------

def glibc_rand(seed: int, count: int) -> np.ndarray:
    """The first ``count`` values of glibc's rand() after srand(seed): its
    default additive feedback generator (TYPE_3, degree 31, separation 3).
    The state r[0..30] is seeded by r[0] = seed, r[i] = 16807 r[i-1] mod
    (2^31 - 1); r[31..33] = r[0..2]; then r[i] = r[i-3] + r[i-31] mod 2^32,
    the first 310 of those are discarded, and each output is r[i] >> 1."""
    seed = int(seed) & 0xFFFFFFFF
    r = [seed if seed != 0 else 1]
    for i in range(1, 31):
        # Schrage's method, signed 32-bit as in glibc
        prev = r[i - 1] if r[i - 1] < 2 ** 31 else r[i - 1] - 2 ** 32
        hi, lo = int(prev / 127773), prev - int(prev / 127773) * 127773
        word = 16807 * lo - 2836 * hi
        if word < 0:
            word += 2147483647
        r.append(word & 0xFFFFFFFF)
    for i in range(31, 34):
        r.append(r[i - 31])
    total = 344 + int(count)
    for i in range(34, total):
        r.append((r[i - 31] + r[i - 3]) & 0xFFFFFFFF)
    return (np.array(r[344:total], dtype=np.uint64) >> np.uint64(1)).astype(np.uint32)


def benchmark_cube(n: int, seed: int = 7331):
    """tests/benchmark.c:158-165: srand(7331); for each particle
    x, y, z = ((float)rand() / RAND_MAX) * 100.0F in turn; mass 1."""
    v = glibc_rand(seed, 3 * int(n)).astype(F32)  # (float)rand(), round to nearest
    u = (v / F32(2147483647.0)) * F32(100.0)        # RAND_MAX converted to float
    u = u.astype(F32).reshape(int(n), 3)
    return (np.ascontiguousarray(u[:, 0]), np.ascontiguousarray(u[:, 1]),
            np.ascontiguousarray(u[:, 2]), np.ones(int(n), F32))


# ---------------------------------------------------------------------------
# the reference's LCG generators (tests/accuracy.c)
# ---------------------------------------------------------------------------

def _lcg_units(seed: int, count: int) -> np.ndarray:
    """``rand_unit`` of tests/accuracy.c:33-40, ``count`` draws in order."""
    s = np.uint32(seed)
    out = np.empty(count, np.float32)
    a, c = np.uint32(1664525), np.uint32(1013904223)
    with np.errstate(over="ignore"):
        for k in range(count):
            s = np.uint32(s * a + c)
            out[k] = F32(np.uint32(s >> np.uint32(8))) / F32(16777215.0)
    return out


def ref_fill_uniform(n: int):
    """tests/accuracy.c:111-121: seed 7331, cube [-50,50)^3, m in [0.5,2)."""
    u = _lcg_units(7331, 4 * n).reshape(n, 4)
    x = u[:, 0] * F32(100.0) - F32(50.0)
    y = u[:, 1] * F32(100.0) - F32(50.0)
    z = u[:, 2] * F32(100.0) - F32(50.0)
    m = F32(0.5) + u[:, 3] * F32(1.5)
    return x, y, z, m


def ref_fill_two_clusters(n: int):
    """tests/accuracy.c:123-137: seed 20260803, clusters at x = -12.75 / +12.75."""
    u = _lcg_units(20260803, 4 * n).reshape(n, 4)
    half = n // 2
    centre = np.where(np.arange(n) < half, F32(-12.0), F32(12.0)).astype(F32)
    skew = np.where(np.arange(n) < half, F32(-0.75), F32(0.75)).astype(F32)
    x = (centre + skew) + ((u[:, 0] - F32(0.5)) * F32(2.0))
    y = (u[:, 1] - F32(0.5)) * F32(2.0)
    z = (u[:, 2] - F32(0.5)) * F32(2.0)
    m = F32(0.5) + (u[:, 3] * F32(2.0))
    return x.astype(F32), y.astype(F32), z.astype(F32), m.astype(F32)


# ---------------------------------------------------------------------------
# counter-based generators
# ---------------------------------------------------------------------------

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def _splitmix64(v: np.ndarray) -> np.ndarray:
    with np.errstate(over="ignore"):
        z = v + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def uniform01(seed: int, stream: int, n: int, offset: int = 0) -> np.ndarray:
    """U[0,1) with 53-bit resolution; value k depends on (seed, stream, k)."""
    idx = np.arange(offset, offset + n, dtype=np.uint64)
    with np.errstate(over="ignore"):
        key = (np.uint64(seed) * np.uint64(0x100000001B3)) ^ (np.uint64(stream) << np.uint64(48))
        bits = _splitmix64(_splitmix64(idx ^ key))
    return (bits >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def uniform_cube(n: int, seed: int = 1, lo: float = -1.0, hi: float = 1.0, mass=1.0):
    """i.i.d. uniform positions in [lo,hi)^3, equal masses."""
    x = (lo + (hi - lo) * uniform01(seed, 0, n)).astype(F32)
    y = (lo + (hi - lo) * uniform01(seed, 1, n)).astype(F32)
    z = (lo + (hi - lo) * uniform01(seed, 2, n)).astype(F32)
    m = np.full(n, mass, F32)
    return x, y, z, m


def plummer(n: int, seed: int = 1, a: float = 1.0, rmax_over_a: float = 20.0,
            centre=(0.0, 0.0, 0.0), mass=None, stream0: int = 0):
    """Plummer sphere: r = a / sqrt(u^(-2/3) - 1) with u = U * M(rmax)/M, so the
    truncation at rmax needs no rejection; isotropic directions."""
    q = rmax_over_a * rmax_over_a
    u_max = (q / (q + 1.0)) ** 1.5
    u = uniform01(seed, stream0 + 0, n) * u_max
    u = np.maximum(u, 1e-300)
    r = a / np.sqrt(u ** (-2.0 / 3.0) - 1.0)
    cth = 2.0 * uniform01(seed, stream0 + 1, n) - 1.0
    phi = 2.0 * np.pi * uniform01(seed, stream0 + 2, n)
    sth = np.sqrt(np.maximum(0.0, 1.0 - cth * cth))
    x = (centre[0] + r * sth * np.cos(phi)).astype(F32)
    y = (centre[1] + r * sth * np.sin(phi)).astype(F32)
    z = (centre[2] + r * cth).astype(F32)
    m = np.full(n, (1.0 / n) if mass is None else mass, F32)
    return x, y, z, m


def two_plummers(n: int, seed: int = 1, a: float = 1.0, sep: float = 5.0, v: float = 0.3):
    """Two Plummer spheres (n/2 each, masses 1/n) centred at (-sep,0,0) and
    (+sep,0,0) with bulk velocities (+v,0,0) and (-v,0,0).  Returns
    (x, y, z, m, vx, vy, vz)."""
    h = n // 2
    xa, ya, za, _ = plummer(h, seed, a, centre=(-sep, 0.0, 0.0), mass=1.0 / n, stream0=0)
    xb, yb, zb, _ = plummer(n - h, seed, a, centre=(sep, 0.0, 0.0), mass=1.0 / n, stream0=8)
    x = np.concatenate([xa, xb])
    y = np.concatenate([ya, yb])
    z = np.concatenate([za, zb])
    m = np.full(n, 1.0 / n, F32)
    vx = np.concatenate([np.full(h, v, F32), np.full(n - h, -v, F32)])
    vy = np.zeros(n, F32)
    vz = np.zeros(n, F32)
    return x, y, z, m, vx, vy, vz


def ensemble_sizes(n_systems: int = 512, seed: int = 1, lo: int = 2048, hi: int = 262144):
    """Per-system N, log-uniform in [lo, hi]."""
    u = uniform01(seed, 100, n_systems)
    return np.rint(np.exp(u * np.log(hi / lo)) * lo).astype(np.int64)


def ensemble(n_systems: int = 512, seed: int = 1, lo: int = 2048, hi: int = 262144):
    """The ragged batch: even systems uniform in [-1,1]^3, odd systems Plummer
    with a ~ U[0.5, 2]; masses 1/N_s.  Returns a list of (x, y, z, m)."""
    sizes = ensemble_sizes(n_systems, seed, lo, hi)
    a_s = 0.5 + 1.5 * uniform01(seed, 101, n_systems)
    out = []
    for s, ns in enumerate(sizes):
        ns = int(ns)
        if s % 2 == 0:
            out.append(uniform_cube(ns, seed * 100003 + s, mass=1.0 / ns))
        else:
            out.append(plummer(ns, seed * 100003 + s, float(a_s[s]), mass=1.0 / ns))
    return out


def sample_targets(n: int, k: int, seed: int = 4096) -> np.ndarray:
    """k distinct original indices by a seeded partial Fisher-Yates shuffle."""
    k = min(int(k), int(n))
    rng = np.random.Generator(np.random.PCG64(seed))
    perm = np.arange(n, dtype=np.int64) if n <= (1 << 22) else None
    if perm is not None:
        for i in range(k):
            j = int(rng.integers(i, n))
            perm[i], perm[j] = perm[j], perm[i]
        return perm[:k].astype(np.uint32)
    chosen = {}
    out = np.empty(k, np.int64)
    for i in range(k):
        j = int(rng.integers(i, n))
        vi, vj = chosen.get(i, i), chosen.get(j, j)
        chosen[j] = vi
        out[i] = vj
    return out.astype(np.uint32)
