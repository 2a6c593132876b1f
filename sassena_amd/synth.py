"""Seeded synthetic trajectories and q-vector sets for the five benchmark cases.

Positions are generated in FP64 and stored as float32 Angstrom without
periodic wrapping.  Coordinates are returned in the device's staging layout,
atom-major SoA ``xyz[3][NA][NF]`` (float32), which ``sx_stage`` accepts as
``SX_LAYOUT_SOA`` without a transpose; ``to_frames_aos`` and ``to_atoms_aos``
give the reference stagers' layouts (data_stager.cpp:111-113, 327-330).

Scattering lengths are neutron constants in fm (database ft=0, no background
correction, database.cpp:474-476).  Orientation sets are seeded normal
triplets normalised row by row, which is what the reference's
``vectors.type=file`` reader does with any file (parameters.cpp:932-944).
"""
from __future__ import annotations

import dataclasses
import math
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np

# Neutron coherent scattering lengths (fm) and the type mix of the coherent cases.
COHERENT_MIX = (("H", 0.5, -3.74), ("C", 0.3, 6.65), ("N", 0.1, 9.36), ("O", 0.1, 5.80))
# Incoherent case: 60 % H with b_inc = 25.27 fm, 40 % C/O with b ~ 0.
INCOHERENT_MIX = (("H", 0.6, 25.27), ("C", 0.2, 0.0), ("O", 0.2, 0.0))


@dataclasses.dataclass
class Trajectory:
    name: str
    xyz: np.ndarray          # float32 [3][NA][NF]
    b: np.ndarray            # float64 [NA]
    qlen: np.ndarray         # float64 [nshells], |q| of each shell (A^-1)
    orient: np.ndarray       # float64 [NM][3], unit rows
    kind: str = "all"        # "all" or "self"
    box: float = 0.0

    @property
    def NA(self):
        return self.xyz.shape[1]

    @property
    def NF(self):
        return self.xyz.shape[2]

    @property
    def NM(self):
        return self.orient.shape[0]

    @property
    def nshells(self):
        return self.qlen.shape[0]

    def terms(self):
        return self.nshells * self.NM * self.NA * self.NF

    def qvectors(self, shell):
        """q_m = |q| u_m (abstract_vectors_scatter_device.cpp:116-122)."""
        return self.qlen[shell] * self.orient


def to_frames_aos(xyz):
    """[3][NA][NF] -> [NF][NA][3] (DataStagerByFrame layout)."""
    return np.ascontiguousarray(np.transpose(xyz, (2, 1, 0)))


def to_atoms_aos(xyz):
    """[3][NA][NF] -> [NA][NF][3] (DataStagerByAtom layout)."""
    return np.ascontiguousarray(np.transpose(xyz, (1, 2, 0)))


def scattering_lengths(rng, NA, mix=COHERENT_MIX):
    counts = [int(round(f * NA)) for _, f, _ in mix]
    counts[0] += NA - sum(counts)
    b = np.concatenate([np.full(c, bl, np.float64) for c, (_, _, bl) in zip(counts, mix)])
    rng.shuffle(b)
    return b


def orientations(rng, NM):
    v = rng.normal(size=(NM, 3))
    n = np.sqrt(np.sum(v * v, axis=1))
    return v / n[:, None]


def linear_shells(q0, q1, n):
    """|q| of a linear scan as the reference builds it (a ``scans`` entry with
    exponent 1, parameters.cpp:1142-1162): the end points exactly, interior
    point j at from + powf(j / (n - 1), 1) * (to - from), where the
    single-precision powf rounds the fraction to float (parameters.cpp:1151),
    so the interior points sit about 6e-8 relative off the exact grid."""
    if n == 1:
        return np.array([(q0 + q1) / 2.0])
    out = []
    for j in range(n):
        if j == 0:
            out.append(q0)
        elif j == n - 1:
            out.append(q1)
        else:
            out.append(q0 + float(np.float32(j * 1.0 / (n - 1))) * (q1 - q0))
    return np.array(out, dtype=np.float64)


def exact_linear_shells(q0, q1, n):
    """|q| on the exact arithmetic grid q0 + j (q1 - q0) / (n - 1) (a q list
    given point by point, e.g. a vectors file)."""
    if n == 1:
        return np.array([(q0 + q1) / 2.0])
    return np.array([q0 + (j / (n - 1)) * (q1 - q0) if 0 < j < n - 1 else (q0 if j == 0 else q1)
                     for j in range(n)], dtype=np.float64)


# Atoms per independently seeded block of random_walk: the result depends
# only on the seed, not on how many threads generate it.
_WALK_BLOCK = 2048


def _threads():
    try:
        return max(1, min(16, len(os.sched_getaffinity(0))))
    except AttributeError:
        return max(1, min(16, os.cpu_count() or 1))


def random_walk(rng, NA, NF, box, sigma, chunk=256):
    """r(0) ~ U[0,box)^3, r(t+1) = r(t) + N(0, sigma^2) per component, in FP64,
    stored as float32.  One seed is drawn from rng; atoms are generated in
    blocks of _WALK_BLOCK, each from its own child of that seed, on up to 16
    threads (numpy releases the GIL while it fills arrays)."""
    xyz = np.empty((3, NA, NF), np.float32)
    seed = int(rng.integers(0, 2 ** 63))
    nblk = (NA + _WALK_BLOCK - 1) // _WALK_BLOCK
    children = np.random.SeedSequence(seed).spawn(nblk)

    def block(k):
        g = np.random.default_rng(children[k])
        a0, a1 = k * _WALK_BLOCK, min(NA, (k + 1) * _WALK_BLOCK)
        pos = g.uniform(0.0, box, size=(3, a1 - a0))
        t = 0
        while t < NF:
            n = min(chunk, NF - t)
            steps = g.normal(0.0, sigma, size=(3, a1 - a0, n))
            if t == 0:
                steps[:, :, 0] = 0.0
            path = pos[:, :, None] + np.cumsum(steps, axis=2)
            xyz[:, a0:a1, t:t + n] = path
            pos = path[:, :, -1]
            t += n

    with ThreadPoolExecutor(_threads()) as ex:
        list(ex.map(block, range(nblk)))
    return xyz


def ou_walk(rng, NA, NF, box, sigma, tau):
    """Ornstein-Uhlenbeck walk around mu ~ U[box], stationary std sigma."""
    xyz = np.empty((3, NA, NF), np.float32)
    mu = rng.uniform(0.0, box, size=(3, NA))
    decay = math.exp(-1.0 / tau)
    kick = sigma * math.sqrt(1.0 - math.exp(-2.0 / tau))
    pos = mu + sigma * rng.normal(size=(3, NA))
    for t in range(NF):
        xyz[:, :, t] = pos
        pos = mu + (pos - mu) * decay + kick * rng.normal(size=(3, NA))
    return xyz


def lattice_vibrations(rng, NA, NF, spacing, sigma, tau):
    """Atoms on a simple cubic lattice (the first NA sites of the smallest
    cube holding them), each doing an Ornstein-Uhlenbeck vibration (stationary
    std sigma per component, time constant tau frames) about its site: a
    crystal, whose orientational anisotropy persists for the whole run."""
    n = int(math.ceil(NA ** (1.0 / 3.0)))
    g = np.stack(np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij"), axis=0).reshape(3, -1)
    sites = spacing * g[:, :NA].astype(np.float64)
    xyz = np.empty((3, NA, NF), np.float32)
    decay = math.exp(-1.0 / tau)
    kick = sigma * math.sqrt(1.0 - math.exp(-2.0 / tau))
    u = sigma * rng.normal(size=(3, NA))
    for t in range(NF):
        xyz[:, :, t] = sites + u
        u = u * decay + kick * rng.normal(size=(3, NA))
    return xyz


def case(name, seed=None, scale=1.0):
    """Build one of c1..c5 (or a scaled-down copy with scale<1 for tests)."""
    seeds = {"c1": 101, "c2": 202, "c3": 303, "c5": 505}
    rng = np.random.default_rng(seeds[name] if seed is None else seed)
    if name == "c1":
        NA, NF, box, ns, NM, q0, q1 = 1000, 200, 30.0, 10, 50, 0.1, 2.0
    elif name == "c2":
        NA, NF, box, ns, NM, q0, q1 = 20000, 2000, 60.0, 40, 200, 0.1, 4.0
    elif name == "c3":
        NA, NF, ns, NM, q0, q1 = 50000, 5000, 20, 100, 0.2, 3.0
        box = (NA / 0.1) ** (1.0 / 3.0)
    elif name == "c5":
        NA, NF, box, ns, NM, q0, q1 = 100000, 10000, 100.0, 60, 1000, 0.05, 5.0
    else:
        raise ValueError(name)
    if scale != 1.0:
        NA = max(1, int(NA * scale))
        NF = max(2, int(NF * scale))
        NM = max(1, int(NM * scale))
        ns = max(1, int(ns * scale))
    if name == "c3":
        b = scattering_lengths(rng, NA, INCOHERENT_MIX)
        xyz = ou_walk(rng, NA, NF, box, 0.2, 50.0)
        kind = "self"
    else:
        b = scattering_lengths(rng, NA)
        xyz = random_walk(rng, NA, NF, box, 0.1)
        kind = "all"
    orient = orientations(rng, NM)
    return Trajectory(name, xyz, b, linear_shells(q0, q1, ns), orient, kind, box)


def c4_batch(seed=404, count=64, scale=1.0, frame_scale=None):
    """The uneven batch: count trajectories of log-uniform size.  scale
    shrinks the atom counts (and the frame counts unless frame_scale is
    given, which shrinks the frame counts alone)."""
    rng = np.random.default_rng(seed)
    fs = scale if frame_scale is None else frame_scale
    out = []
    for i in range(count):
        NA = int(round(math.exp(rng.uniform(math.log(500), math.log(50000))) * scale))
        NF = int(rng.integers(256, 4097) * fs)
        NA, NF = max(NA, 1), max(NF, 2)
        box = (NA / 0.1) ** (1.0 / 3.0)
        ns = int(rng.integers(10, 61))
        NM = int(rng.integers(50, 301))
        b = scattering_lengths(rng, NA)
        xyz = random_walk(rng, NA, NF, box, 0.1)
        orient = orientations(rng, NM)
        out.append(Trajectory(f"c4_{i}", xyz, b, linear_shells(0.1, 4.0, ns), orient, "all", box))
    return out
