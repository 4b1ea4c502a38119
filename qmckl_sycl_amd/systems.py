"""Molecular systems for the per-electron basis path: the ChBrClF golden
system of the reference tests and deterministic synthetic systems.

A :class:`System` carries exactly the arrays the qmckl setters take, in the
reference's layouts (``include/qmckl_gpu.hpp:513-801``):

* nucleus: ``charge[N]``, ``nucl_coord[3][N]`` (the ``'T'`` layout of
  ``qmckl_set_nucleus_coord_device``, ``src/qmckl_nucleus.cpp:542-610``);
* AO basis (type ``'G'``): ``nucleus_index[N]``, ``nucleus_shell_num[N]``,
  ``shell_ang_mom[S]`` (int32), ``shell_prim_num[S]``, ``shell_prim_index[S]``,
  ``shell_factor[S]``, ``exponent/coefficient/prim_factor[Np]``,
  ``ao_factor[A]`` (``src/qmckl_ao.cpp:402-1236``);
* MO basis: ``mo_coef[M][A]`` (``src/qmckl_mo.cpp:526-583``);
* electrons: ``elec_coord[walk][elec][3]`` (the ``'N'`` layout of
  ``qmckl_set_electron_coord_device``, ``src/qmckl_electron.cpp:50-90``;
  point ``p = w * elec_num + e``).

Synthetic systems follow the workload descriptions with fixed seeds:
N2-like (seed 1001), ChBrClF-sized (1002), 64-water cluster (1003), the
48-molecule ragged batch (1004) and the 64-water differentiation system
(1005).  Normalisation: primitives by
``(2a/pi)^(3/4) (4a)^(l/2) / sqrt((2l-1)!!)``, contracted shells to unit
norm, Cartesian components by ``sqrt((2l-1)!!/((2a-1)!!(2b-1)!!(2c-1)!!))``
(the convention of the ChBrClF fixture, ``tests/chbrclf.hpp:42-130``).
"""
from __future__ import annotations

import dataclasses
import math
import os
from typing import List, Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
REPO_ROOT = os.path.dirname(_HERE)
GOLDEN = os.path.join(REPO_ROOT, "tests", "golden", "chbrclf.npz")

BOHR_PER_ANGSTROM = 1.0 / 0.529177210903


def ncart(l: int) -> int:
    return (l + 1) * (l + 2) // 2


def cart_exponents(l: int):
    """Cartesian (a, b, c) of shell l in the reference order: a descending,
    then b descending (``src/qmckl_ao_sycl.cpp:291-331``)."""
    out = []
    for a in range(l, -1, -1):
        for b in range(l - a, -1, -1):
            out.append((a, b, l - a - b))
    return out


def dfact(n: int) -> int:
    r = 1
    while n > 1:
        r *= n
        n -= 2
    return r


def prim_norm(alpha: np.ndarray, l: int) -> np.ndarray:
    return (2.0 * alpha / math.pi) ** 0.75 * (4.0 * alpha) ** (0.5 * l) / math.sqrt(dfact(2 * l - 1))


def shell_norm(alpha: np.ndarray, coef: np.ndarray, l: int) -> float:
    ai = alpha[:, None]
    aj = alpha[None, :]
    ov = (2.0 * np.sqrt(ai * aj) / (ai + aj)) ** (l + 1.5)
    s = float(coef @ ov @ coef)
    return 1.0 / math.sqrt(s)


def ao_factors(l: int) -> List[float]:
    return [math.sqrt(dfact(2 * l - 1) / (dfact(2 * a - 1) * dfact(2 * b - 1) * dfact(2 * c - 1)))
            for (a, b, c) in cart_exponents(l)]


@dataclasses.dataclass
class System:
    name: str
    charge: np.ndarray
    nucl_coord: np.ndarray          # [3][N]
    nucleus_index: np.ndarray       # int64 [N]
    nucleus_shell_num: np.ndarray   # int64 [N]
    shell_ang_mom: np.ndarray       # int32 [S]
    shell_prim_num: np.ndarray      # int64 [S]
    shell_prim_index: np.ndarray    # int64 [S]
    shell_factor: np.ndarray        # [S]
    exponent: np.ndarray            # [Np]
    coefficient: np.ndarray         # [Np]
    prim_factor: np.ndarray         # [Np]
    ao_factor: np.ndarray           # [A]
    mo_coef: np.ndarray             # [M][A]
    elec_up_num: int
    elec_dn_num: int
    elec_coord: np.ndarray          # [walk][elec][3]
    note: str = ""

    @property
    def nucl_num(self) -> int:
        return int(self.charge.shape[0])

    @property
    def shell_num(self) -> int:
        return int(self.shell_ang_mom.shape[0])

    @property
    def prim_num(self) -> int:
        return int(self.exponent.shape[0])

    @property
    def ao_num(self) -> int:
        return int(self.ao_factor.shape[0])

    @property
    def mo_num(self) -> int:
        return int(self.mo_coef.shape[0])

    @property
    def elec_num(self) -> int:
        return int(self.elec_up_num + self.elec_dn_num)

    @property
    def walk_num(self) -> int:
        return int(self.elec_coord.shape[0])

    @property
    def point_num(self) -> int:
        return self.walk_num * self.elec_num

    def coord_T(self) -> np.ndarray:
        """Point coordinates ``[3][P]`` (the stored layout)."""
        return np.ascontiguousarray(self.elec_coord.reshape(-1, 3).T)

    def with_walkers(self, elec_coord: np.ndarray, name: Optional[str] = None) -> "System":
        return dataclasses.replace(self, elec_coord=np.ascontiguousarray(elec_coord),
                                   name=name or self.name)

    def summary(self) -> dict:
        return dict(name=self.name, nucl_num=self.nucl_num, shell_num=self.shell_num,
                    prim_num=self.prim_num, ao_num=self.ao_num, mo_num=self.mo_num,
                    elec_num=self.elec_num, walk_num=self.walk_num, point_num=self.point_num)

    def validate(self) -> None:
        N, S = self.nucl_num, self.shell_num
        assert self.nucl_coord.shape == (3, N)
        assert self.nucleus_index.shape == (N,) and self.nucleus_shell_num.shape == (N,)
        assert self.shell_prim_num.shape == (S,) and self.shell_prim_index.shape == (S,)
        assert self.shell_factor.shape == (S,)
        assert self.coefficient.shape == self.exponent.shape == self.prim_factor.shape
        assert int(sum(ncart(int(l)) for l in self.shell_ang_mom)) == self.ao_num
        assert self.mo_coef.shape[1] == self.ao_num
        assert self.elec_coord.shape[1:] == (self.elec_num, 3)


def chbrclf() -> System:
    """The ChBrClF system of ``tests/chbrclf.hpp`` (2 walkers x 68 electrons,
    263 AOs, 224 MOs)."""
    d = np.load(GOLDEN, allow_pickle=False)
    s = System(
        name="chbrclf",
        charge=d["charge"].astype(np.float64),
        nucl_coord=np.ascontiguousarray(d["nucl_coord_T"]),
        nucleus_index=d["nucleus_index"].astype(np.int64),
        nucleus_shell_num=d["nucleus_shell_num"].astype(np.int64),
        shell_ang_mom=d["shell_ang_mom"].astype(np.int32),
        shell_prim_num=d["shell_prim_num"].astype(np.int64),
        shell_prim_index=d["shell_prim_index"].astype(np.int64),
        shell_factor=d["shell_factor"].astype(np.float64),
        exponent=d["exponent"].astype(np.float64),
        coefficient=d["coefficient"].astype(np.float64),
        prim_factor=d["prim_factor"].astype(np.float64),
        ao_factor=d["ao_factor"].astype(np.float64),
        mo_coef=np.ascontiguousarray(d["mo_coef"]),
        elec_up_num=int(d["elec_up_num"]),
        elec_dn_num=int(d["elec_dn_num"]),
        elec_coord=np.ascontiguousarray(d["elec_coord"]),
        note="tests/chbrclf.hpp",
    )
    s.validate()
    return s


def chbrclf_mo_reference() -> np.ndarray:
    """``mo_vgl[136][5][224]`` of ``tests/mo_reference.txt``."""
    return np.load(GOLDEN, allow_pickle=False)["mo_reference"]


# ---------------------------------------------------------------------------
# Synthetic systems
# ---------------------------------------------------------------------------

class _BasisBuilder:
    def __init__(self):
        self.nucleus_index: List[int] = []
        self.nucleus_shell_num: List[int] = []
        self.l: List[int] = []
        self.prim_num: List[int] = []
        self.prim_index: List[int] = []
        self.shell_factor: List[float] = []
        self.expo: List[float] = []
        self.coef: List[float] = []
        self.pfac: List[float] = []
        self.ao_factor: List[float] = []

    def add_nucleus(self, shells, rng, expo_range, coef_range=(0.05, 1.0)):
        """shells: list of (l, nprim)."""
        self.nucleus_index.append(len(self.l))
        self.nucleus_shell_num.append(len(shells))
        lo, hi = math.log(expo_range[0]), math.log(expo_range[1])
        for (l, nprim) in shells:
            alpha = np.exp(rng.uniform(lo, hi, size=nprim))
            alpha = np.sort(alpha)[::-1].copy()
            c = rng.uniform(coef_range[0], coef_range[1], size=nprim)
            pf = prim_norm(alpha, l)
            sf = shell_norm(alpha, c, l)
            self.l.append(l)
            self.prim_num.append(nprim)
            self.prim_index.append(len(self.expo))
            self.shell_factor.append(sf)
            self.expo.extend(alpha.tolist())
            self.coef.extend(c.tolist())
            self.pfac.extend(pf.tolist())
            self.ao_factor.extend(ao_factors(l))

    def arrays(self):
        return dict(
            nucleus_index=np.array(self.nucleus_index, dtype=np.int64),
            nucleus_shell_num=np.array(self.nucleus_shell_num, dtype=np.int64),
            shell_ang_mom=np.array(self.l, dtype=np.int32),
            shell_prim_num=np.array(self.prim_num, dtype=np.int64),
            shell_prim_index=np.array(self.prim_index, dtype=np.int64),
            shell_factor=np.array(self.shell_factor, dtype=np.float64),
            exponent=np.array(self.expo, dtype=np.float64),
            coefficient=np.array(self.coef, dtype=np.float64),
            prim_factor=np.array(self.pfac, dtype=np.float64),
            ao_factor=np.array(self.ao_factor, dtype=np.float64),
        )


def _electrons_near(rng, centres: np.ndarray, owners: Sequence[int], walk_num: int, sigma=1.0):
    """[walk][elec][3]: electron e sits at centre[owners[e]] + N(0, sigma)."""
    base = centres[np.asarray(owners, dtype=np.int64)]          # [E][3]
    noise = rng.normal(0.0, sigma, size=(walk_num, len(owners), 3))
    return np.ascontiguousarray(base[None, :, :] + noise)


def n2_like(walk_num: int = 1, seed: int = 1001) -> System:
    """2 nuclei Z=7 at (0,0,+-1.0375) bohr; per nucleus 3s2p1d, 1-8 primitives,
    exponents log-uniform in [0.1, 1e4]; 30 AOs, 30 MOs ~ N(0, 0.3);
    14 electrons (7 up, 7 down) at nucleus + N(0, 1 bohr)."""
    rng = np.random.default_rng(seed)
    xyz = np.array([[0.0, 0.0, 1.0375], [0.0, 0.0, -1.0375]])
    bb = _BasisBuilder()
    for _ in range(2):
        shells = [(0, int(rng.integers(1, 9))) for _ in range(3)] + \
                 [(1, int(rng.integers(1, 9))) for _ in range(2)] + [(2, int(rng.integers(1, 9)))]
        bb.add_nucleus(shells, rng, (0.1, 1e4))
    arr = bb.arrays()
    A = int(sum(ncart(int(l)) for l in arr["shell_ang_mom"]))
    assert A == 30
    mo = rng.normal(0.0, 0.3, size=(30, A))
    owners = [0] * 7 + [1] * 7
    elec = _electrons_near(rng, xyz, owners, walk_num)
    s = System(name="n2_like", charge=np.array([7.0, 7.0]), nucl_coord=np.ascontiguousarray(xyz.T),
               mo_coef=mo, elec_up_num=7, elec_dn_num=7, elec_coord=elec, **arr)
    s.validate()
    return s


def chbrclf_sized(walk_num: int = 256, seed: int = 1002) -> System:
    """5 nuclei (Z = 6, 1, 35, 17, 9) uniform in a 4-bohr box, shells up to
    l=3 with 1-10 primitives (exponents log-uniform in [0.05, 5e5]),
    145 Cartesian AOs, 145 MOs ~ N(0, 0.2), 67 electrons (34 up, 33 down)
    Gaussian N(0, 1) around the nuclei (Z electrons per nucleus),
    ``walk_num`` walkers (default 256 -> 17,152 points)."""
    rng = np.random.default_rng(seed)
    Z = [6, 1, 35, 17, 9]
    xyz = rng.uniform(-2.0, 2.0, size=(5, 3))
    # shell recipe, 145 Cartesian AOs in total:
    #   C  3s2p1d      -> 3 + 6 + 6           = 15
    #   H  4s1p        -> 4 + 3               = 7
    #   Br 5s4p3d2f    -> 5 + 12 + 18 + 20    = 55
    #   Cl 4s3p2d1f    -> 4 + 9 + 12 + 10     = 35
    #   F  5s2p2d1f    -> 5 + 6 + 12 + 10     = 33
    recipe = {
        6: [0, 0, 0, 1, 1, 2],
        1: [0, 0, 0, 0, 1],
        35: [0, 0, 0, 0, 0, 1, 1, 1, 1, 2, 2, 2, 3, 3],
        17: [0, 0, 0, 0, 1, 1, 1, 2, 2, 3],
        9: [0, 0, 0, 0, 0, 1, 1, 2, 2, 3],
    }
    bb = _BasisBuilder()
    for z in Z:
        shells = [(l, int(rng.integers(1, 11))) for l in recipe[z]]
        bb.add_nucleus(shells, rng, (0.05, 5e5))
    arr = bb.arrays()
    A = int(sum(ncart(int(l)) for l in arr["shell_ang_mom"]))
    assert A == 145, A
    mo = rng.normal(0.0, 0.2, size=(145, A))
    owners = []
    for n, z in enumerate(Z):
        owners += [n] * z
    owners = owners[:67]
    # 6+1+35+17+9 = 68 -> drop the last F electron: 67 electrons
    elec = _electrons_near(rng, xyz, owners, walk_num)
    s = System(name="chbrclf_sized", charge=np.array(Z, dtype=np.float64),
               nucl_coord=np.ascontiguousarray(xyz.T), mo_coef=mo,
               elec_up_num=34, elec_dn_num=33, elec_coord=elec, **arr)
    s.validate()
    return s


def _random_rotation(rng) -> np.ndarray:
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    a, b, c, d = q
    return np.array([
        [a * a + b * b - c * c - d * d, 2 * (b * c - a * d), 2 * (b * d + a * c)],
        [2 * (b * c + a * d), a * a - b * b + c * c - d * d, 2 * (c * d - a * b)],
        [2 * (b * d - a * c), 2 * (c * d + a * b), a * a - b * b - c * c + d * d]])


def water_cluster(n_side: int = 4, walk_num: int = 128, seed: int = 1003,
                  mo_num: Optional[int] = None, zero_fraction: float = 0.7,
                  name: str = "water64") -> System:
    """n_side^3 waters (64 for n_side=4) at random orientations on a 3.0 A
    lattice (O-H 1.8088 bohr, H-O-H 104.52 deg).  Per O 4s3p2d1f shells, per
    H 3s2p1d, 1-9 primitives each, exponents log-uniform in [0.1, 1e4]
    (35 + 2*15 = 65 Cartesian AOs per water: 4,160 for 64 waters).
    MOs: one per electron pair (320 for 64 waters), coefficients N(0, 0.1)
    of which ``zero_fraction`` are zeroed by locality: MO i keeps only the AOs
    on the 30% of nuclei nearest a centre drawn uniformly among the nuclei.
    Electrons: 8 per O and 1 per H at the nucleus + N(0, 1 bohr)."""
    rng = np.random.default_rng(seed)
    spacing = 3.0 * BOHR_PER_ANGSTROM
    r_oh = 1.8088
    half = math.radians(104.52) / 2.0
    h1 = np.array([r_oh * math.sin(half), 0.0, r_oh * math.cos(half)])
    h2 = np.array([-r_oh * math.sin(half), 0.0, r_oh * math.cos(half)])
    xyz, Z = [], []
    for i in range(n_side):
        for j in range(n_side):
            for k in range(n_side):
                o = np.array([i, j, k], dtype=np.float64) * spacing
                R = _random_rotation(rng)
                xyz += [o, o + R @ h1, o + R @ h2]
                Z += [8, 1, 1]
    xyz = np.array(xyz)
    xyz -= xyz.mean(axis=0)
    bb = _BasisBuilder()
    o_shells = [0, 0, 0, 0, 1, 1, 1, 2, 2, 3]
    h_shells = [0, 0, 0, 1, 1, 2]
    for z in Z:
        ls = o_shells if z == 8 else h_shells
        bb.add_nucleus([(l, int(rng.integers(1, 10))) for l in ls], rng, (0.1, 1e4))
    arr = bb.arrays()
    A = int(sum(ncart(int(l)) for l in arr["shell_ang_mom"]))
    N = len(Z)
    elec_num = int(sum(Z))
    M = mo_num if mo_num is not None else elec_num // 2
    # AO -> nucleus map
    ao_nucl = np.empty(A, dtype=np.int64)
    k = 0
    for n in range(N):
        for s in range(arr["nucleus_index"][n], arr["nucleus_index"][n] + arr["nucleus_shell_num"][n]):
            nc = ncart(int(arr["shell_ang_mom"][s]))
            ao_nucl[k:k + nc] = n
            k += nc
    mo = rng.normal(0.0, 0.1, size=(M, A))
    keep_n = max(1, int(round((1.0 - zero_fraction) * N)))
    centres = rng.integers(0, N, size=M)
    for i in range(M):
        dist = np.linalg.norm(xyz - xyz[centres[i]], axis=1)
        near = np.argsort(dist, kind="stable")[:keep_n]
        mask = np.zeros(N, dtype=bool)
        mask[near] = True
        mo[i, ~mask[ao_nucl]] = 0.0
    owners = []
    for n, z in enumerate(Z):
        owners += [n] * z
    elec = _electrons_near(rng, xyz, owners, walk_num)
    s = System(name=name, charge=np.array(Z, dtype=np.float64),
               nucl_coord=np.ascontiguousarray(xyz.T), mo_coef=mo,
               elec_up_num=elec_num - elec_num // 2, elec_dn_num=elec_num // 2,
               elec_coord=elec, **arr,
               note=f"{N} nuclei, {A} Cartesian AOs, {M} MOs, "
                    f"{float((mo == 0).mean()):.3f} of MO coefficients zero")
    s.validate()
    return s


def water64(walk_num: int = 128) -> System:
    return water_cluster(4, walk_num=walk_num, seed=1003)


def water64_diff(walk_num: int = 32, seed: int = 1005):
    """The 64-water system at 32 walkers with weights w ~ N(0, 1) in the
    layout of mo_vgl [P][5][M]."""
    s = water_cluster(4, walk_num=walk_num, seed=1003, name="water64_diff")
    rng = np.random.default_rng(seed)
    w = rng.normal(0.0, 1.0, size=(s.point_num, 5, s.mo_num))
    return s, w


# cc-pVDZ shell shapes: (l, nprim) per element
_CCPVDZ = {
    1: [(0, 4), (0, 1), (1, 1)],
    6: [(0, 9), (0, 9), (0, 1), (1, 4), (1, 1), (2, 1)],
    7: [(0, 9), (0, 9), (0, 1), (1, 4), (1, 1), (2, 1)],
    8: [(0, 9), (0, 9), (0, 1), (1, 4), (1, 1), (2, 1)],
}


def ragged_batch(n_mol: int = 48, seed: int = 1004, max_ao: int = 1000,
                 walkers=(16, 512), scale: float = 1.0) -> List[System]:
    """48 molecules of 2-64 nuclei (Z in {1, 6, 7, 8}) on random 3D lattices,
    cc-pVDZ-shaped shells (AO count kept at or below ``max_ao``), electrons
    = sum of Z (2-256, kept by dropping nuclei), 16-512 walkers each,
    MOs = one per electron pair.  ``scale`` < 1 shrinks walker counts for
    CPU tests."""
    rng = np.random.default_rng(seed)
    out = []
    for m in range(n_mol):
        while True:
            nn = int(rng.integers(2, 65))
            Z = rng.choice([1, 6, 7, 8], size=nn, p=[0.45, 0.25, 0.15, 0.15]).tolist()
            A = sum(sum(ncart(l) for (l, _) in _CCPVDZ[z]) for z in Z)
            E = sum(Z)
            if 30 <= A <= max_ao and 2 <= E <= 256:
                break
        side = int(math.ceil(nn ** (1.0 / 3.0))) + 1
        cells = rng.choice(side ** 3, size=nn, replace=False)
        xyz = np.stack([cells // (side * side), (cells // side) % side, cells % side], axis=1)
        xyz = xyz.astype(np.float64) * 2.6 + rng.normal(0.0, 0.15, size=(nn, 3))
        bb = _BasisBuilder()
        for z in Z:
            bb.add_nucleus(list(_CCPVDZ[z]), rng, (0.1, 1e4) if z > 1 else (0.1, 100.0))
        arr = bb.arrays()
        E = int(sum(Z))
        M = max(1, (E + 1) // 2)
        A = int(sum(ncart(int(l)) for l in arr["shell_ang_mom"]))
        mo = rng.normal(0.0, 0.1, size=(M, A))
        W = int(rng.integers(walkers[0], walkers[1] + 1))
        W = max(1, int(round(W * scale)))
        owners = []
        for n, z in enumerate(Z):
            owners += [n] * z
        elec = _electrons_near(rng, xyz, owners, W)
        s = System(name=f"mol{m:02d}", charge=np.array(Z, dtype=np.float64),
                   nucl_coord=np.ascontiguousarray(xyz.T), mo_coef=mo,
                   elec_up_num=E - E // 2, elec_dn_num=E // 2, elec_coord=elec, **arr)
        s.validate()
        out.append(s)
    return out
