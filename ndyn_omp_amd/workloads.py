"""Synthetic MD inputs: FCC lattices, seeded velocities and the benchmark configs.

Everything here is deterministic from an integer seed (splitmix64 streams, Box-Muller
normals), so the same system can be rebuilt on any host.  Units follow Ndyn
(README.md:21-33 of the reference): eV, Angstrom, amu, time unit 10.1805057 fs,
k_B = 8.6173303e-5 eV/K.  Positions are scaled coordinates in [-0.5, 0.5) as the
reference stores them (src/read_input.f90:380).
"""
from __future__ import annotations

import dataclasses
import math
from typing import List, Optional, Sequence

import numpy as np

KB_EV = 8.6173303e-5          # eV/K, reference README.md:27
TIME_UNIT_FS = 10.1805057     # fs per internal time unit, reference README.md:25
DT_1FS = 0.0982269            # 1 fs in internal units (deck value)
DT_2FS = 0.196454             # 2 fs

# pair kinds (same ids as the C ABI / oracle)
MORSE, BUCK, LJ, MORSE_PHYS, SPRING = 1, 2, 3, 4, 5
SINE, COSINE, SOFT, JOHNSON, MORSE_C, BUCK_C, LJ_C = 6, 7, 8, 9, 10, 11, 12
# added to a kind: energy table V(r) - V(rc), same forces (not in the reference, which
# truncates every potential at its cutoff)
ESHIFT = 0x100

CU_MORSE = (7.0, 0.3429, 1.3588, 2.866)        # rc, D, alpha, r0
# peridynamic bond of examples/testPDspringTension/input.sh: rc, C, max strain s0, Vi
PD_SPRING = (6.0, 2.152295e-3, 0.1, 12.459008)
AR_LJ = (8.5125, 0.0104, 3.405, 0.0)           # rc, eps, sigma
ALLOY_BUCK = (8.0, 1000.0, 0.3, 10.0)          # rc, A, rho, C (code order, src/potentials.f90:115-117)


# --------------------------------------------------------------------------- RNG
_MASK = (1 << 64) - 1


def _splitmix64(x: np.ndarray) -> np.ndarray:
    """splitmix64 finaliser on a uint64 counter array (Steele, Lea, Flood 2014)."""
    z = (x + np.uint64(0x9E3779B97F4A7C15)) & np.uint64(_MASK)
    z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & np.uint64(_MASK)
    z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & np.uint64(_MASK)
    return z ^ (z >> np.uint64(31))


def uniform(seed: int, stream: int, n: int) -> np.ndarray:
    """n doubles in (0,1) from counter-based splitmix64 (seed, stream, index)."""
    with np.errstate(over="ignore"):
        base = np.uint64((seed * 0x632BE59BD9B4E019 + stream * 0x8CB92BA72F3D8DD7) & _MASK)
        ctr = base + np.arange(n, dtype=np.uint64)
        bits = _splitmix64(ctr) >> np.uint64(11)
    return (bits.astype(np.float64) + 0.5) * (1.0 / 9007199254740992.0)


def normal(seed: int, stream: int, n: int) -> np.ndarray:
    """n standard normals by Box-Muller on two uniform streams."""
    m = (n + 1) // 2
    u1 = uniform(seed, 2 * stream + 1, m)
    u2 = uniform(seed, 2 * stream + 2, m)
    r = np.sqrt(-2.0 * np.log(u1))
    z = np.concatenate([r * np.cos(2 * math.pi * u2), r * np.sin(2 * math.pi * u2)])
    return z[:n]


# --------------------------------------------------------------------------- systems
@dataclasses.dataclass
class System:
    """One periodic system in the reference's layout (pos(3,np) scaled, typ 1-based)."""
    pos: np.ndarray                 # (np,3) float64, scaled
    vel: np.ndarray                 # (np,3) float64, A / time unit
    typ: np.ndarray                 # (np,) int32, 1-based
    box: np.ndarray                 # (3,) float64, A
    masses: np.ndarray              # (ntypes,) amu
    pairmat: np.ndarray             # (ntypes,ntypes) int32, 1-based pair ids, 0 = none
    kinds: np.ndarray               # (npairs,) int32
    params: np.ndarray              # (npairs,4) float64 {rc, p2, p3, p4}
    ptrmin: float = 0.25
    ptsize: int = 5000
    name: str = ""
    # initial positions (scaled, (np,3)) and box of initPositionSet, for 'spring' pairs
    ipos: Optional[np.ndarray] = None
    init_box: Optional[np.ndarray] = None
    # charge product of each pair (PairChargeProd), for the Coulomb kinds
    qprd: Optional[np.ndarray] = None

    @property
    def np(self) -> int:
        return int(self.pos.shape[0])

    @property
    def ntypes(self) -> int:
        return int(self.masses.shape[0])

    @property
    def npairs(self) -> int:
        return int(self.params.shape[0])

    @property
    def springs(self) -> bool:
        return bool(np.any((np.asarray(self.kinds) & ~ESHIFT) == SPRING))

    def copy(self) -> "System":
        return dataclasses.replace(self, pos=self.pos.copy(), vel=self.vel.copy(),
                                   typ=self.typ.copy(), box=self.box.copy(),
                                   ipos=None if self.ipos is None else self.ipos.copy(),
                                   init_box=None if self.init_box is None else self.init_box.copy())


def fcc(nx: int, ny: int, nz: int, a: float, fp32_like_dump: bool = False):
    """FCC positions (scaled, [-0.5,0.5)) and box; atom order as the reference's
    examples/MorseTest/model.dump: basis (0,0,0),(1/2,1/2,0),(1/2,0,1/2),(0,1/2,1/2),
    x fastest.  fp32_like_dump rounds like the dump reader (src/read_input.f90:380)."""
    basis = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    iz, iy, ix = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    cells = np.stack([ix.ravel(), iy.ravel(), iz.ravel()], axis=1).astype(np.float64)
    pts = (cells[:, None, :] + basis[None, :, :]).reshape(-1, 3)
    n = np.array([nx, ny, nz], dtype=np.float64)
    if fp32_like_dump:
        xs = (pts / n).astype(np.float32)
        pos = (xs - np.float32(0.5)).astype(np.float64)
    else:
        pos = pts / n - 0.5
    box = n * a
    return pos, box


def thermal_velocities(typ: np.ndarray, masses: np.ndarray, temp_k: float, seed: int,
                       exact_ke: bool = True) -> np.ndarray:
    """Maxwell-Boltzmann velocities at temp_k with zero net momentum, then rescaled so
    that KE = 3/2 N k T exactly (what the reference's Velocity directive intends,
    src/nd.f90:1598-1622)."""
    npart = typ.shape[0]
    m = masses[typ - 1]
    z = normal(seed, 7, 3 * npart).reshape(npart, 3)
    v = z * np.sqrt(KB_EV * temp_k / m)[:, None]
    p = (v * m[:, None]).sum(axis=0) / m.sum()
    v -= p[None, :]
    if exact_ke and temp_k > 0:
        ke = 0.5 * (m[:, None] * v * v).sum()
        v *= math.sqrt(1.5 * npart * KB_EV * temp_k / ke)
    return v


def jitter(pos: np.ndarray, box: np.ndarray, sigma_a: float, seed: int) -> np.ndarray:
    """Gaussian displacement of sigma_a Angstrom, kept in scaled units."""
    z = normal(seed, 3, pos.size).reshape(pos.shape)
    return pos + z * sigma_a / box[None, :]


def cu_morse(nx: int, a: float = 3.615, temp_k: float = 300.0, seed: int = 12345,
             jitter_a: float = 0.0, kind: int = MORSE, params=CU_MORSE,
             ptsize: int = 5000, ptrmin: float = 0.25, fp32_like_dump=False) -> System:
    pos, box = fcc(nx, nx, nx, a, fp32_like_dump)
    if jitter_a > 0:
        pos = jitter(pos, box, jitter_a, seed)
    typ = np.ones(pos.shape[0], dtype=np.int32)
    masses = np.array([63.546])
    vel = thermal_velocities(typ, masses, temp_k, seed) if temp_k > 0 else np.zeros_like(pos)
    return System(pos, vel, typ, box, masses, np.array([[1]], dtype=np.int32),
                  np.array([kind], dtype=np.int32), np.array([params], dtype=np.float64),
                  ptrmin, ptsize, f"cu_morse_{nx}")


def ar_lj(nx: int, a: float = 5.26, temp_k: float = 90.0, seed: int = 12346,
          jitter_a: float = 0.005, ptsize: int = 5000) -> System:
    pos, box = fcc(nx, nx, nx, a)
    pos = jitter(pos, box, jitter_a, seed)
    typ = np.ones(pos.shape[0], dtype=np.int32)
    masses = np.array([39.948])
    vel = thermal_velocities(typ, masses, temp_k, seed)
    return System(pos, vel, typ, box, masses, np.array([[1]], dtype=np.int32),
                  np.array([LJ], dtype=np.int32), np.array([AR_LJ], dtype=np.float64),
                  0.25, ptsize, f"ar_lj_{nx}")


def buck_alloy(n: Sequence[int], a: float = 5.6, temp_k: float = 600.0, seed: int = 12347,
               jitter_a: float = 0.0, ptsize: int = 5000) -> System:
    """Two-species Buckingham alloy (A-A, A-B, B-B all with the same parameters),
    50/50 random species, masses 27 and 63.5 amu."""
    nx, ny, nz = n
    pos, box = fcc(nx, ny, nz, a)
    if jitter_a > 0:
        pos = jitter(pos, box, jitter_a, seed)
    u = uniform(seed, 11, pos.shape[0])
    typ = np.where(u < 0.5, 1, 2).astype(np.int32)
    masses = np.array([27.0, 63.5])
    vel = thermal_velocities(typ, masses, temp_k, seed)
    pairmat = np.array([[1, 2], [2, 3]], dtype=np.int32)
    params = np.array([ALLOY_BUCK] * 3, dtype=np.float64)
    return System(pos, vel, typ, box, masses, pairmat, np.array([BUCK] * 3, dtype=np.int32),
                  params, 0.25, ptsize, f"buck_alloy_{nx}x{ny}x{nz}")


def asym_morse(nx: int, variant: int = 1, a: float = 3.615, temp_k: float = 300.0,
               seed: int = 12350, jitter_a: float = 0.0, ptsize: int = 5000) -> System:
    """Two species (50/50, masses 63.5 and 27 amu) on an FCC lattice with directed Morse
    pairs, as written by the reference's `asym` pair lines (src/read_input.f90:127-133:
    only PairMat(A,B) is set).  An atom of type t feels a neighbour of type u when
    PairMat(u,t) > 0 (src/nd.f90:474-483):
      variant 1: PairMat = [[1, 2], [0, 3]] -- A-A; PairMat(A,B) = 2, PairMat(B,A) = 0, so
                 B atoms feel A atoms and A atoms do not feel B atoms; B-B;
      variant 2: PairMat = [[1, 2], [3, 4]] -- A-A; both directions set, with different
                 ids (the pair uses table 3 both ways); B-B.
    Each pair has its own Morse parameters, so the table choice (the larger id of the two
    directions) is visible in the forces."""
    pos, box = fcc(nx, nx, nx, a)
    if jitter_a > 0:
        pos = jitter(pos, box, jitter_a, seed)
    u = uniform(seed, 13, pos.shape[0])
    typ = np.where(u < 0.5, 1, 2).astype(np.int32)
    masses = np.array([63.5, 27.0])
    vel = thermal_velocities(typ, masses, temp_k, seed)
    prm = [CU_MORSE, (6.8, 0.25, 1.30, 2.90), (7.0, 0.40, 1.40, 2.85), (6.9, 0.30, 1.35, 2.88)]
    if variant == 1:
        pairmat = np.array([[1, 2], [0, 3]], dtype=np.int32)
        prm = [prm[0], prm[1], prm[3]]
    else:
        pairmat = np.array([[1, 2], [3, 4]], dtype=np.int32)
    n = len(prm)
    return System(pos, vel, typ, box, masses, pairmat, np.array([MORSE] * n, dtype=np.int32),
                  np.array(prm, dtype=np.float64), 0.25, ptsize, f"asym_morse{variant}_{nx}")


def spring_cu(nx: int, a: float = 3.68, stretch=(0.0, 0.0, 0.0), jitter_a: float = 0.0,
              temp_k: float = 0.0, seed: int = 12360, ptsize: int = 1000) -> System:
    """A peridynamic FCC block: one 'spring' pair (PD_SPRING, as in
    examples/testPDspringTension/input.sh), initial positions = the perfect lattice in its
    own box, current state = the lattice stretched by `stretch` (relative, per axis) and
    jittered by jitter_a (A).  Bonds whose strain exceeds PairParams(3) carry no force."""
    ipos, box0 = fcc(nx, nx, nx, a)
    box = box0 * (1.0 + np.asarray(stretch, dtype=np.float64))
    pos = jitter(ipos, box, jitter_a, seed) if jitter_a > 0 else ipos.copy()
    typ = np.ones(pos.shape[0], dtype=np.int32)
    masses = np.array([63.546])
    vel = thermal_velocities(typ, masses, temp_k, seed)
    return System(pos, vel, typ, box, masses, np.array([[1]], dtype=np.int32),
                  np.array([SPRING], dtype=np.int32), np.array([PD_SPRING], dtype=np.float64),
                  0.25, ptsize, f"spring_cu_{nx}", ipos=ipos, init_box=box0.copy())


def spring_mix(nx: int, variant: int = 0, a: float = 3.615, stretch=(0.02, 0.0, -0.01),
               jitter_a: float = 0.1, temp_k: float = 0.0, seed: int = 12361,
               ptsize: int = 2000) -> System:
    """Two species (50/50, 63.5 / 27 amu) on an FCC lattice, Morse within each species and
    springs between them, against the unstretched lattice as initial positions:
      variant 0: PairMat = [[1, 2], [2, 3]] -- symmetric, A-B is a spring (compute2);
      variant 1: PairMat = [[1, 2], [0, 3]] -- A-B spring one way only: B atoms feel A atoms
                 through it, A atoms do not feel B (compute3, as the GP deck's
                 'Cu1 Cu-2 asym spring', examples/testGP/GP/input.sh:27);
      variant 2: PairMat = [[1, 2], [3, 4]] -- springs both ways with different ids, the
                 pair uses spring 3 (the larger id) both ways (compute3)."""
    ipos, box0 = fcc(nx, nx, nx, a)
    box = box0 * (1.0 + np.asarray(stretch, dtype=np.float64))
    pos = jitter(ipos, box, jitter_a, seed) if jitter_a > 0 else ipos.copy()
    u = uniform(seed, 13, pos.shape[0])
    typ = np.where(u < 0.5, 1, 2).astype(np.int32)
    masses = np.array([63.5, 27.0])
    vel = thermal_velocities(typ, masses, temp_k, seed)
    mA, mB = CU_MORSE, (6.8, 0.25, 1.30, 2.90)
    sp1, sp2 = (6.0, 2.152295e-3, 0.06, 12.459008), (6.5, 3.0e-3, 0.08, 10.0)
    if variant == 0:
        pm = [[1, 2], [2, 3]]
        kinds, prm = [MORSE, SPRING, MORSE], [mA, sp1, mB]
    elif variant == 1:
        pm = [[1, 2], [0, 3]]
        kinds, prm = [MORSE, SPRING, MORSE], [mA, sp1, mB]
    else:
        pm = [[1, 2], [3, 4]]
        kinds, prm = [MORSE, SPRING, SPRING, MORSE], [mA, sp1, sp2, mB]
    return System(pos, vel, typ, box, masses, np.array(pm, dtype=np.int32),
                  np.array(kinds, dtype=np.int32), np.array(prm, dtype=np.float64), 0.25,
                  ptsize, f"spring_mix{variant}_{nx}", ipos=ipos, init_box=box0.copy())


def near_cubic_cells(n_atoms: float) -> tuple:
    """Near-cubic (nx,ny,nz) with 4*nx*ny*nz closest to n_atoms."""
    target = max(1.0, n_atoms / 4.0)
    best = None
    c = round(target ** (1.0 / 3.0))
    for nx in range(max(1, c - 2), c + 3):
        for ny in range(max(1, nx - 1), nx + 2):
            for nz in range(max(1, ny - 1), ny + 2):
                cells = nx * ny * nz
                key = (abs(cells - target), max(nx, ny, nz) - min(nx, ny, nz))
                if best is None or key < best[0]:
                    best = (key, (nx, ny, nz))
    return best[1]


def replica_mix(n_replicas: int = 512, lo: float = 500, hi: float = 20000,
                seed: int = 12348, cu_a: float = 3.615, alloy_a: float = 5.6,
                t_lo: float = 100.0, t_hi: float = 1200.0) -> List[System]:
    """Uneven replica mix: atom counts log-uniform in [lo,hi] rounded to whole near-cubic
    4-atom FCC cells, each replica Morse Cu or Buckingham alloy, T uniform in [t_lo,t_hi]."""
    u_size = uniform(seed, 21, n_replicas)
    u_kind = uniform(seed, 22, n_replicas)
    u_temp = uniform(seed, 23, n_replicas)
    out = []
    for r in range(n_replicas):
        n = math.exp(math.log(lo) + u_size[r] * (math.log(hi) - math.log(lo)))
        cells = near_cubic_cells(n)
        t = t_lo + u_temp[r] * (t_hi - t_lo)
        if u_kind[r] < 0.5:
            pos, box = fcc(*cells, cu_a)
            typ = np.ones(pos.shape[0], dtype=np.int32)
            masses = np.array([63.546])
            vel = thermal_velocities(typ, masses, t, seed + 1000 + r)
            s = System(pos, vel, typ, box, masses, np.array([[1]], dtype=np.int32),
                       np.array([MORSE], dtype=np.int32),
                       np.array([CU_MORSE], dtype=np.float64), 0.25, 5000,
                       f"rep{r}_cu_{cells}")
        else:
            s = buck_alloy(cells, a=alloy_a, temp_k=t, seed=seed + 1000 + r)
            s.name = f"rep{r}_alloy_{cells}"
        out.append(s)
    return out


def serving_jobs(n_jobs: int, seed: int = 12349, lo: float = 1000, hi: float = 200000,
                 steps_lo: float = 1000, steps_hi: float = 50000, rate_hz: float = 1.0,
                 step_scale: float = 1.0, size_scale: float = 1.0):
    """Serving-mode job stream: (arrival_s, System, nsteps, ensemble, temp_k).
    step_scale / size_scale shrink the stream for bounded runs."""
    u_size = uniform(seed, 31, n_jobs)
    u_steps = uniform(seed, 32, n_jobs)
    u_kind = uniform(seed, 33, n_jobs)
    u_ens = uniform(seed, 34, n_jobs)
    u_temp = uniform(seed, 35, n_jobs)
    u_arr = uniform(seed, 36, n_jobs)
    t = 0.0
    jobs = []
    for k in range(n_jobs):
        t += -math.log(u_arr[k]) / rate_hz
        n = math.exp(math.log(lo) + u_size[k] * (math.log(hi) - math.log(lo))) * size_scale
        steps = math.exp(math.log(steps_lo) + u_steps[k] * (math.log(steps_hi) - math.log(steps_lo)))
        steps = max(1, int(round(steps * step_scale)))
        temp = 100.0 + u_temp[k] * 1100.0
        cells = near_cubic_cells(n)
        if u_kind[k] < 0.5:
            pos, box = fcc(*cells, 3.615)
            typ = np.ones(pos.shape[0], dtype=np.int32)
            masses = np.array([63.546])
            vel = thermal_velocities(typ, masses, temp, seed + 5000 + k)
            s = System(pos, vel, typ, box, masses, np.array([[1]], dtype=np.int32),
                       np.array([MORSE], dtype=np.int32),
                       np.array([CU_MORSE], dtype=np.float64), 0.25, 5000, f"job{k}_cu")
        else:
            s = buck_alloy(cells, temp_k=temp, seed=seed + 5000 + k)
            s.name = f"job{k}_alloy"
        ens = "nvt" if u_ens[k] < 0.5 else "nve"
        jobs.append((t, s, steps, ens, temp))
    return jobs
