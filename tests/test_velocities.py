"""Initial velocities (normalVelocityDistribution, src/nd.f90:1500-1624).

The reference draws v = N(0,1) sqrt(2 KE_t / (3 m)) per component (src/nd.f90:1598-1601,
KE_t = 3/2 k T per atom) and subtracts the mean velocity (src/nd.f90:1608-1622).  Its random
stream depends on the compiler and the thread count, so it cannot be reproduced; what is
pinned here are the properties the stream must have:

CPU -- the seeded generator of the workloads (thermal_velocities, used by every test,
benchmark and job): total momentum zero to 1e-12 relative, kinetic energy 3/2 N k T to 1e-14
relative, per type the components of v sqrt(m / kT) standard normal (mean, variance and
kurtosis inside 5 sigma of their sampling distributions, Kolmogorov-Smirnov p > 1e-4); the
Box-Muller source itself; different seeds give independent draws.
GPU -- the deck driver's Velocity directive (its own Box-Muller, then the reference's
arithmetic-mean removal): from its xyz frame (4 significant digits) the velocity sum is zero
to 1e-5 of sum |v| and the scaled components are standard normal as above; the KE of its
first thermo line lies within 5 sigma of 3/2 (N - 1) k T.
"""
import math
import os

import numpy as np
import pytest
from scipy import stats

from ndyn_omp_amd import workloads as W


def _normal_checks(u, label):
    n = u.size
    assert abs(u.mean()) < 5.0 / math.sqrt(n), (label, u.mean())
    assert abs(u.var() - 1.0) < 5.0 * math.sqrt(2.0 / n), (label, u.var())
    assert abs(stats.kurtosis(u)) < 5.0 * math.sqrt(24.0 / n), (label, stats.kurtosis(u))
    assert stats.kstest(u, "norm").pvalue > 1e-4, label


def test_box_muller_source_is_standard_normal():
    z = W.normal(2024, 5, 200001)
    assert z.shape == (200001,)
    _normal_checks(z, "normal")
    # streams and seeds are independent of each other
    a, b = W.normal(2024, 5, 50000), W.normal(2024, 6, 50000)
    c = W.normal(2025, 5, 50000)
    for x, y in ((a, b), (a, c)):
        assert abs(np.corrcoef(x, y)[0, 1]) < 5.0 / math.sqrt(50000)
    np.testing.assert_array_equal(W.normal(2024, 5, 1000), W.normal(2024, 5, 1000))


@pytest.mark.parametrize("case", ["cu", "alloy", "cu_cold"])
def test_thermal_velocities(case):
    if case == "alloy":
        s = W.buck_alloy((12, 12, 12), temp_k=900.0, seed=77)   # masses 27 and 63.5
        temp = 900.0
    elif case == "cu":
        s = W.cu_morse(14, temp_k=300.0, seed=78)
        temp = 300.0
    else:
        s = W.cu_morse(14, temp_k=10.0, seed=79)
        temp = 10.0
    m = s.masses[s.typ - 1]
    v = s.vel
    p = (m[:, None] * v).sum(axis=0)
    pscale = (m[:, None] * np.abs(v)).sum(axis=0)
    assert np.all(np.abs(p) < 1e-12 * pscale), p / pscale
    ke = 0.5 * (m[:, None] * v * v).sum()
    assert ke == pytest.approx(1.5 * s.np * W.KB_EV * temp, rel=1e-14)
    for t in range(1, s.ntypes + 1):
        sel = s.typ == t
        u = (v[sel] * np.sqrt(s.masses[t - 1] / (W.KB_EV * temp))).ravel()
        _normal_checks(u, f"{case} type {t}")
    # the same seed gives the same velocities; another seed others
    s2 = W.cu_morse(14, temp_k=300.0, seed=78) if case == "cu" else None
    if s2 is not None:
        np.testing.assert_array_equal(s2.vel, s.vel)
        s3 = W.cu_morse(14, temp_k=300.0, seed=80)
        assert abs(np.corrcoef(s3.vel.ravel(), s.vel.ravel())[0, 1]) < 5.0 / math.sqrt(s.vel.size)


DRIVER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                      "ndyn_omp_amd", "bin", "ndyn_gpu")


def _xyz_last_frame(path):
    lines = open(path).read().splitlines()
    starts = [k for k, l in enumerate(lines) if l.strip().isdigit()]
    k = starts[-1]
    n = int(lines[k])
    rows = [l.split() for l in lines[k + 2:k + 2 + n]]
    return np.array([[float(x) for x in r[4:7]] for r in rows])


@pytest.mark.gpu
@pytest.mark.skipif(not os.path.exists(DRIVER), reason="driver not built")
def test_driver_velocity_directive(tmp_path):
    """Deck: Cu FCC 10^3 (4000 atoms), Velocity 0.038778 (KE per atom of 300 K), one step of
    dt 0 (velocities unchanged, forces vanish on the perfect lattice), an xyz frame."""
    from test_driver import run_deck, thermo, write_fcc_dump
    write_fcc_dump(tmp_path / "model.dump", n=10, a=3.615)
    kt = 1.5 * W.KB_EV * 300.0
    deck = f"""START
Model 	model.dump dump
Types 1
 Cu 63.546
Pairs 1 0.25 1000
 Cu Cu sym morse 7.0 0.3429 1.3588 2.866
Seed 4242
Velocity {kt:.6f}
OutFreq 1
ConfOutFreq 0 xyz
Run 1 0.0
END
"""
    print("case driver Velocity", flush=True)
    p = run_deck(deck, tmp_path)
    assert p.returncode == 0, p.stderr
    v = _xyz_last_frame(tmp_path / "out.xyz")
    assert v.shape == (4000, 3)
    assert np.all(np.abs(v.sum(axis=0)) < 1e-5 * np.abs(v).sum(axis=0))
    kT = W.KB_EV * 300.0
    _normal_checks((v * math.sqrt(63.546 / kT)).ravel(), "driver Velocity")
    rows = thermo(p.stdout)
    ke = rows[0][2]
    n = 4000
    mean = 1.5 * (n - 1) * kT
    assert abs(ke - mean) < 5.0 * math.sqrt(1.5 * (n - 1)) * kT, (ke, mean)
