"""Pins the CPU oracle (oracle/ndyn_oracle.c) before it is trusted as the checker.

Checks, per oracle function:
  * golden vectors written by the compiled reference (tests/golden/make_golden.py);
  * known answers printed in the reference's own example logs;
  * closed forms that do not go through either code: FCC shell sums, Newton 3,
    free-particle drift, exact velocity rescale, a one-step barostat by hand,
    brute-force pair sets.
"""
import math

import numpy as np
import pytest

from conftest import golden, system_from_golden
from ndyn_omp_amd import workloads as W


# ----------------------------------------------------------------- tables
@pytest.mark.parametrize("name,kind", [("morse", 1), ("lj", 3), ("buck", 2)])
def test_tables_match_reference(oracle, name, kind):
    g = golden("tables.npz")
    P, F, rc2, dr2 = oracle.pair_tables([kind], g[f"{name}_params"], float(g["ptrmin"]),
                                        int(g["ptsize"]))
    assert rc2[0] == g[f"{name}_rc2"] and dr2[0] == g[f"{name}_dr2"]
    # absolute floor for the rows where the two LJ/Buck terms cancel (|F| ~ 1e-3)
    np.testing.assert_allclose(P[0], g[f"{name}_P"], rtol=1e-13, atol=1e-14)
    np.testing.assert_allclose(F[0], g[f"{name}_F"], rtol=1e-13, atol=1e-14)


MORE = ["sine", "cosine", "soft", "john", "morsec", "buckc", "ljc"]


@pytest.mark.parametrize("name", MORE)
def test_more_tables_match_reference(oracle, name):
    """The other potentials of buildPairTables (src/potentials.f90:82-147, 172-241) against
    the compiled reference's tables: sine, cosine, soft, John and the Coulomb variants
    MorseC / BuckC / LJC with a nonzero charge product."""
    g = golden("tables_more.npz")
    P, F, rc2, dr2 = oracle.pair_tables([int(g[f"{name}_kind"])], g[f"{name}_params"],
                                        float(g["ptrmin"]), int(g["ptsize"]),
                                        qprd=[float(g[f"{name}_qprd"])])
    assert rc2[0] == g[f"{name}_rc2"] and dr2[0] == g[f"{name}_dr2"]
    for got, want in ((P[0], g[f"{name}_P"]), (F[0], g[f"{name}_F"])):
        np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-13 * np.abs(want).max())


def test_more_tables_closed_form(oracle):
    """Independent forms (numpy, scipy's erfc): the Coulomb variants add
    q (erfc(a r)/r - vcon + fcon (r - rc)) to the energy and r q (erfc(a r)/r^2 +
    2a/sqrt(pi) exp(-a^2 r^2)/r - fcon) to the force column, a = 0.2
    (src/potentials.f90:11,23-26,106-108,199-201); soft's force column is
    r pi p1 sin(pi r/p2)/p2 while its energy is p2 (1 + cos(pi r/p3)) (:86, :176)."""
    from scipy.special import erfc
    prm, q = W.CU_MORSE, 0.8
    P0, F0, _, dr2 = oracle.pair_tables([W.MORSE], prm, 0.25, 5000)
    P1, F1, _, _ = oracle.pair_tables([W.MORSE_C], prm, 0.25, 5000, qprd=[q])
    r = np.sqrt(0.25 + np.arange(5001) * dr2[0])
    a, rc = 0.2, prm[0]
    vcon = erfc(a * rc) / rc
    fcon = erfc(a * rc) / rc ** 2 + 2 * a / np.sqrt(np.pi) * np.exp(-a * a * rc * rc) / rc
    np.testing.assert_allclose(P1[0] - P0[0], q * (erfc(a * r) / r - vcon + fcon * (r - rc)),
                               rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(F1[0] - F0[0], r * q * (erfc(a * r) / r ** 2 + 2 * a / np.sqrt(np.pi)
                                                       * np.exp(-a * a * r * r) / r - fcon),
                               rtol=1e-9, atol=1e-12)
    sp = (6.0, 0.5, 6.0, 0.0)
    P, F, _, dr2 = oracle.pair_tables([W.SOFT], sp, 0.25, 5000)
    r = np.sqrt(0.25 + np.arange(5001) * dr2[0])
    np.testing.assert_allclose(P[0], 0.5 * (1 + np.cos(np.pi * r / 6.0)), rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(F[0], r * np.pi * 6.0 * np.sin(np.pi * r / 0.5) / 0.5,
                               rtol=1e-9, atol=1e-9)


def test_tables_closed_form(oracle):
    """Grid is uniform in r^2 and values are the analytic forms; Morse F is -V'(r)
    (reference quirk), LJ/Buck F are -r V'(r)."""
    prm = W.CU_MORSE
    P, F, rc2, dr2 = oracle.pair_tables([1], prm, 0.25, 5000)
    _, D, al, r0 = prm
    s = 0.25 + np.arange(5001) * dr2[0]
    r = np.sqrt(s)
    e = np.exp(-al * (r - r0))
    np.testing.assert_allclose(P[0], D * (e * e - 2 * e), rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(F[0], 2 * D * al * (e * e - e), rtol=1e-12, atol=1e-15)
    _, _, _, _ = oracle.pair_tables([4], prm, 0.25, 5000)
    P4, F4, _, _ = oracle.pair_tables([4], prm, 0.25, 5000)
    np.testing.assert_allclose(F4[0], r * 2 * D * al * (e * e - e), rtol=1e-12, atol=1e-15)
    _, eps, sig, _ = W.AR_LJ
    P, F, rc2, dr2 = oracle.pair_tables([3], W.AR_LJ, 0.25, 5000)
    s = 0.25 + np.arange(5001) * dr2[0]
    r = np.sqrt(s)
    x = (sig / r) ** 6
    np.testing.assert_allclose(P[0], 4 * eps * (x * x - x), rtol=1e-11)
    np.testing.assert_allclose(F[0], 24 * eps * (2 * x * x - x), rtol=1e-11)
    assert dr2[0] == (8.5125 ** 2 - 0.25) / 4999


# ----------------------------------------------------------------- lists
@pytest.mark.parametrize("fname", ["list_cu7.npz", "list_cu10.npz", "list_alloy5.npz",
                                   "list_ar6.npz", "list_asym1.npz"])
def test_half_list_matches_reference_exactly(oracle, fname):
    g = golden(fname)
    s = system_from_golden(g)
    neimax = int(g["neimax"])
    err, vnl, _ = oracle.half_list(s, float(g["skin"]), neimax=neimax)
    assert err == 0
    np.testing.assert_array_equal(vnl[:, 0], g["counts"])
    ent = np.concatenate([vnl[i, 1:1 + vnl[i, 0]] for i in range(s.np)])
    np.testing.assert_array_equal(ent, g["entries"])        # same entries, same order


@pytest.mark.parametrize("fname", ["list_cu7.npz", "list_alloy5.npz"])
def test_half_list_equals_brute_force_pair_set(oracle, fname):
    g = golden(fname)
    s = system_from_golden(g)
    skin = float(g["skin"])
    _, vnl, neimax = oracle.half_list(s, skin)
    _, vnlb, _ = oracle.half_list(s, skin, neimax=neimax, brute=True)
    assert oracle.pair_set_from_half(vnl) == oracle.pair_set_from_half(vnlb)
    # and both equal the set defined directly from the minimum-image distance
    rc = s.params[:, 0].max() + skin
    d = s.pos[:, None, :] - s.pos[None, :, :]
    d -= np.round(d)
    d2 = ((d * s.box) ** 2).sum(-1)
    i, j = np.nonzero(np.triu(d2 <= rc * rc, 1))
    assert set(zip(i.tolist(), j.tolist())) == oracle.pair_set_from_half(vnl)


def test_neighbour_counts_match_reference_logs(oracle):
    """examples/MorseTest/out:23-25 (skin 1.2) and examples/testT/out:23-25 (skin 0.2):
    MCELL 5, NeiMax estimate 106 for skin 0.2, count+1 min/max 15/135 and 12/79."""
    s = W.cu_morse(10, a=3.68, temp_k=0.0, params=(6.0, 0.3429, 1.359, 2.866), ptsize=1000,
                   fp32_like_dump=True)
    assert list(oracle.cell_grid(s.box, 6.0, 1.2)) == [5, 5, 5]
    assert oracle.neimax_estimate(6.0, 0.2, s.np, s.box) == 106
    _, vnl, _ = oracle.half_list(s, 1.2, neimax=200)
    assert (vnl[:, 0].min() + 1, vnl[:, 0].max() + 1) == (15, 135)
    _, vnl, _ = oracle.half_list(s, 0.2, neimax=106)
    assert (vnl[:, 0].min() + 1, vnl[:, 0].max() + 1) == (12, 79)


def test_neimax_overflow_is_reported(oracle):
    s = W.cu_morse(7, temp_k=0.0)
    err, _, _ = oracle.half_list(s, 0.5, neimax=20)
    assert err == 1


# ----------------------------------------------------------------- forces
@pytest.mark.parametrize("fname", ["force_cu10.npz", "force_ar6.npz", "force_alloy5.npz"])
def test_compute0_matches_reference(oracle, fname):
    g = golden(fname)
    s = system_from_golden(g)
    skin = float(g["skin"])
    tables = oracle.pair_tables(s.kinds, s.params, s.ptrmin, s.ptsize)
    _, vnl, neimax = oracle.half_list(s, skin)
    f, pe, vir, pene = oracle.compute0(s, vnl, neimax, tables)
    scale = np.abs(g["force"]).max()
    np.testing.assert_allclose(f, g["force"], rtol=1e-11, atol=1e-13 * scale)
    assert pe == pytest.approx(float(g["pe"]), rel=1e-13)
    np.testing.assert_allclose(vir, g["virial"], rtol=1e-12, atol=1e-12 * np.abs(g["virial"]).max())
    # invariants: Newton 3 and sum of per-atom energies
    assert np.abs(f.sum(axis=0)).max() < 1e-11 * max(1.0, scale) * math.sqrt(s.np)
    assert pene.sum() == pytest.approx(pe, rel=1e-12)


@pytest.mark.parametrize("fname", ["force_asym1.npz", "force_asym2.npz"])
def test_compute1_matches_reference(oracle, fname):
    """Asymmetric pair matrices: the reference's compute1 (src/nd.f90:364-530).  compute0
    would give other forces (both directions feel every pair); the invariants that hold
    are the per-atom energy sum and, with both directions set (asym2), Newton 3."""
    g = golden(fname)
    s = system_from_golden(g)
    assert oracle.asymmetric(s)
    skin = float(g["skin"])
    tables = oracle.pair_tables(s.kinds, s.params, s.ptrmin, s.ptsize)
    _, vnl, neimax = oracle.half_list(s, skin)
    f, pe, vir, pene = oracle.compute(s, vnl, neimax, tables)
    scale = np.abs(g["force"]).max()
    np.testing.assert_allclose(f, g["force"], rtol=1e-11, atol=1e-13 * scale)
    assert pe == pytest.approx(float(g["pe"]), rel=1e-13)
    np.testing.assert_allclose(vir, g["virial"], rtol=1e-12, atol=1e-12 * np.abs(g["virial"]).max())
    assert pene.sum() == pytest.approx(pe, rel=1e-12)
    f0, _, _, _ = oracle.compute0(s, vnl, neimax, tables)
    assert np.abs(f0 - g["force"]).max() > 1e-3 * scale    # the fixture tells the two apart
    if (s.pairmat > 0).all():
        assert np.abs(f.sum(axis=0)).max() < 1e-11 * max(1.0, scale) * math.sqrt(s.np)
    else:
        # one-way pairs: the net force is the sum of the unanswered pair forces
        assert np.abs(f.sum(axis=0)).max() > 1e-6


SPRING_FORCES = ["force_spring_cu7.npz", "force_spring_mix0.npz", "force_spring_mix1.npz",
                 "force_spring_mix2.npz"]


def _spring_strains(s, vnl):
    """Strain r/r0 - 1 of every listed spring pair, from the definitions (numpy)."""
    stride = vnl.shape[1]
    ii = np.repeat(np.arange(s.np), vnl[:, 0])
    jj = np.concatenate([vnl[i, 1:1 + vnl[i, 0]] for i in range(s.np)])
    ti, tj = s.typ[ii] - 1, s.typ[jj] - 1
    pm = np.maximum(s.pairmat[ti, tj], s.pairmat[tj, ti])
    sp = s.kinds[pm - 1] == W.SPRING
    def dist(p, box):
        d = p[ii] - p[jj]
        d -= np.round(d)
        return np.sqrt(((d * box) ** 2).sum(axis=1))
    st = dist(s.pos, s.box) / dist(s.ipos, s.init_box) - 1.0
    return st[sp], s.params[pm[sp] - 1, 2]


@pytest.mark.parametrize("fname", SPRING_FORCES)
def test_compute23_matches_reference(oracle, fname):
    """Spring pairs: the reference's compute2 (symmetric PairMat, src/nd.f90:531-727) and
    compute3 (asymmetric, src/nd.f90:728-933) against initial positions, bit for bit.
    Every fixture has listed springs on both sides of the strain cut PairParams(3)."""
    g = golden(fname)
    s = system_from_golden(g)
    assert s.springs and s.ipos is not None
    skin = float(g["skin"])
    tables = oracle.pair_tables(s.kinds, s.params, s.ptrmin, s.ptsize)
    _, vnl, neimax = oracle.half_list(s, skin)
    f, pe, vir, pene = oracle.compute(s, vnl, neimax, tables, ipos=s.ipos, init_box=s.init_box)
    scale = np.abs(g["force"]).max()
    np.testing.assert_allclose(f, g["force"], rtol=1e-11, atol=1e-13 * scale)
    assert pe == pytest.approx(float(g["pe"]), rel=1e-13)
    np.testing.assert_allclose(vir, g["virial"], rtol=1e-12, atol=1e-12 * np.abs(g["virial"]).max())
    assert pene.sum() == pytest.approx(pe, rel=1e-12)
    strain, s0 = _spring_strains(s, vnl)
    assert (strain > s0).sum() > 10 and (strain <= s0).sum() > 100
    if not oracle.asymmetric(s):
        assert np.abs(f.sum(axis=0)).max() < 1e-11 * max(1.0, scale) * math.sqrt(s.np)


def test_spring_energy_and_virial_closed_form(oracle):
    """A perfect lattice stretched affinely (no jitter): every pair within the list cutoff
    (rc + skin = 6 A, current distances) carries e = C Vi^2 s^2 r0 / 2 and the virial
    r (x) r f/r with f = -C s Vi^2, s = r/r0 - 1, unless s > s0.  Brute force over all
    minimum-image pairs in numpy."""
    s = W.spring_cu(5, stretch=(0.07, -0.02, 0.11))
    rc, C, s0, Vi = W.PD_SPRING
    tables = oracle.pair_tables(s.kinds, s.params, s.ptrmin, s.ptsize)
    _, vnl, neimax = oracle.half_list(s, 0.0)
    _, pe, vir, _ = oracle.compute(s, vnl, neimax, tables, ipos=s.ipos, init_box=s.init_box)
    i, j = np.triu_indices(s.np, 1)
    d = s.pos[i] - s.pos[j]
    d -= np.round(d)
    rv = d * s.box
    r = np.sqrt((rv ** 2).sum(axis=1))
    d0 = s.ipos[i] - s.ipos[j]
    d0 -= np.round(d0)
    r0 = np.sqrt(((d0 * s.init_box) ** 2).sum(axis=1))
    keep = r <= rc
    assert np.abs(r[keep] - rc).min() > 1e-6 and np.abs(r[~keep] - rc).min() > 1e-6
    st = r / r0 - 1.0
    on = keep & (st <= s0)
    assert (keep & (st > s0)).sum() > 0     # some listed bonds are past the strain cut
    e_ref = (0.5 * C * Vi * Vi * st[on] ** 2 * r0[on]).sum()
    fr = -C * st[on] * Vi * Vi / r[on]
    v = rv[on]
    vir_ref = np.array([(v[:, 0] ** 2 * fr).sum(), (v[:, 1] ** 2 * fr).sum(),
                        (v[:, 2] ** 2 * fr).sum(), (v[:, 0] * v[:, 1] * fr).sum(),
                        (v[:, 0] * v[:, 2] * fr).sum(), (v[:, 1] * v[:, 2] * fr).sum()])
    assert pe == pytest.approx(e_ref, rel=1e-12)
    np.testing.assert_allclose(vir, vir_ref, rtol=1e-11, atol=1e-11 * np.abs(vir_ref).max())


def test_spring_tension_log(oracle):
    """examples/testPDspringTension/input.sh (4000-atom lattice at rest, one spring pair,
    Skin 0, ReNeighbor F, BoxVel 1e-2 0 0, Run 40000 0.1; initial positions set by dynamics)
    for 300 steps against examples/testPDspringTension/out:39-41: PE 0.191554, pressure
    -0.145643E-02, V 49970.1 at step 100; 0.774476, -0.292621E-02, 50105.5 at 200;
    1.74955, -0.439464E-02, 50240.9 at 300 (6 significant digits)."""
    s = W.spring_cu(10, a=3.68)
    s.ipos = s.init_box = None
    st, thermo, err = oracle.run(s, 300, 0.1, skin=0.0, out_steps=100, renei=False,
                                 box_vel=(1.0e-2, 0.0, 0.0))
    assert err == 0
    log = {100: (0.191554, -0.145643e-02, 49970.1), 200: (0.774476, -0.292621e-02, 50105.5),
           300: (1.74955, -0.439464e-02, 50240.9)}
    got = {r["step"]: (r["pe"], r["pressure"], r["volume"]) for r in thermo}
    for step, ref in log.items():
        for a, b in zip(got[step], ref):
            unit = 10.0 ** (math.floor(math.log10(abs(b))) - 5)
            assert abs(a - b) <= unit, (step, a, b)
    # dynamics set the initial positions from the Run's first state
    np.testing.assert_array_equal(st["ipos"], s.pos)
    np.testing.assert_array_equal(st["init_box"], s.box)


def test_compute1_equals_compute0_on_symmetric_matrix(oracle):
    g = golden("force_alloy5.npz")
    s = system_from_golden(g)
    tables = oracle.pair_tables(s.kinds, s.params, s.ptrmin, s.ptsize)
    _, vnl, neimax = oracle.half_list(s, float(g["skin"]))
    a = oracle.compute0(s, vnl, neimax, tables)
    b = oracle.compute1(s, vnl, neimax, tables)
    np.testing.assert_array_equal(a[0], b[0])
    assert a[1] == b[1]
    np.testing.assert_array_equal(a[2], b[2])


def _fcc_shells(a, rmax):
    n = int(math.ceil(rmax / a)) + 1
    rng = np.arange(-n, n + 1)
    basis = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    g = np.stack(np.meshgrid(rng, rng, rng, indexing="ij"), -1).reshape(-1, 3)
    pts = ((g[:, None, :] + basis[None]) * a).reshape(-1, 3)
    r = np.sqrt((pts ** 2).sum(1))
    r = r[(r > 1e-9) & (r <= rmax)]
    vals, counts = np.unique(np.round(r, 9), return_counts=True)
    return vals, counts


def test_perfect_lattice_closed_form(oracle):
    """PE = N/2 sum_s n_s V(r_s); reference virial trace = N/2 sum_s n_s (-V'(r_s)) for its
    Morse (the /r quirk); forces vanish by symmetry."""
    a = 3.615
    s = W.cu_morse(10, a=a, temp_k=0.0)
    rc, D, al, r0 = W.CU_MORSE
    tables = oracle.pair_tables(s.kinds, s.params, s.ptrmin, s.ptsize)
    _, vnl, neimax = oracle.half_list(s, 0.5)
    f, pe, vir, _ = oracle.compute0(s, vnl, neimax, tables)
    rs, ns = _fcc_shells(a, rc)
    e = np.exp(-al * (rs - r0))
    pe_cf = s.np / 2 * (ns * D * (e * e - 2 * e)).sum()
    tr_cf = s.np / 2 * (ns * 2 * D * al * (e * e - e)).sum()
    # analytic: only the linear-interpolation error of the r^2 table remains (~3e-6)
    assert pe == pytest.approx(pe_cf, rel=1e-5)
    assert vir[:3].sum() == pytest.approx(tr_cf, rel=1e-4)   # shells cancel: 2e-5 here
    # the same shell sums through an independent emulation of the table lookup
    # (single-precision index, src/nd.f90:299-308) agree to rounding
    P, F, rc2, dr2 = tables
    d2 = rs * rs
    rk = (d2 - 0.25).astype(np.float32) * np.float32(1.0 / dr2[0]) + np.float32(1.0)
    k = rk.astype(np.int64)
    w = rk.astype(np.float64) - k
    v = w * P[0][k] + (1 - w) * P[0][k - 1]
    fv = w * F[0][k] + (1 - w) * F[0][k - 1]
    assert pe == pytest.approx(s.np / 2 * (ns * v).sum(), rel=1e-10)
    assert vir[:3].sum() == pytest.approx(s.np / 2 * (ns * fv).sum(), rel=1e-10)
    assert np.abs(f).max() < 1e-9


def test_reference_log_known_answers(oracle):
    """examples/MorseTest/out:38: PE -13581.1, pressure 0.358006E-01, volume 49836.0
    (a=3.68, rc 6, PTsize 1000, skin 1.2);  examples/testCAF/OMP_4/out:38: 23,328 atoms
    a=3.629, skin 0: PE -79142.9, pressure 0.548932E-01, volume 278727."""
    for nx, a, skin, pe_ref, p_ref, v_ref in [(10, 3.68, 1.2, -13581.1, 0.358006e-1, 49836.0),
                                              (18, 3.629, 0.0, -79142.9, 0.548932e-1, 278727.0)]:
        s = W.cu_morse(nx, a=a, temp_k=0.0, params=(6.0, 0.3429, 1.359, 2.866), ptsize=1000,
                       fp32_like_dump=(nx == 10))
        tables = oracle.pair_tables(s.kinds, s.params, s.ptrmin, s.ptsize)
        _, vnl, neimax = oracle.half_list(s, skin, neimax=200)
        _, pe, vir, _ = oracle.compute0(s, vnl, neimax, tables)
        vol = float(np.prod(s.box))
        # printed with 6 significant digits: agree to one unit in the last digit (the
        # 23,328-atom model.xyz is not shipped; the regenerated lattice gives
        # 0.05489327 where the log shows 0.548932E-01)
        for got, ref in ((pe, pe_ref), (vir[:3].sum() / (3 * vol), p_ref), (vol, v_ref)):
            unit = 10.0 ** (math.floor(math.log10(abs(ref))) - 5)
            assert abs(got - ref) <= unit, (got, ref)


# ----------------------------------------------------------------- dynamics
@pytest.mark.parametrize("fname", ["run_cu10_nve.npz", "run_ar6_nvt.npz", "run_cu7_npt.npz",
                                   "run_alloy5_nve.npz", "run_asym2_nve.npz",
                                   "run_spring_pd.npz", "run_spring_mix1_nve.npz",
                                   "run_hold_cu7.npz"])
def test_dynamics_matches_reference(oracle, fname):
    g = golden(fname)
    s = system_from_golden(g)
    kw = {}
    if "iso_ke" in g.files:
        ke = g["iso_ke"]
        kw["isokinetic"] = (float(ke[0]), float(ke[1]), int(ke[2]))
    if "iso_p" in g.files:
        p = g["iso_p"]
        kw["isobaric"] = ((p[0], p[1], p[2]), float(p[3]), int(p[4]))
    if "iso_e" in g.files:
        e = g["iso_e"]
        kw["isoenergetic"] = (float(e[0]), int(e[1]))
    if "box_vel" in g.files:
        kw["box_vel"] = tuple(float(v) for v in g["box_vel"])
    if "renei" in g.files:
        kw["renei"] = bool(int(g["renei"]))
    if s.ipos is not None:
        kw.update(ipos=s.ipos, init_box=s.init_box)
    if "hold" in g.files:
        kw["holds"] = (g["hold"], g["bc_disp"], g["bc_dir"], g["bc_parm"])
    st, thermo, err = oracle.run(s, int(g["nsteps"]), float(g["dt"]), skin=float(g["skin"]),
                                 out_steps=int(g["outsteps"]), **kw)
    assert err == 0
    np.testing.assert_allclose(st["box"], g["out_box"], rtol=1e-13)
    np.testing.assert_allclose(st["pos"], g["out_pos"], rtol=0, atol=1e-12)
    np.testing.assert_allclose(st["vel"], g["out_vel"], rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(st["acc"], g["out_acc"], rtol=1e-8, atol=1e-11)
    assert st["nupdates"] == int(g["out_nupdates"])
    # thermo lines printed by the reference (6 significant digits)
    t = g["thermo"]
    assert len(thermo) == t.shape[0]
    for row, rec in zip(t, thermo):
        assert rec["step"] == int(row[0])
        for k, col in (("pe", 1), ("ke", 2), ("volume", 5)):
            assert rec[k] == pytest.approx(row[col], rel=1e-5, abs=1e-9)


def test_hold_displacements_exact(oracle):
    """Hold/Disp (src/nd.f90:1055-1064): a frozen group does not move, a pulled group moves
    by n * Disp along its constrained axis (the step's scaled displacement replaces the
    integrated one) and freely along the others; a hold group without Disp and the other
    atoms integrate as usual."""
    g = golden("run_hold_cu7.npz")
    s = system_from_golden(g)
    hold = g["hold"]
    st, _, err = oracle.run(s, 60, float(g["dt"]), skin=float(g["skin"]), out_steps=20,
                            holds=(hold, g["bc_disp"], g["bc_dir"], g["bc_parm"]))
    assert err == 0
    d = st["pos"] - s.pos
    assert np.abs(d[hold == 1]).max() == 0.0
    np.testing.assert_allclose(d[hold == 2, 0], 60 * 2.0e-5, rtol=0, atol=1e-13)
    assert np.abs(d[hold == 2, 1:]).max() > 1e-3
    assert np.abs(d[hold == 3]).max() > 1e-3 and np.abs(d[hold == 0]).max() > 1e-3
    free, _, _ = oracle.run(s, 60, float(g["dt"]), skin=float(g["skin"]), out_steps=20)
    assert np.abs(free["pos"][hold == 3] - st["pos"][hold == 3]).max() > 0.0   # coupled


def test_testT_barostat_log(oracle):
    """examples/testT/input.sh run 1 (perfect lattice at rest, Isobaric 0 0 0 0.1 10,
    ReNeighbor F, Run 2000 0.1) against examples/testT/out:38-40: PE -13581.1 V 49836.0
    at step 1; PE -12798.9 V 56637.3 at step 100; PE -12798.8 V 56637.5 at step 200."""
    s = W.cu_morse(10, a=3.68, temp_k=0.0, params=(6.0, 0.3429, 1.359, 2.866), ptsize=1000,
                   fp32_like_dump=True)
    _, thermo, err = oracle.run(s, 200, 0.1, skin=0.2, out_steps=100, renei=False,
                                isobaric=((0.0, 0.0, 0.0), 0.1, 10))
    assert err == 0
    got = [(r["step"], float(f"{r['pe']:.6g}"), float(f"{r['volume']:.6g}")) for r in thermo]
    assert got == [(1, -13581.1, 49836.0), (100, -12798.9, 56637.3), (200, -12798.8, 56637.5)]


def test_free_particle_drift(oracle):
    """Two atoms beyond the cutoff: x(t) = x0 + v t exactly up to rounding."""
    s = W.cu_morse(1, temp_k=0.0)
    s.pos = np.array([[-0.25, 0.0, 0.0], [0.25, 0.0, 0.0]])
    s.vel = np.array([[0.01, -0.02, 0.005], [0.0, 0.03, 0.0]])
    s.typ = np.ones(2, np.int32)
    s.box = np.array([100.0, 100.0, 100.0])
    st, _, err = oracle.run(s, 50, 0.1, skin=0.5, brute=True)
    assert err == 0
    np.testing.assert_allclose(st["pos"], s.pos + s.vel * 5.0 / s.box, rtol=0, atol=1e-14)
    np.testing.assert_array_equal(st["force"], 0.0)


def test_isokinetic_exact_rescale(oracle):
    """With tau = dt, mkIsoKE sets KE to KEtarget*np exactly (src/nd.f90:1142-1143)."""
    v = W.normal(1, 1, 300).reshape(100, 3)
    m = np.full(1, 2.0)
    ke = 0.5 * 2.0 * (v * v).sum()
    vv = np.ascontiguousarray(v.copy())
    oracle.lib().ndo_mk_iso_ke(100, vv, 0.3, 0.3, 0.05, ke)
    assert 0.5 * 2.0 * (vv * vv).sum() == pytest.approx(0.05 * 100, rel=1e-13)


def test_barostat_single_step_by_hand(oracle):
    """mkIsoP src/nd.f90:1183-1213: sequential update, box(2) uses the new box(1)."""
    box = np.array([10.0, 11.0, 12.0])
    vir = np.array([30.0, -20.0, 10.0, 5.0, -3.0, 2.0])
    b = box.copy()
    oracle.lib().ndo_mk_iso_p(b, vir, 0.1, 2.0, np.array([0.001, 0.0, -0.001]))
    vol = 1320.0
    p = vir / vol
    l1, l2, l3 = (1 - 0.05 * (0.001 - p[0]), 1 - 0.05 * (0.0 - p[1]), 1 - 0.05 * (-0.001 - p[2]))
    l4, l5, l6 = 0.05 * p[3], 0.05 * p[4], 0.05 * p[5]
    b1 = 10 * l1 + 11 * l6 + 12 * l5
    b2 = b1 * l6 + 11 * l2 + 12 * l4
    b3 = b1 * l5 + b2 * l4 + 12 * l3
    np.testing.assert_allclose(b, [b1, b2, b3], rtol=1e-15)


def test_skin_trigger_rebuilds_keep_lists_complete(oracle):
    """SKIN-mode trigger (not in the reference): after a hot run with rebuilds the list
    still holds every pair inside the cutoff, i.e. forces equal a fresh-list evaluation."""
    s = W.cu_morse(7, temp_k=1500.0, seed=9)
    st, _, err = oracle.run(s, 300, 2 * W.DT_1FS, skin=0.3, trigger=oracle.TRIGGER_SKIN)
    assert err == 0 and st["nupdates"] > 0
    s2 = s.copy()
    s2.pos = st["pos"]
    tables = oracle.pair_tables(s.kinds, s.params, s.ptrmin, s.ptsize)
    _, vnl, neimax = oracle.half_list(s2, 0.3)
    f, _, _, _ = oracle.compute0(s2, vnl, neimax, tables)
    np.testing.assert_allclose(st["force"], f, rtol=1e-12, atol=1e-12)
    # the compat trigger never fires here (scaled vs A^2, src/nd.f90:1493)
    st2, _, _ = oracle.run(s, 300, 2 * W.DT_1FS, skin=0.3, trigger=oracle.TRIGGER_COMPAT)
    assert st2["nupdates"] == 0
