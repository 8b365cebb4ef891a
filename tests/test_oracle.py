"""Pins of the CPU oracle: Philox known answers, deterministic exp/log against
libm, the reference's own known-answer tests, closed forms and invariants.
A plausible slip in the oracle (a dropped term, a wrong power, a transposed
operand, a wrong image) fails at least one of these."""
import ctypes
import json
import math

import numpy as np
import pytest

from conftest import GOLDEN


# ------------------------------------------------------------------ Philox

# Random123 kat_vectors for philox4x32_10 (counter, key -> output)
PHILOX_KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


@pytest.mark.parametrize("ctr,key,out", PHILOX_KAT)
def test_philox4x32_kat(oracle, ctr, key, out):
    assert tuple(oracle.philox4x32(ctr, key)) == out


def test_philox_round_structure_against_numpy(oracle):
    """The 4x64 variant shares the round/key-schedule code shape; numpy's
    Philox (4x64-10) increments the counter before its first block."""
    import ctypes as C

    src = (oracle.HERE / "orc_math.h").read_text()
    assert "0xD2511F53u" in src and "0x9E3779B9u" in src
    bg = np.random.Philox(key=np.array([7, 11], dtype=np.uint64), counter=np.array([5, 0, 0, 0], dtype=np.uint64))
    raw = [int(x) for x in bg.random_raw(4)]
    # compile-free check: replicate the 4x64 rounds in Python
    M0, M1 = 0xD2E7470EE14C6C93, 0xCA5A826395121157
    W0, W1 = 0x9E3779B97F4A7C15, 0xBB67AE8584CAA73B
    c = [6, 0, 0, 0]
    k0, k1 = 7, 11
    mask = (1 << 64) - 1
    for _ in range(10):
        p0, p1 = M0 * c[0], M1 * c[2]
        c = [((p1 >> 64) ^ c[1] ^ k0) & mask, p1 & mask, ((p0 >> 64) ^ c[3] ^ k1) & mask, p0 & mask]
        k0, k1 = (k0 + W0) & mask, (k1 + W1) & mask
    assert raw == c


# -------------------------------------------------------- deterministic math


def _ulp_err(a, b):
    return abs(a - b) / np.spacing(abs(b)) if b != 0 else abs(a)


def test_det_exp_matches_libm(oracle):
    rng = np.random.default_rng(1)
    xs = np.concatenate([rng.uniform(-745, 709, 20000), rng.uniform(-1, 1, 20000), [0.0, -1e-300, 1e-300, -744.9,
                                                                                       -708.5, 709.0]])
    worst = max(_ulp_err(oracle.lib().orc_exp_det(float(x)), math.exp(x)) for x in xs if math.exp(x) > 2.3e-308)
    assert worst <= 2.0
    assert oracle.lib().orc_exp_det(0.0) == 1.0
    assert oracle.lib().orc_exp_det(-800.0) == 0.0
    # subnormal results stay within a few units of the last place of the subnormal grid
    for x in (-710.0, -720.0, -740.0):
        assert abs(oracle.lib().orc_exp_det(x) - math.exp(x)) <= 4 * 4.9e-324


def test_det_log_matches_libm(oracle):
    rng = np.random.default_rng(2)
    xs = np.concatenate([10 ** rng.uniform(-300, 300, 20000), rng.uniform(0.5, 2, 20000), [1.0, 2.0, 1e-310]])
    worst = max(_ulp_err(oracle.lib().orc_log_det(float(x)), math.log(x)) for x in xs if x != 1.0)
    assert worst <= 2.0
    assert oracle.lib().orc_log_det(1.0) == 0.0


def test_metropolis_threshold_form(oracle):
    """The FP64 pair-potential sweeps decide min(1, e^{-beta dE}) >= u
    (src/Exchange.jl:53-57) as dE <= -ln(u) / beta.  Over random moves near
    and far from the boundary the two forms agree except where e^{-beta dE}
    and u are within rounding of each other; a threshold with beta
    multiplied instead of divided, or with the sign of ln u flipped, does not."""
    rng = np.random.default_rng(5)
    L = oracle.lib()
    n = 20000
    beta = 10 ** rng.uniform(1, 5, n)
    u = rng.uniform(0, 1, n)
    # half the moves right at the boundary dE = -ln(u)/beta (times 1 +- 1e-9), half spread
    de = np.where(np.arange(n) % 2 == 0, -np.log(u) / beta * (1 + rng.uniform(-1e-9, 1e-9, n)),
                  rng.uniform(-5, 5, n) / beta)
    exp_form = np.array([min(1.0, L.orc_exp_det(float(-b * d))) >= x for b, d, x in zip(beta, de, u)])
    lnu = np.array([L.orc_log_det(float(x)) for x in u])
    thr_form = de <= -lnu / beta
    p = np.array([min(1.0, L.orc_exp_det(float(-b * d))) for b, d in zip(beta, de)])
    tie = np.abs(p - u) <= 1e-13 * u
    assert np.all(thr_form[~tie] == exp_form[~tie])
    assert tie.mean() < 0.01 and exp_form.mean() > 0.2
    assert np.any((de <= -lnu * beta)[~tie] != exp_form[~tie])
    assert np.any((de <= lnu / beta)[~tie] != exp_form[~tie])


# ------------------------------------------------------ reference KATs


def test_elj_known_answers(oracle):
    """test/test_potentials.jl:3-15: e(1) = sum(c); even and all-powers forms agree where they must."""
    kat = json.loads((GOLDEN / "known_answers.json").read_text())["elj"]
    for case in kat["cases"]:
        pot = oracle.make_pot(case["kind"], case["coeff"])
        for form in (0, 1):
            assert oracle.lib().orc_dimer_energy(pot, case["r2"], form) == pytest.approx(case["energy"], abs=1e-15)
    # ELJPotentialEven([-2, 1]) == ELJPotential([-2, 0, 1]) at r2 = 2 (test_potentials.jl:10)
    pe = oracle.make_pot("elj_even", [-2.0, 1.0])
    po = oracle.make_pot("elj", [-2.0, 0.0, 1.0])
    assert oracle.lib().orc_dimer_energy(pe, 2.0, 0) == oracle.lib().orc_dimer_energy(po, 2.0, 0)


def test_elj_powers_closed_form(oracle):
    """Each coefficient multiplies exactly one power: c_k r^-(4+2k) (even) and c_k r^-(5+k)."""
    rng = np.random.default_rng(3)
    for k in range(6):
        c = [0.0] * 6
        c[k] = 1.0
        pe = oracle.make_pot("elj_even", c)
        po = oracle.make_pot("elj", c)
        for r2 in rng.uniform(2.0, 60.0, 20):
            for form in (0, 1):
                assert oracle.lib().orc_dimer_energy(pe, r2, form) == pytest.approx(r2 ** (-(6 + 2 * k) / 2), rel=1e-14)
                assert oracle.lib().orc_dimer_energy(po, r2, form) == pytest.approx(r2 ** (-(6 + k) / 2), rel=1e-14)


def test_horner_form_agrees_with_reference_form(oracle, ne13):
    pot = oracle.make_pot("elj_even", ne13["coeff"])
    rng = np.random.default_rng(4)
    for r2 in rng.uniform(15.0, 400.0, 5000):
        e0 = oracle.lib().orc_dimer_energy(pot, r2, 0)
        e1 = oracle.lib().orc_dimer_energy(pot, r2, 1)
        mag = sum(abs(c) * r2 ** (-(6 + 2 * k) / 2) for k, c in enumerate(ne13["coeff"]))
        assert abs(e0 - e1) <= 1e-14 * mag


def test_distance_known_answers(oracle):
    """test/runtests.jl:74-100."""
    import ctypes as C

    kat = json.loads((GOLDEN / "known_answers.json").read_text())["distance"]
    P = lambda v: (C.c_double * 3)(*v)
    d = oracle.lib().orc_distance2
    assert d(P(kat["v1"]), P(kat["v2"]), 0, 0.0, 0) == kat["plain_12"]
    assert d(P(kat["v1"]), P(kat["v3_sphere"]), 0, 0.0, 0) == kat["d2mat_13"]
    assert d(P(kat["v1"]), P(kat["v3_cubic"]), 0, 0.0, 0) == kat["plain_13"]
    for form in (0, 1):
        assert d(P(kat["v1"]), P(kat["v3_cubic"]), 1, kat["box"], form) == kat["cubic_13"]
        assert d(P(kat["v1"]), P(kat["v2"]), 1, kat["box"], form) == kat["plain_12"]


def test_minimum_image_invariance(oracle):
    """Shifting either atom by whole box vectors leaves the cubic distance unchanged
    and it never exceeds 3 (L/2)^2."""
    import ctypes as C

    rng = np.random.default_rng(5)
    L = 7.3
    for _ in range(2000):
        a, b = rng.uniform(-L / 2, L / 2, 3), rng.uniform(-L / 2, L / 2, 3)
        sh = rng.integers(-3, 4, 3) * L
        P = lambda v: (C.c_double * 3)(*v)
        d0 = oracle.lib().orc_distance2(P(a), P(b), 1, L, 1)
        d1 = oracle.lib().orc_distance2(P(a), P(b + sh), 1, L, 1)
        assert d0 == pytest.approx(d1, rel=1e-12, abs=1e-12)
        assert d0 <= 3 * (L / 2) ** 2 + 1e-9
        brute = min(np.sum((a - b - L * np.array(s)) ** 2) for s in np.ndindex(3, 3, 3) for s in [np.array(s) - 1])
        assert d0 == pytest.approx(brute, rel=1e-12)


def test_eam_known_answer(oracle):
    """test/test_potentials.jl:44-56: three atoms, n = m = eps = C = a = 1 -> E = -1.3495549581716526."""
    kat = json.loads((GOLDEN / "known_answers.json").read_text())["eam_three_atoms"]
    pot = oracle.make_pot("eam", eam=(kat["n"], kat["m"], kat["ean"], 1.0))
    for mode in (oracle.REFERENCE, oracle.engine_mode(1)):
        run = oracle.OracleRun(3, [1.0], np.array(kat["positions"]), pot=pot, radius=2.0, mode=mode)
        assert run.en_tot[0] == pytest.approx(kat["energy"], rel=1e-14)
        assert np.allclose(run.comp[0][:, 0], run.comp[0][:, 1])  # n == m: both components equal (:53)


def test_lrc_closed_form(oracle):
    """lrc = (pi N^2 / (4 rc^3)) sum_k c_k rc^{-(2k+1)}... equals the tail integral
    2 pi rho N int_rc^inf r^2 e(r) dr with rho = N / (2 rc)^3 (rc = r_cut^(1/2))."""
    from scipy.integrate import quad

    c = [-10.5097942564988, 989.725135614556, -101383.865938807, 3918846.12841668, -56234083.4334278,
         288738837.441765]
    pot = oracle.make_pot("elj_even", c)
    for N, rc2 in ((32, 60.0), (108, 150.0), (256, 400.0)):
        rc = math.sqrt(rc2)
        e = lambda r: sum(ck * r ** (-(6 + 2 * k)) for k, ck in enumerate(c))
        rho = N / (2 * rc) ** 3
        tail = 2 * math.pi * rho * N * quad(lambda r: r * r * e(r), rc, np.inf, epsabs=0, epsrel=1e-13)[0]
        assert oracle.lib().orc_lrc(N, rc2, pot) == pytest.approx(tail, rel=1e-10)
    assert oracle.lib().orc_lrc(32, 49.9, pot) == 0.0  # only applied above r_cut = 50 (EnergyEvaluation.jl:969)


def test_ne13_start_energy(oracle, ne13):
    """Full pair sum of the icosahedral start equals an independent numpy sum,
    and matches the minimum energy in the reference checkpoint (-0.0056336 Ha,
    test/testing_data/checkpoint/hist_info.data:2) to its printed digits."""
    pot = oracle.make_pot("elj_even", ne13["coeff"])
    pos = ne13["pos"]
    d = pos[:, None, :] - pos[None, :, :]
    r2 = (d**2).sum(-1)[np.triu_indices(13, 1)]
    ref = sum(sum(c * r2 ** (-(3 + k)) for k, c in enumerate(ne13["coeff"])))
    for mode in (oracle.REFERENCE, oracle.engine_mode(1), oracle.engine_mode(4)):
        run = oracle.OracleRun(13, [1.0], pos, pot=pot, radius=ne13["radius"], mode=mode)
        assert run.en_tot[0] == pytest.approx(ref, rel=1e-13)
    assert ref == pytest.approx(-0.0056347, abs=2e-6)


def test_checkpoint_container_radius(ne13):
    kat = json.loads((GOLDEN / "known_answers.json").read_text())["checkpoint_r2"]
    assert ne13["radius"] ** 2 == pytest.approx(kat["r2"], rel=1e-15)


# ------------------------------------------------------- invariants of runs


def _ne13_run(oracle, ne13, mode, **kw):
    pot = oracle.make_pot("elj_even", ne13["coeff"])
    from paralleltemperingmontecarlo.api import TempGrid

    temp = TempGrid(4.0, 16.0, 8)
    return oracle.OracleRun(13, temp.beta_grid, ne13["pos"], pot=pot, radius=ne13["radius"], mode=mode, **kw)


@pytest.mark.parametrize("mode_name", ["reference", "engine1", "engine4"])
def test_running_energy_equals_recompute(oracle, ne13, mode_name):
    mode = {"reference": oracle.REFERENCE, "engine1": oracle.engine_mode(1), "engine4": oracle.engine_mode(4)}[mode_name]
    run = _ne13_run(oracle, ne13, mode, exchange_mode=2)
    run.cycles(300)
    pot = oracle.make_pot("elj_even", ne13["coeff"])
    for t in range(run.R):
        xt = np.ascontiguousarray(run.x[t])
        e = oracle.lib().orc_energy_config(ctypes.byref(run.p), ctypes.byref(pot), oracle._p(xt), 0.0, 0.0, None)
        assert run.en_tot[t] == pytest.approx(e, rel=1e-11)
        assert np.all(np.sum(run.x[t] ** 2, axis=1) <= ne13["radius"] ** 2 + 1e-12)
    assert run.count_atom[:, 0].sum() > 0 and run.count_exc[:, 0].sum() > 0


def test_engine_and_reference_arithmetic_agree(oracle, ne13):
    """Per-move dE of the engine's arithmetic (Horner, distances from positions,
    4-wide sums, deterministic exp) agrees with the reference arithmetic to
    1e-10 of the moved atom's interaction scale, and the two runs take the same
    accept/reject decisions."""
    a = _ne13_run(oracle, ne13, oracle.REFERENCE, exchange_mode=2)
    b = _ne13_run(oracle, ne13, oracle.engine_mode(4), exchange_mode=2)
    ta = a.cycles(200, trace=True)
    tb = b.cycles(200, trace=True)
    assert np.array_equal(ta["idx"], tb["idx"])
    assert np.array_equal(ta["acc"], tb["acc"])
    scale = 1e-3  # |sum_j e_ij| of a neon atom in the cluster is ~1e-3 Ha
    assert np.max(np.abs(ta["de"] - tb["de"])) <= 1e-10 * scale
    assert np.allclose(a.x, b.x, rtol=0, atol=1e-12)


def test_exchange_modes(oracle, ne13):
    """Reference exchange: attempts happen with probability 0.1 per cycle and
    involve two neighbours; even/odd: every cycle, every other pair."""
    run = _ne13_run(oracle, ne13, oracle.engine_mode(1), exchange_mode=0)
    run.cycles(2000)
    att = run.count_exc[:, 0].sum() / 2
    assert 150 < att < 260  # binomial(2000, 0.1) = 200 +- 13
    run2 = _ne13_run(oracle, ne13, oracle.engine_mode(1), exchange_mode=2)
    run2.cycles(100)
    # 8 temperatures: pairs (0,1),(2,3),(4,5),(6,7) on even cycles, (1,2),(3,4),(5,6) on odd
    assert list(run2.count_exc[:, 0]) == [50, 100, 100, 100, 100, 100, 100, 50]


def test_exchange_cadence(oracle, ne13):
    """exchange_every = 3: attempts only after the cycles with (cycle + 1) % 3
    == 0 (cycles 2, 5, 8, ...), the even/odd pairs by that cycle's parity;
    between them no temperature changes hands, and cadence 1 is the default."""
    run = _ne13_run(oracle, ne13, oracle.engine_mode(1), exchange_mode=2, exchange_every=3)
    run.cycles(30)
    # cycles 2, 8, 14, 20, 26 (even): pairs (0,1),(2,3),(4,5),(6,7); 5, 11, 17, 23, 29 (odd): (1,2),(3,4),(5,6)
    assert list(run.count_exc[:, 0]) == [5, 10, 10, 10, 10, 10, 10, 5]
    a = _ne13_run(oracle, ne13, oracle.engine_mode(1), exchange_mode=1)
    b = _ne13_run(oracle, ne13, oracle.engine_mode(1), exchange_mode=1, exchange_every=1)
    a.cycles(40)
    b.cycles(40)
    assert np.array_equal(a.x, b.x) and np.array_equal(a.count_exc, b.count_exc)


def test_exchange_uniform_and_near_tie_check(oracle, ne13):
    """The replay tests' exchange uniform (tests/decisions.py) is the
    oracle's: over 40 even/odd cycles every attempted pair was accepted iff
    min(1, e^{dbeta dE}) > u with the energies before the exchange.  The
    configuration check accepts a swapped pair only on a near tie: planted
    swaps with energies that put p on u pass, far from u they fail, and a
    configuration that matches no oracle one fails."""
    from decisions import assert_same_configurations, exchange_uniform

    from paralleltemperingmontecarlo.api import TempGrid

    beta = TempGrid(4.0, 16.0, 8).beta_grid
    R = len(beta)
    run = _ne13_run(oracle, ne13, oracle.engine_mode(1), exchange_mode=2, seed=2468, run_key=8)
    n_ok = 0
    prev = run.count_exc[:, 1].copy()
    for c in range(40):
        run.cycles(1)
        acc = run.count_exc[:, 1] - prev
        prev = run.count_exc[:, 1].copy()
        e = run.en_tot
        for n in range(c & 1, R - 1, 2):
            swapped = acc[n] == 1
            e1, e2 = (e[n + 1], e[n]) if swapped else (e[n], e[n + 1])
            p = min(1.0, math.exp((beta[n] - beta[n + 1]) * (e1 - e2)))
            assert (p > exchange_uniform(oracle, 2468, 8, c, n)) == swapped
            n_ok += 1
    assert n_ok == 20 * 3 + 20 * 4

    class Fake:
        pass

    rng = np.random.default_rng(1)
    R, N = 6, 5
    x = rng.normal(size=(R, N, 3))
    seed, key, sweep = 99, 4, 7
    b = np.linspace(2.0, 1.0, R)
    u = exchange_uniform(oracle, seed, key, sweep, 2)
    e = rng.normal(size=R)
    # energies of the pair (2, 3) with p = u: dbeta (e2 - e3) = ln u
    e[2] = e[3] + math.log(u) / (b[2] - b[3])
    orc = Fake()
    orc.x, orc.en_tot = x.copy(), e.copy()
    xs = x.copy()
    xs[[2, 3]] = xs[[3, 2]]
    es = e.copy()
    es[[2, 3]] = es[[3, 2]]
    assert assert_same_configurations(oracle, x, e, orc, b, seed=seed, run_key=key, sweep=sweep, e_tol=1e-6) == 0
    assert assert_same_configurations(oracle, xs, es, orc, b, seed=seed, run_key=key, sweep=sweep, e_tol=1e-6) == 1
    orc.en_tot = e + np.where(np.arange(R) == 2, 0.5, 0.0)  # p far from u
    with pytest.raises(AssertionError):
        assert_same_configurations(oracle, xs, es, orc, b, seed=seed, run_key=key, sweep=sweep, e_tol=1e-6)
    orc.en_tot = e.copy()
    xb = x.copy()
    xb[4, 0, 0] += 1e-12
    with pytest.raises(AssertionError):
        assert_same_configurations(oracle, xb, e, orc, b, seed=seed, run_key=key, sweep=sweep, e_tol=1e-6)


def test_step_adaptation(oracle, ne13):
    """update_max_stepsize!: after n_adjust cycles delta is scaled by exactly 0.9 or 1.1 (or kept)."""
    run = _ne13_run(oracle, ne13, oracle.engine_mode(1), n_adjust=50)
    run.cycles(50)
    ratio = run.max_displ[:, 0] / 0.1
    assert all(any(abs(r - v) < 1e-15 for v in (0.9, 1.0, 1.1)) for r in ratio)
    assert np.all(run.count_atom[:, 1] == 0)
