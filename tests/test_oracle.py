"""Pins of the CPU restatement (oracle/lgf_oracle.c) -- CPU only.

Each check is chosen so that a plausible slip in the restatement fails it:
  * golden vectors from the reference program itself (tests/golden/ref_*.json,
    made by tests/golden/make_golden.py from oracle/_ref/lgf_ref);
  * exp: X = 0, diagonal closed form, and scipy.linalg.expm on random X in
    all three branches of LGF/Math/SU3_mat.cpp:216-289;
  * gradient: zero on a cold start, gauge covariance, and the finite-
    difference derivative of sum_P Re Tr P (catches sign/factor slips of
    LGF/Action/Gauge/Wilson.cpp:55-71 and the staple);
  * RK3: cold start stays fixed; global error scales as eps^3;
  * clovers: E and Q from a two-plane abelian field in closed form (pins the
    leaf orientation, the improved-F coefficients incl. the sign of k3 =
    -32/45, the 1/(4 pi^2) and the pi^2/4 file scale), gauge invariance and
    the sign flip of Q under an axis reflection.
"""
import numpy as np
import pytest
import scipy.linalg

from oracle import oracle as O
from tests.common import (TOL_FP64, config1_inputs, golden, q_err, rel_err, small_input, twisted_flux_field,
                          twisted_flux_observables)

RNG = np.random.default_rng(7)


def to_c(m):
    m = np.asarray(m).reshape(9, 2)
    return (m[:, 0] + 1j * m[:, 1]).reshape(3, 3)


def to_r(c):
    c = np.asarray(c).reshape(9)
    out = np.empty(18)
    out[0::2], out[1::2] = c.real, c.imag
    return out


def rand_su3_alg(scale):
    """random traceless anti-Hermitian 3x3 with Frobenius norm `scale`"""
    A = RNG.normal(size=(3, 3)) + 1j * RNG.normal(size=(3, 3))
    H = A - A.conj().T
    H -= np.trace(H) / 3 * np.eye(3)
    return H * (scale / np.linalg.norm(H))


def rand_su3():
    A = RNG.normal(size=(3, 3)) + 1j * RNG.normal(size=(3, 3))
    Q, R = np.linalg.qr(A)
    Q = Q @ np.diag(np.diag(R) / np.abs(np.diag(R)))
    return Q / np.linalg.det(Q) ** (1 / 3)


# ---------------------------------------------------------------- golden
def test_oracle_matches_reference_small():
    d, U = small_input()
    shape = tuple(d["shape"])
    c = d["configs"][0]
    t, E, Q = O.flow_run(shape, U, d["eps"], d["steps"], d["every"])
    assert np.allclose(t, c["t"], rtol=0, atol=1e-15)
    assert rel_err(E, c["E"]) < 1e-13
    assert np.max(np.abs(np.asarray(Q) - c["Q"])) < 1e-13
    assert np.array_equal(U.reshape(-1)[:18], np.asarray(c["flowed_first_link"]))
    assert np.array_equal(U.reshape(-1)[-18:], np.asarray(c["flowed_last_link"]))
    _, Qa, qx = O.top_charge(shape, U, want_qx=True)
    assert abs(qx.sum() - c["qx_sum"]) < 1e-12 * c["qx_abs_sum"]
    # MeasureFlowedGaugeConfigurations on the flowed configuration
    assert rel_err(O.energy_density(shape, U, improved=True), c["measured_E_improved"]) < 1e-13
    assert q_err(O.top_charge_unimproved(shape, U), c["measured_Q_unimproved"], Qa) < 1e-13
    for k, v in c["qx_samples"].items():
        assert abs(qx[int(k)] - v) < 1e-13 * c["qx_abs_sum"]


def test_oracle_matches_reference_config1():
    d, shape, cfgs = config1_inputs()
    import hashlib
    for U, h in zip(cfgs, d["input_sha256"]):
        assert hashlib.sha256(U.tobytes()).hexdigest() == h, "generator no longer reproduces the inputs"
    for U, c in zip(cfgs, d["configs"]):
        t, E, Q = O.flow_run(shape, U, d["eps"], d["steps"], d["every"])
        assert rel_err(E, c["E"]) < 1e-13
        Qa = O.top_charge(shape, U)[1]
        assert q_err(Q, c["Q"], Qa) < 1e-12
        assert rel_err(O.energy_density(shape, U, improved=True), c["measured_E_improved"]) < 1e-12
        assert q_err(O.top_charge_unimproved(shape, U), c["measured_Q_unimproved"], Qa) < 1e-12
        # the restatement follows the reference's operation order (no FP contraction),
        # so its flowed links are the reference's saved configuration bit for bit
        assert hashlib.sha256(U.tobytes()).hexdigest() == c["flowed_sha256"]


# ---------------------------------------------------------------- exp
def test_exp_zero_is_identity():
    assert np.array_equal(O.exp_ah(np.zeros(18)), to_r(np.eye(3)))


def test_exp_diagonal_closed_form():
    for a, b in [(0.3, -0.1), (1e-3, 2e-3), (1.2, 0.7), (2.5, -1.0)]:
        X = np.diag([1j * a, 1j * b, -1j * (a + b)])
        got = to_c(O.exp_ah(to_r(X)))
        want = np.diag(np.exp([1j * a, 1j * b, -1j * (a + b)]))
        assert np.max(np.abs(got - want)) < 1e-14


@pytest.mark.parametrize("scale", [1e-8, 1e-4, 5e-3, 0.05, 0.3, 1.0, 3.0])
def test_exp_matches_expm_all_branches(scale):
    for _ in range(20):
        X = rand_su3_alg(scale)
        got = to_c(O.exp_ah(to_r(X)))
        want = scipy.linalg.expm(X)
        assert np.max(np.abs(got - want)) < 2e-14
        assert np.max(np.abs(got.conj().T @ got - np.eye(3))) < 1e-14
        assert abs(np.linalg.det(got) - 1) < 1e-14


def test_exp_negative_c0_branch():
    # X and -X give c0 of opposite signs: exp(-X) = exp(X)^-1
    for _ in range(10):
        X = rand_su3_alg(0.7)
        a = to_c(O.exp_ah(to_r(X)))
        b = to_c(O.exp_ah(to_r(-X)))
        assert np.max(np.abs(a @ b - np.eye(3))) < 1e-14


# ---------------------------------------------------------------- gradient
def cold(shape):
    U = O.empty_field(shape)
    U[:, :, 0] = U[:, :, 8] = U[:, :, 16] = 1.0
    return U


def test_gradient_cold_start_is_zero():
    shape = (4, 4, 4, 4)
    U = cold(shape)
    assert np.array_equal(O.staple(shape, U, 5, 2), to_r(6 * np.eye(3)))
    assert np.max(np.abs(O.flow_gradient(shape, U))) == 0.0


def site_coords(shape, g):
    c = []
    for n in reversed(shape):
        c.append(g % n)
        g //= n
    return tuple(reversed(c))


def site_index(shape, c):
    g = 0
    for n, x in zip(shape, c):
        g = g * n + (x % n)
    return g


def gauge_transform(shape, U, G):
    """U'_mu(x) = g(x) U_mu(x) g(x+mu)^dag"""
    V = U.shape[0]
    out = np.empty_like(U)
    for g in range(V):
        c = site_coords(shape, g)
        for mu in range(4):
            cp = list(c)
            cp[mu] += 1
            out[g, mu] = to_r(G[g] @ to_c(U[g, mu]) @ G[site_index(shape, cp)].conj().T)
    return out


def test_gradient_gauge_covariant():
    shape = (4, 4, 2, 4)
    U = O.generate(shape, 6.0, seed=11, start="hot", sweeps=2)
    G = [rand_su3() for _ in range(U.shape[0])]
    Z = O.flow_gradient(shape, U)
    Zg = O.flow_gradient(shape, gauge_transform(shape, U, G))
    for g in range(0, U.shape[0], 7):
        for mu in range(4):
            want = G[g] @ to_c(Z[g, mu]) @ G[g].conj().T
            assert np.max(np.abs(to_c(Zg[g, mu]) - want)) < 1e-13


def gell_mann():
    l = np.zeros((8, 3, 3), dtype=complex)
    l[0][0, 1] = l[0][1, 0] = 1
    l[1][0, 1], l[1][1, 0] = -1j, 1j
    l[2][0, 0], l[2][1, 1] = 1, -1
    l[3][0, 2] = l[3][2, 0] = 1
    l[4][0, 2], l[4][2, 0] = -1j, 1j
    l[5][1, 2] = l[5][2, 1] = 1
    l[6][1, 2], l[6][2, 1] = -1j, 1j
    l[7] = np.diag([1, 1, -2]) / np.sqrt(3)
    return l


def plaq_sum_around(shape, U, g, mu):
    """Re Tr(U_mu(x) A_mu(x)) = sum of Re Tr P over the 6 plaquettes containing U_mu(x)"""
    return np.trace(to_c(U[g, mu]) @ to_c(O.staple(shape, U, g, mu))).real


def test_gradient_is_minus_derivative_of_plaquettes():
    """d/dh Re Tr(e^{hT} U A) = -Re Tr(T Z) for T in su(3)"""
    shape = (4, 4, 4, 4)
    U = O.generate(shape, 6.0, seed=3, start="hot", sweeps=3)
    Z = O.flow_gradient(shape, U)
    h = 1e-5
    for g, mu in [(0, 0), (37, 2), (200, 3)]:
        for lam in gell_mann():
            T = 0.5j * lam
            vals = []
            for s in (+1, -1):
                V = U.copy()
                V[g, mu] = to_r(scipy.linalg.expm(s * h * T) @ to_c(U[g, mu]))
                vals.append(plaq_sum_around(shape, V, g, mu))
            fd = (vals[0] - vals[1]) / (2 * h)
            assert abs(fd - (-np.trace(T @ to_c(Z[g, mu])).real)) < 1e-8


# ---------------------------------------------------------------- RK3
def test_rk3_cold_start_fixed():
    shape = (4, 4, 4, 4)
    U = cold(shape)
    O.rk3_step(shape, U, 0.02)
    assert np.array_equal(U, cold(shape))


def test_rk3_third_order():
    shape = (4, 4, 4, 4)
    U0 = O.generate(shape, 6.0, seed=5, start="hot", sweeps=5)
    for _ in range(10):  # smooth first so the asymptotic regime is reached at these eps
        O.rk3_step(shape, U0, 0.02)
    T = 0.16

    def flow(eps):
        U = U0.copy()
        for _ in range(int(round(T / eps))):
            O.rk3_step(shape, U, eps)
        return U

    ref = flow(0.0025)
    errs = [np.max(np.abs(flow(e) - ref)) for e in (0.04, 0.02, 0.01)]
    r1, r2 = errs[0] / errs[1], errs[1] / errs[2]
    # third order: halving eps divides the error by ~8 (second order: 4, fourth: 16)
    assert 6.5 < r1 < 11 and 6.5 < r2 < 11, errs


# ---------------------------------------------------------------- clovers
def abelian_field(shape, kb, kc):
    """U_1(x) = diag(e^{i b x0}, e^{-i b x0}, 1), U_3(x) = diag(e^{i c x2}, e^{-i c x2}, 1)"""
    b, c = 2 * np.pi * kb / shape[0], 2 * np.pi * kc / shape[2]
    U = cold(shape)
    for g in range(U.shape[0]):
        x = site_coords(shape, g)
        U[g, 1] = to_r(np.diag([np.exp(1j * b * x[0]), np.exp(-1j * b * x[0]), 1]))
        U[g, 3] = to_r(np.diag([np.exp(1j * c * x[2]), np.exp(-1j * c * x[2]), 1]))
    return U, b, c


def improved_profile(a):
    return 19 / 9 * np.sin(a) + 1 / 36 * np.sin(4 * a) - 32 / 45 * np.sin(2 * a) + 1 / 15 * np.sin(3 * a)


@pytest.mark.parametrize("shape,kb,kc", [((8, 4, 6, 4), 1, 1), ((6, 4, 8, 4), 2, -1), ((10, 4, 4, 4), 3, 1)])
def test_clover_abelian_closed_form(shape, kb, kc):
    U, b, c = abelian_field(shape, kb, kc)
    V = U.shape[0]
    E = O.energy_density(shape, U)
    assert rel_err(E, 2 * np.sin(b) ** 2 + 2 * np.sin(c) ** 2) < 1e-14
    Q, _, qx = O.top_charge(shape, U, want_qx=True)
    q_site = -2 * improved_profile(b) * improved_profile(c)
    assert abs(Q - V * q_site / (4 * np.pi ** 2)) < 1e-12 * V
    # the reference writes q(x) scaled by (1.0/4.0*PI2) = pi^2/4
    assert np.max(np.abs(qx - np.pi ** 2 / 4 * q_site)) < 1e-13
    F = to_c(O.field_strength(shape, U, 3, 0, 1, 0))
    assert np.max(np.abs(F - np.diag([-np.sin(b), np.sin(b), 0]))) < 1e-15
    F = to_c(O.field_strength(shape, U, 3, 0, 1, 4))
    assert np.max(np.abs(F - improved_profile(b) * np.diag([-1, 1, 0]))) < 1e-14


@pytest.mark.parametrize("shape,n01,n23", [((8, 8, 6, 6), 1, 1), ((6, 4, 4, 6), 1, -2), ((12, 12, 8, 8), 2, 1)])
def test_twisted_flux_closed_form_and_integer_charge(shape, n01, n23):
    """integer flux through the (0,1) and (2,3) tori: the observables follow
    their closed forms, and the improved charge approaches the integer
    -2 n01 n23 with O(p^4) corrections (the improvement cancels O(p^2))"""
    U, p, q = twisted_flux_field(shape, n01, n23)
    ref = twisted_flux_observables(shape, p, q)
    # the twisted links carry phases up to 2 pi n: rounding of ~1e-15 per link
    assert rel_err(O.energy_density(shape, U), ref["E"]) < 1e-12
    assert rel_err(O.energy_density(shape, U, improved=True), ref["E_improved"]) < 1e-12
    Q, _, qx = O.top_charge(shape, U, want_qx=True)
    assert abs(Q - ref["Q"]) < 1e-12
    assert np.max(np.abs(qx - np.pi ** 2 / 4 * ref["qraw"])) < 1e-13
    assert abs(O.top_charge_unimproved(shape, U) - ref["Q_unimproved"]) < 1e-12
    target = -2 * n01 * n23
    assert abs(ref["Q"] - target) < 2 * abs(target) * (max(abs(p), abs(q)) ** 4)
    assert abs(ref["Q"] - target) < abs(ref["Q_unimproved"] - target)


def test_flow_keeps_constant_abelian_field():
    """a constant abelian field strength is a stationary point of the Wilson
    flow: the upper (phase p) and lower (phase -p) plaquettes of every link
    cancel in the projection, so Z = 0 and the links do not move; a wrong
    sign or neighbour in the staple sum breaks this"""
    shape = (6, 4, 4, 6)
    U, p, q = twisted_flux_field(shape, 1, -2)
    V = U.copy()
    for _ in range(3):
        O.rk3_step(shape, V, 0.05)
    assert np.max(np.abs(V - U)) < 1e-13


def test_clover_cold_zero():
    shape = (4, 4, 4, 4)
    U = cold(shape)
    assert O.energy_density(shape, U) == 0.0
    assert O.top_charge(shape, U)[0] == 0.0


def test_observables_gauge_invariant():
    shape = (4, 4, 4, 4)
    U = O.generate(shape, 6.0, seed=9, start="hot", sweeps=4)
    G = [rand_su3() for _ in range(U.shape[0])]
    Ug = gauge_transform(shape, U, G)
    assert rel_err(O.energy_density(shape, Ug), O.energy_density(shape, U)) < 1e-13
    Q, Qa, _ = O.top_charge(shape, U)
    assert abs(O.top_charge(shape, Ug)[0] - Q) < 1e-13 * Qa


def reflect(shape, U, a):
    """x_a -> -x_a: U'_a(x) = U_a(R(x+a))^dag, U'_mu(x) = U_mu(Rx) otherwise"""
    out = np.empty_like(U)
    for g in range(U.shape[0]):
        x = list(site_coords(shape, g))
        rx = list(x)
        rx[a] = -x[a]
        for mu in range(4):
            if mu == a:
                y = list(x)
                y[a] += 1
                y[a] = -y[a]
                out[g, mu] = to_r(to_c(U[site_index(shape, y), mu]).conj().T)
            else:
                out[g, mu] = U[site_index(shape, rx), mu]
    return out


@pytest.mark.parametrize("axis", [0, 2])
def test_charge_flips_under_reflection(axis):
    shape = (4, 4, 4, 4)
    U = O.generate(shape, 6.0, seed=21, start="hot", sweeps=3)
    for _ in range(10):
        O.rk3_step(shape, U, 0.02)
    Q, Qa, _ = O.top_charge(shape, U)
    R = reflect(shape, U, axis)
    Qr, _, _ = O.top_charge(shape, R)
    assert abs(Qr + Q) < 1e-13 * Qa
    assert rel_err(O.energy_density(shape, R), O.energy_density(shape, U)) < 1e-13


def test_generator_plaquette_beta6():
    """<P> at beta = 6.0 is 0.5937 (large-volume literature value); 8^4 agrees to ~1%."""
    d = golden("config1")
    assert all(abs(p - 0.5937) < 0.006 for p in d["plaquette"])


def test_plaquette_closed_form_twisted_flux():
    """avg_plaquette (SU3_field::Avg_Plaquette, LGF/Field/SU3_field.cpp:485-499:
    sum over sites and the 6 planes of Re Tr P, divided by 3 * 6 * V) on the
    twisted abelian field: the (0,1) plaquettes are diag(e^{ip}, e^{-ip}, 1),
    the (2,3) ones diag(e^{iq}, e^{-iq}, 1), the four mixed planes the unit
    matrix, so <P> = (2 cos p + 1 + 2 cos q + 1 + 4 * 3) / 18 exactly."""
    from tests.common import twisted_flux_field
    shape = (8, 6, 6, 4)
    U, p, q = twisted_flux_field(shape, 1, -2)
    want = (2 * np.cos(p) + 1 + 2 * np.cos(q) + 1 + 12) / 18
    assert abs(O.avg_plaquette(shape, U) - want) < 1e-13
    mx, mn = O.unitarity(shape, U)
    assert mx < 1e-15
