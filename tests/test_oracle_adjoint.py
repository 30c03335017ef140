"""Pins the FP64 oracle adjoint of Cabs by central finite differences of the
dense forward model (independent of the adjoint algebra), and checks its
Cabs against EVALQ's oracle formula."""
import numpy as np

from oracle import ddscat_oracle as O
from oracle import adjoint as ADJ


def small_problem(seed=0, offdiag=True):
    rng = np.random.default_rng(seed)
    nx, ny, nz = 3, 3, 2
    iocc = (rng.random(nx * ny * nz) < 0.8).astype(np.int16)
    iocc[0] = 1
    n0 = int(iocc.sum())
    akd = 0.4
    cx = lambda s, *sh: s * (rng.normal(size=sh) + 1j * rng.normal(size=sh))
    adia = 2.0 + 0.3 * rng.normal(size=3 * n0) + 1j * (0.05 + 0.2 * rng.random(3 * n0)) * -1.0
    aoff = cx(0.2, 3 * n0) if offdiag else np.zeros(3 * n0, complex)
    e = cx(1.0, 3 * n0)
    Ao = ADJ.compact_dense_operator(O.dense_operator, nx, ny, nz, akd, adia, aoff, iocc)
    return dict(nx=nx, ny=ny, nz=nz, iocc=iocc, akd=akd, adia=adia, aoff=aoff, e=e, Ao=Ao)


def fd(P, which, idx, dz, h=1e-6):
    def cab(delta):
        ad, ao = P["adia"].copy(), P["aoff"].copy()
        (ad if which == "adia" else ao)[idx] += delta * dz
        Ao = ADJ.compact_dense_operator(O.dense_operator, P["nx"], P["ny"], P["nz"], P["akd"], ad, ao, P["iocc"])
        return ADJ.cabs_dense(Ao, P["e"], ad, ao, P["akd"])[0]
    return (cab(h) - cab(-h)) / (2 * h)


def test_adjoint_gradient_matches_central_differences():
    P = small_problem(1)
    c0, p = ADJ.cabs_dense(P["Ao"], P["e"], P["adia"], P["aoff"], P["akd"])
    ga, go = ADJ.cabs_adjoint(P["Ao"], p, P["adia"], P["aoff"], P["akd"])
    rng = np.random.default_rng(5)
    for which, g in (("adia", ga), ("aoff", go)):
        for idx in rng.choice(g.size, 6, replace=False):
            for dz in (1.0, 1j):
                num = fd(P, which, idx, dz)
                ana = np.real(g[idx] * dz)
                assert abs(num - ana) <= 1e-6 * max(1.0, abs(c0)), (which, idx, dz, num, ana)


def test_adjoint_catches_dropped_terms():
    """Without the implicit (mu) term or the explicit term the check fails."""
    P = small_problem(2)
    _, p = ADJ.cabs_dense(P["Ao"], P["e"], P["adia"], P["aoff"], P["akd"])
    ga, _ = ADJ.cabs_adjoint(P["Ao"], p, P["adia"], P["aoff"], P["akd"])
    cf = 4.0 * np.pi * P["akd"]
    explicit = 1j * cf * np.abs(p) ** 2
    idx = int(np.argmax(np.abs(ga - explicit)))
    num = fd(P, "adia", idx, 1.0)
    assert abs(num - np.real(ga[idx])) < 1e-6 * max(1.0, abs(num))
    assert abs(num - np.real(explicit[idx])) > 1e-3 * max(1.0, abs(num))
    assert abs(num - np.real(ga[idx] - explicit[idx])) > 1e-3 * max(1.0, abs(num)) or abs(np.real(explicit[idx])) < 1e-9


def test_cabs_of_matches_evalq_formula():
    P = small_problem(3)
    _, p = ADJ.cabs_dense(P["Ao"], P["e"], P["adia"], P["aoff"], P["akd"])
    cabs, _, _ = O.evalq(P["e"], p, P["adia"], P["aoff"], (P["akd"], 0.0, 0.0), 1.0)
    assert np.isclose(ADJ.cabs_of(p, P["adia"], P["aoff"], P["akd"]), cabs, rtol=1e-6)
