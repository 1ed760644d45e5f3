"""The CPU restatement of the Ewald surrogate (tests/sur_ref.py) pinned by
what fixes it independently: the moment-tensor angular features equal the
explicit sums over neighbour pairs of w w (1 +- cos theta)^zeta, the
analytic forces equal minus the central finite differences of the summed
per-particle energies, and the analytic training gradient equals central
finite differences of the loss."""
import numpy as np

import sur_ref as R


def config(N=24, L=9.0, seed=3):
    rng = np.random.default_rng(seed)
    return rng.uniform(0, L, size=(N, 3)), L


def random_model(seed=5, scale=0.3):
    rng = np.random.default_rng(seed)
    par = rng.normal(size=R.NPAR) * scale
    norm = (rng.normal(size=R.ND) * 0.1, np.exp(rng.normal(size=R.ND) * 0.2) * 0.5, 0.3, 0.7)
    return par, norm


def test_angular_features_are_triplet_sums():
    x, L = config()
    D, _ = R.descriptors(x, L, 4.0)
    bf = R.angular_bruteforce(x, L, 4.0)
    np.testing.assert_allclose(D[:, R.NR:], bf, rtol=1e-12, atol=1e-12)


def test_forces_are_minus_gradient():
    x, L = config()
    par, norm = random_model()
    rc = 4.0
    F = R.forces(x, L, rc, par, norm)
    h = 1e-6
    for b in (0, 7, 19):
        for d in range(3):
            xp, xm = x.copy(), x.copy()
            xp[b, d] += h
            xm[b, d] -= h
            fd = -(R.energies(xp, L, rc, par, norm).sum() - R.energies(xm, L, rc, par, norm).sum()) / (2 * h)
            assert abs(F[b, d] - fd) < 1e-6 * max(1.0, abs(fd)), (b, d, F[b, d], fd)
    # translation invariance: the forces of a slice sum to zero
    assert np.abs(F.sum(0)).max() < 1e-10 * np.abs(F).max()


def test_training_gradient_matches_finite_differences():
    rng = np.random.default_rng(9)
    Dn = rng.normal(size=(40, R.ND))
    t = rng.normal(size=40)
    par, _ = random_model(seed=11, scale=0.1)
    loss, g = R.loss_and_gradient(Dn, t, par)
    h = 1e-6
    for k in list(rng.choice(R.NPAR, 25, replace=False)) + [R.NPAR - 1, R.ND * R.H]:
        pp, pm = par.copy(), par.copy()
        pp[k] += h
        pm[k] -= h
        fd = (R.loss_and_gradient(Dn, t, pp)[0] - R.loss_and_gradient(Dn, t, pm)[0]) / (2 * h)
        assert abs(g[k] - fd) < 1e-6 * max(1.0, abs(fd)), (k, g[k], fd)
