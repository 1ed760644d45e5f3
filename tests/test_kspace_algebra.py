"""The conjugate-pair form of the reciprocal-space sums that the structure-
factor and force kernels evaluate on the matrix cores (pair_forces.hip),
checked in numpy against the direct sums over every m_z on random phases.

With Z(m) = c_m + i s_m = e^{i m theta_z} and Z(-m) = Z(m)*:
  S(k_xy, +-m) = X_c +- i X_s,  X_c = sum_beads A c_m,  X_s = sum_beads A s_m;
  with W = coef conj(S), P_m = W(+m) + W(-m), Q_m = W(+m) - W(-m)
  (P_0 = W(0), Q_0 = 0):
  U = sum_{m_z} Z W = sum_{m >= 0} (c_m P_m + i s_m Q_m),
  V = sum_{m_z} m_z Z W = sum_{m >= 0} m (c_m Q_m + i s_m P_m).
A dropped term, a swapped sign or a swapped P/Q fails one of the checks.
"""
import numpy as np


def test_conjugate_pair_sums():
    rng = np.random.default_rng(5)
    nbead, ncol, mz = 37, 11, 9
    ms = np.arange(-mz, mz + 1)
    tz = rng.uniform(0, 2 * np.pi, nbead)
    A = np.exp(1j * rng.uniform(0, 2 * np.pi, (ncol, nbead)))        # e^{i(mx tx + my ty)}
    Zfull = np.exp(1j * np.outer(tz, ms))                             # [bead][m_z]
    coef = rng.uniform(0.1, 1.0, (ncol, ms.size))
    coef[0, ms <= 0] = 0.0                                            # column (0, 0): m_z > 0 only
    # direct sums over every m_z
    S = A @ Zfull                                                     # [col][m_z]
    W = coef * np.conj(S)
    U = Zfull @ W.T                                                   # [bead][col]
    V = (Zfull * ms) @ W.T
    # conjugate-pair form over m >= 0
    m0 = np.arange(0, mz + 1)
    c, s = np.cos(np.outer(tz, m0)), np.sin(np.outer(tz, m0))
    Xc, Xs = A @ c, A @ s
    Sp, Sm = Xc + 1j * Xs, Xc - 1j * Xs
    ip, im = mz + m0, mz - m0                                          # columns of +m and -m in ms
    np.testing.assert_allclose(Sp, S[:, ip], rtol=0, atol=1e-12)
    np.testing.assert_allclose(Sm, S[:, im], rtol=0, atol=1e-12)
    Wp, Wm = coef[:, ip] * np.conj(Sp), coef[:, im] * np.conj(Sm)
    Pm, Qm = Wp + Wm, Wp - Wm
    Pm[:, 0], Qm[:, 0] = Wp[:, 0], 0.0
    U2 = c @ Pm.T + 1j * (s @ Qm.T)
    V2 = (c * m0) @ Qm.T + 1j * ((s * m0) @ Pm.T)
    np.testing.assert_allclose(U2, U, rtol=0, atol=1e-11)
    np.testing.assert_allclose(V2, V, rtol=0, atol=1e-10)
    # the real parts the kernels accumulate: Re U = c.Pr - s.Qi, Im U = c.Pi + s.Qr,
    # Re V = m (c.Qr - s.Pi), Im V = m (c.Qi + s.Pr)
    np.testing.assert_allclose(c @ Pm.real.T - s @ Qm.imag.T, U.real, rtol=0, atol=1e-11)
    np.testing.assert_allclose(c @ Pm.imag.T + s @ Qm.real.T, U.imag, rtol=0, atol=1e-11)
    np.testing.assert_allclose((c * m0) @ Qm.real.T - (s * m0) @ Pm.imag.T, V.real, rtol=0, atol=1e-10)
    np.testing.assert_allclose((c * m0) @ Qm.imag.T + (s * m0) @ Pm.real.T, V.imag, rtol=0, atol=1e-10)
