"""ORACLE — test infrastructure only. The Mie series of a perfectly conducting
sphere (textbook; Bohren & Huffman, "Absorption and Scattering of Light by
Small Particles", 1983, sec. 4.4, with the conductor limit of the
coefficients) and the radiated far field of RWG currents, for the end-to-end
check the survey names for the element's sign and normalisation: solve
Z J = -V for a plane wave on a PEC sphere and compare the bistatic radar
cross section with the series (the reference's solve J = -Z \\ Ei at
codes/GPU_implementation/seed/compres_single_block_spheres.jl:3773).

Pins (tests/test_oracle.py): the Rayleigh limit sigma_back = 9 pi a^2 (ka)^4
and the optical limit sigma_back -> pi a^2, and the optical theorem for the
sum of the series."""
import numpy as np
from scipy import special


def _coefficients(x, nmax):
    n = np.arange(1, nmax + 1)
    jn, jnp = special.spherical_jn(n, x), special.spherical_jn(n, x, derivative=True)
    yn, ynp = special.spherical_yn(n, x), special.spherical_yn(n, x, derivative=True)
    hn, hnp = jn + 1j * yn, jnp + 1j * ynp
    psi, dpsi = x * jn, jn + x * jnp  # Riccati-Bessel psi_n = x j_n, xi_n = x h_n^(1)
    xi, dxi = x * hn, hn + x * hnp
    return n, dpsi / dxi, psi / xi  # a_n, b_n of a perfect conductor


def _pi_tau(mu, nmax):
    pi = np.zeros((nmax + 1,) + np.shape(mu))
    tau = np.zeros_like(pi)
    pi[1] = 1.0
    tau[1] = mu
    for n in range(2, nmax + 1):
        pi[n] = (2 * n - 1) / (n - 1) * mu * pi[n - 1] - n / (n - 1) * pi[n - 2]
        tau[n] = n * mu * pi[n] - (n + 1) * pi[n - 1]
    return pi[1:], tau[1:]


def amplitudes(ka, scatter_angle):
    """S1, S2 of Bohren & Huffman for a PEC sphere of size parameter ka."""
    nmax = int(ka + 4 * ka ** (1 / 3) + 10)
    n, a, b = _coefficients(ka, nmax)
    mu = np.cos(np.asarray(scatter_angle, dtype=np.float64))
    pi, tau = _pi_tau(mu, nmax)
    c = ((2 * n + 1) / (n * (n + 1)))[:, None] if mu.ndim else (2 * n + 1) / (n * (n + 1))
    aa = a[:, None] if mu.ndim else a
    bb = b[:, None] if mu.ndim else b
    S1 = np.sum(c * (aa * pi + bb * tau), axis=0)
    S2 = np.sum(c * (aa * tau + bb * pi), axis=0)
    return S1, S2


def rcs(k, a, scatter_angle, plane="E"):
    """Bistatic radar cross section (m^2) at scattering angles Theta (0 =
    forward): E-plane (incident E in the scattering plane) 4 pi |S2|^2 / k^2,
    H-plane 4 pi |S1|^2 / k^2."""
    S1, S2 = amplitudes(k * a, scatter_angle)
    S = S2 if plane == "E" else S1
    return 4 * np.pi * np.abs(S) ** 2 / k ** 2


def extinction(k, a):
    """Extinction cross section from the series, (2 pi / k^2) sum (2n+1) Re(a_n + b_n)."""
    x = k * a
    n, an, bn = _coefficients(x, int(x + 4 * x ** (1 / 3) + 10))
    return 2 * np.pi / k ** 2 * np.sum((2 * n + 1) * np.real(an + bn))


_D7 = np.array([
    [1 / 3, 1 / 3, 1 / 3, 0.225],
    [0.059715871789770, 0.470142064105115, 0.470142064105115, 0.132394152788506],
    [0.470142064105115, 0.059715871789770, 0.470142064105115, 0.132394152788506],
    [0.470142064105115, 0.470142064105115, 0.059715871789770, 0.132394152788506],
    [0.797426985353087, 0.101286507323456, 0.101286507323456, 0.125939180544827],
    [0.101286507323456, 0.797426985353087, 0.101286507323456, 0.125939180544827],
    [0.101286507323456, 0.101286507323456, 0.797426985353087, 0.125939180544827],
])


def rwg_far_field(vertex, topol, edges, ln, J, k, rhat):
    """N(rhat) = sum_n J_n int f_n(r') e^{jk rhat.r'} dS' (7-point rule per
    triangle), f = s l/(2A)(r' - v_free); vertex 3 x nv, topol 3 x nt, edges
    4 x ne (1-based), rhat 3 x nd. Returns 3 x nd."""
    V = np.asarray(vertex, float)
    T = np.asarray(topol).astype(int) - 1
    E = np.asarray(edges).astype(int) - 1
    rhat = np.asarray(rhat, float)
    N = np.zeros((3, rhat.shape[1]), complex)
    for h, s in ((0, 1.0), (1, -1.0)):
        t = E[h]
        vf = V[:, E[2 + h]]
        P = V[:, T[:, t]]  # 3 x 3 x ne: coordinate, corner, RWG
        cr = np.cross((P[:, 1] - P[:, 0]).T, (P[:, 2] - P[:, 0]).T).T
        A = 0.5 * np.linalg.norm(cr, axis=0)
        coef = s * ln * J / (2 * A) * A  # quadrature weights carry the area
        for a, b, c, w in _D7:
            q = a * P[:, 0] + b * P[:, 1] + c * P[:, 2]  # 3 x ne
            ph = np.exp(1j * k * (rhat.T @ q))  # nd x ne
            N += ((q - vf) * (w * coef)) @ ph.T
    return N


def mom_rcs(vertex, topol, edges, ln, J, k, eta, rhat):
    """sigma = k^2 eta^2 |N_perp|^2 / (4 pi) for unit incident amplitude, the
    scattered field -j k eta e^{-jkr} / (4 pi r) N_perp of the currents J."""
    N = rwg_far_field(vertex, topol, edges, ln, J, k, rhat)
    Np = N - rhat * np.sum(N * rhat, axis=0)
    return k ** 2 * eta ** 2 * np.sum(np.abs(Np) ** 2, axis=0) / (4 * np.pi)
