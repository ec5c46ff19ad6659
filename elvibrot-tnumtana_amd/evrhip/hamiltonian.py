"""Seeded synthetic Hamiltonians of the benchmark configurations.

Counter-based generator: splitmix64, one stream per parameter family, seed
0xE7B00000 + configuration index.  Energies in cm^-1, dimensionless normal
coordinates q.

  V(q)  = sum_i w_i/2 q_i^2 + sum_{i<=j<=k} c_ijk q_i q_j q_k
          + sum_{i<j} d_ij q_i^2 q_j^2 + sum_i e_i q_i^4
  T     = -1/2 sum_ij d_i G_ij(q) d_j,  G_ij(q) = w_i delta_ij + sum_k a_ijk q_k
"""
import numpy as np

MASK = (1 << 64) - 1
FAM_OMEGA, FAM_CUBIC, FAM_QUART, FAM_GLIN, FAM_PSI, FAM_MU1, FAM_MU2, FAM_FIELD = range(8)
AU_TO_CM = 219474.6313705  # reference default (CODATA 2006)


class SplitMix64:
    def __init__(self, seed, family):
        self.state = ((seed & 0xFFFFFFFF) << 8 | (family & 0xFF)) & MASK

    def next_u64(self):
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
        return z ^ (z >> 31)

    def uniform(self, n):
        """n values uniform in [-1, 1)."""
        out = np.empty(n)
        for i in range(n):
            out[i] = (self.next_u64() >> 11) * (2.0 ** -53) * 2.0 - 1.0
        return out


def config_seed(index):
    return 0xE7B00000 + index


class Hamiltonian:
    """Coefficients of V and G for one configuration."""

    def __init__(self, D, omega, c3, d22, e4, Ga=None):
        self.D = D
        self.omega = np.asarray(omega, float)
        self.c3 = c3          # D^3, only i<=j<=k used
        self.d22 = d22        # D x D, only i<j used
        self.e4 = np.asarray(e4, float)
        self.Ga = Ga          # D^3 fully symmetric or None

    @property
    def G0(self):
        return np.diag(self.omega)

    @property
    def v2(self):
        return 0.5 * self.omega

    def potential(self, q):
        """V at points q (npts x D)."""
        D = self.D
        q = np.asarray(q, float)
        V = (q * q) @ self.v2 + (q ** 4) @ self.e4
        for i in range(D):
            for j in range(i, D):
                if j > i:
                    V += self.d22[i, j] * q[:, i] ** 2 * q[:, j] ** 2
                for k in range(j, D):
                    V += self.c3[i, j, k] * q[:, i] * q[:, j] * q[:, k]
        return V

    def Ga_fortran(self):
        """a_ijk flattened as i + D*(j + D*k) (symmetric, so order is immaterial)."""
        return None if self.Ga is None else np.ascontiguousarray(self.Ga).ravel(order="F")

    def scaled(self, f):
        """The same Hamiltonian with every coefficient times f (a change of
        energy unit, e.g. 1 / AU_TO_CM from cm^-1 to hartree)."""
        return Hamiltonian(self.D, f * self.omega, f * self.c3, f * self.d22, f * self.e4,
                           None if self.Ga is None else f * self.Ga)


def _upper_cubic(D, rng, scale):
    c3 = np.zeros((D, D, D))
    for i in range(D):
        for j in range(i, D):
            for k in range(j, D):
                c3[i, j, k] = scale * rng.uniform(1)[0]
    return c3


def _upper_pairs(D, rng, scale):
    d = np.zeros((D, D))
    for i in range(D):
        for j in range(i + 1, D):
            d[i, j] = scale * rng.uniform(1)[0]
    return d


def _sym3(D, rng, scale):
    a = np.zeros((D, D, D))
    for i in range(D):
        for j in range(i, D):
            for k in range(j, D):
                v = scale * rng.uniform(1)[0]
                for (x, y, z) in {(i, j, k), (i, k, j), (j, i, k), (j, k, i), (k, i, j), (k, j, i)}:
                    a[x, y, z] = v
    return a


def config0():
    """D=3 direct-product anharmonic oscillator (G = diag(w))."""
    seed = config_seed(0)
    D = 3
    omega = np.array([1200.0, 1500.0, 2100.0])
    c3 = _upper_cubic(D, SplitMix64(seed, FAM_CUBIC), 30.0)
    d22 = _upper_pairs(D, SplitMix64(seed, FAM_QUART), 5.0)
    return Hamiltonian(D, omega, c3, d22, np.full(D, 20.0))


def smolyak_config(index, D):
    """Configurations 1, 2, 4: w ~ U[800, 3000), same polynomial families,
    G_ij = w_i delta_ij + sum_k a_ijk q_k with a ~ U(+-0.02 min w), symmetric."""
    seed = config_seed(index)
    omega = 800.0 + 2200.0 * 0.5 * (SplitMix64(seed, FAM_OMEGA).uniform(D) + 1.0)
    c3 = _upper_cubic(D, SplitMix64(seed, FAM_CUBIC), 30.0)
    d22 = _upper_pairs(D, SplitMix64(seed, FAM_QUART), 5.0)
    Ga = _sym3(D, SplitMix64(seed, FAM_GLIN), 0.02 * omega.min())
    return Hamiltonian(D, omega, c3, d22, np.full(D, 20.0), Ga)


def control_config(D, nk=64, nf=32, index=3):
    """Optimal-control inputs of configuration 3: the dipole surface
    mu(q) = sum_i m_i q_i + sum_{i<=j} m_ij q_i q_j with m_i, m_ij uniform in
    +-0.1 au, and nf fields of nk sine coefficients a_kf uniform in +-1e-3 au
    (nk x nf).  Returns (m1, m2 upper triangle, amp)."""
    seed = config_seed(index)
    m1 = 0.1 * SplitMix64(seed, FAM_MU1).uniform(D)
    r2 = SplitMix64(seed, FAM_MU2)
    m2 = np.zeros((D, D))
    for i in range(D):
        for j in range(i, D):
            m2[i, j] = 0.1 * r2.uniform(1)[0]
    amp = 1e-3 * SplitMix64(seed, FAM_FIELD).uniform(nk * nf).reshape(nf, nk).T.copy(order="F")
    return m1, m2, amp


def dipole(q, m1, m2):
    """mu at points q (npts x D) for control_config's coefficients."""
    q = np.asarray(q, float)
    return q @ m1 + np.einsum("pi,ij,pj->p", q, m2, q)


def test_vectors(n, nvec, index=0):
    """Normalised random test vectors (n x nvec, column-major)."""
    rng = SplitMix64(config_seed(index), FAM_PSI)
    # vectorised splitmix64 for large n
    s0 = rng.state
    k = np.arange(1, n * nvec + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(s0) + k * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    u = (z >> np.uint64(11)).astype(np.float64) * (2.0 ** -53) * 2.0 - 1.0
    X = u.reshape(nvec, n).T.copy(order="F")
    X /= np.linalg.norm(X, axis=0)
    return X
