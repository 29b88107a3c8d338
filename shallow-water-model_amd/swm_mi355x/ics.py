"""Seeded synthetic initial states (host, numpy) for grids up to about g = 7.

The reference has no seeded initial conditions (it runs shallow-water test
case 5, swm/source/initialize.F:330-513); the benchmark workloads use
random smooth fields instead:

* vorticity: real spherical-harmonic sum, degrees 1..L, coefficients
  N(0,1) * l^-1.5; streamfunction psi = -a^2 sum c_lm / (l(l+1)) Y_lm;
* divergence: an independent sum scaled to 10 % of the vorticity RMS, giving
  the velocity potential chi the same way;
* corner velocity u = grad(chi) + k x grad(psi) with the model's own discrete
  gradient (operators.F:592-682, :1402-1514), scaled to max |u| = umax;
* thickness h0 plus a degree 1..L harmonic sum of the given RMS;
* tracers: sums of Gaussian bumps exp(-(d/500 km)^2) at random centres.

Random numbers come from splitmix64 (counter based, reproducible anywhere);
stream seeds are s0+1 (vorticity), s0+2 (divergence), s0+3 (thickness),
s0+4 (tracer centres).
"""
import numpy as np

MASK64 = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed):
        self.state = seed & MASK64

    def next(self):
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def uniform(self):
        return (self.next() >> 11) * (1.0 / 9007199254740992.0)

    def normal(self):
        u1 = max(self.uniform(), 1e-300)
        u2 = self.uniform()
        return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)


def harmonic_sum(xyz, lmax, rng, power=-1.5, laplace_inverse=False):
    """sum_{l=1..lmax} sum_{m=-l..l} c_lm l^power Y_lm at unit vectors xyz (3, n);
    with laplace_inverse each term is divided by -l(l+1) (unit sphere).
    Y_lm are 4-pi normalised real harmonics."""
    x, y, z = xyz
    mu = np.clip(z, -1.0, 1.0)
    cphi = np.sqrt(np.maximum(0.0, 1.0 - mu * mu))
    lon = np.arctan2(y, x)
    coef = {}
    for l in range(1, lmax + 1):
        for m in range(-l, l + 1):
            coef[(l, m)] = rng.normal() * float(l) ** power
    out = np.zeros_like(mu)
    pmm = np.ones_like(mu)                       # P_0^0
    for m in range(0, lmax + 1):
        if m == 1:
            pmm = np.sqrt(3.0) * cphi
        elif m > 1:
            pmm = np.sqrt((2.0 * m + 1.0) / (2.0 * m)) * cphi * pmm
        p_lm2, p_lm1 = None, None
        cm, sm = np.cos(m * lon), np.sin(m * lon)
        for l in range(m, lmax + 1):
            if l == m:
                p = pmm
            elif l == m + 1:
                p = np.sqrt(2.0 * m + 3.0) * mu * pmm
            else:
                a = np.sqrt((2.0 * l - 1.0) * (2.0 * l + 1.0) / ((l - m) * (l + m)))
                b = np.sqrt((2.0 * l + 1.0) * (l + m - 1.0) * (l - m - 1.0) / ((l - m) * (l + m) * (2.0 * l - 3.0)))
                p = a * mu * p_lm1 - b * p_lm2
            p_lm2, p_lm1 = p_lm1, p
            if l >= 1:
                scale = -1.0 / (l * (l + 1.0)) if laplace_inverse else 1.0
                if m == 0:
                    out += coef[(l, 0)] * scale * p
                else:
                    out += scale * p * (coef[(l, m)] * cm + coef[(l, -m)] * sm)
    return out


def corner_gradient(grid_arrays, x):
    """Discrete gradient of a cell field at the two corners of every interior
    cell (operators.F:592-682); returns (2, 2, iim, jjm, 10)."""
    vw = grid_arrays["v_weights"]          # (3, 2, 2, iim, jjm, 10)
    cai = grid_arrays["corner_area_inv"]   # (2, iim, jjm, 10)
    iim, jjm = x.shape[0], x.shape[1]
    out = np.zeros((2, 2, iim, jjm, 10), order="F")
    I = slice(1, iim - 1)
    J = slice(1, jjm - 1)
    x0 = x[1:iim - 1, 1:jjm - 1]
    xw = x[0:iim - 2, 1:jjm - 1]
    xsw = x[0:iim - 2, 0:jjm - 2]
    xs = x[1:iim - 1, 0:jjm - 2]
    for sp in range(2):
        out[sp, 0, I, J] = (vw[0, sp, 0, I, J] * x0 + vw[1, sp, 0, I, J] * xw + vw[2, sp, 0, I, J] * xsw) * cai[0, I, J]
        out[sp, 1, I, J] = (vw[0, sp, 1, I, J] * x0 + vw[1, sp, 1, I, J] * xsw + vw[2, sp, 1, I, J] * xs) * cai[1, I, J]
    return out


def synthetic_state(grid, ntracer=1, lmax=20, seed_index=0, umax=40.0, h0=5000.0, h_rms=50.0,
                    nbumps=4, bump_radius=500.0e3):
    """Returns dict(h, hs, u, tracer) in the reference layout for a Grid."""
    s0 = 20261203 + 100 * seed_index
    p = grid.array("p")                       # (3, iim, jjm, 10), halo filled
    mask = grid.array("grid_center_mask")
    iim, jjm = grid.iim, grid.jjm
    xyz = p.reshape(3, -1, order="F")
    a = grid.radius
    real = mask.reshape(-1, order="F") == 1.0

    psi = (a * a) * harmonic_sum(xyz, lmax, SplitMix64(s0 + 1), laplace_inverse=True)
    zeta = harmonic_sum(xyz, lmax, SplitMix64(s0 + 1))
    chi = (a * a) * harmonic_sum(xyz, lmax, SplitMix64(s0 + 2), laplace_inverse=True)
    delta = harmonic_sum(xyz, lmax, SplitMix64(s0 + 2))
    rz = np.sqrt(np.mean(zeta[real] ** 2))
    rd = np.sqrt(np.mean(delta[real] ** 2))
    chi *= 0.1 * rz / rd
    arrays = {"v_weights": grid.array("v_weights"), "corner_area_inv": grid.array("corner_area_inv")}
    shp = (iim, jjm, 10)
    gp = corner_gradient(arrays, psi.reshape(shp, order="F"))
    gc = corner_gradient(arrays, chi.reshape(shp, order="F"))
    u = np.zeros((2, 2) + shp, order="F")
    u[0] = gc[0] - gp[1]
    u[1] = gc[1] + gp[0]
    speed = np.sqrt(u[0] ** 2 + u[1] ** 2)
    u *= umax / speed.max()

    hp = harmonic_sum(xyz, lmax, SplitMix64(s0 + 3))
    hp *= h_rms / np.sqrt(np.mean(hp[real] ** 2))
    h = np.where(real, h0 + hp, 0.0).reshape(shp, order="F")
    hs = np.zeros(shp, order="F")

    rng = SplitMix64(s0 + 4)
    tracer = np.zeros((iim, jjm, ntracer, 10), order="F")
    for q in range(ntracer):
        t = np.zeros(xyz.shape[1])
        for _ in range(nbumps):
            zc = 2.0 * rng.uniform() - 1.0
            lc = 2.0 * np.pi * rng.uniform()
            rc = np.sqrt(max(0.0, 1.0 - zc * zc))
            c = np.array([rc * np.cos(lc), rc * np.sin(lc), zc])
            d = a * np.arccos(np.clip(c @ xyz, -1.0, 1.0))
            t += np.exp(-(d / bump_radius) ** 2)
        tracer[:, :, q, :] = t.reshape(shp, order="F")
    return {"h": np.asfortranarray(h), "hs": hs, "u": u, "tracer": tracer}
