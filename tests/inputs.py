"""Seeded synthetic inputs shared by the parity tests."""
import numpy as np

SIGMA_R = 1.444 / (2.0 * np.sqrt(2.0 * np.log(2.0)))


def particles(seed, np_, nn, radius, kmax=1.3, spill=0.0):
    """Positions uniform in a ball (a fraction `spill` pushed outside the grid),
    momenta uniform in a ball."""
    rng = np.random.default_rng(seed)
    tot = np_ + nn

    def ball(r):
        out = np.empty((tot, 3))
        got = 0
        while got < tot:
            c = rng.uniform(-r, r, (2 * tot + 16, 3))
            c = c[(c ** 2).sum(1) < r * r]
            m = min(tot - got, len(c))
            out[got:got + m] = c[:m]
            got += m
        return out

    r = ball(radius)
    k = ball(kmax)
    if spill > 0:
        idx = rng.choice(tot, max(1, int(spill * tot)), replace=False)
        r[idx] *= 100.0
    st = {n: np.ascontiguousarray(a) for n, a in zip(("x", "y", "z"), r.T)}
    st.update({n: np.ascontiguousarray(a) for n, a in zip(("kx", "ky", "kz"), k.T)})
    st.update({n: rng.normal(0, 5.0, tot) for n in ("fx", "fy", "fz")})
    return st


def smooth_field(seed, n, scale=1.0, positive=True):
    """A smooth random field on an n grid (sum of a few Gaussians)."""
    rng = np.random.default_rng(seed)
    i, j, k = np.meshgrid(*[np.arange(a) for a in n], indexing="ij")
    f = np.zeros(n)
    for _ in range(4):
        c = rng.uniform(0.3, 0.7, 3) * np.array(n)
        s = rng.uniform(0.1, 0.3) * min(n)
        f += rng.uniform(0.5, 1.0) * np.exp(-((i - c[0]) ** 2 + (j - c[1]) ** 2 + (k - c[2]) ** 2) / (2 * s * s))
    if not positive:
        f -= f.mean()
    return np.ascontiguousarray((scale * f).ravel())
