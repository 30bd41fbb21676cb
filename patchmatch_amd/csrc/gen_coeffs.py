"""Polynomial coefficients of the elementary functions in pmp_math.h.

Each polynomial is fitted by weighted least squares on Chebyshev nodes and
refined towards minimax by Lawson's iteration (iteratively reweighted least
squares), in double precision, then rounded to float:

  asin(a) = a + a^3 P(a^2),  a in [0, 1/2]        P degree 4 (rel. weight a^2)
  asin(s) = s + s^3 P(s^2),  s in [0, 1/sqrt 2]   P degree 5 (rel. weight s^2;
                                                   the longitude's octant angle)
  atan(q) = q P(q^2),        q in [0, 1]          P degree 8
  exp(r)  = P(r),            r in [-ln2/2, ln2/2] P degree 6
  log1p(f) = f P(f),         f in [sqrt(1/2)-1, sqrt(2)-1]  P degree 9
  sin(pi f/2) = f S(f^2), cos(pi f/2) = C(f^2), f in [0, 1]  S, C degree 5

    python patchmatch_amd/csrc/gen_coeffs.py      # prints the tables
"""
import numpy as np


def _nodes(a, b, n):
    k = np.arange(n)
    return 0.5 * (a + b) + 0.5 * (b - a) * np.cos((2 * k + 1) * np.pi / (2 * n))


def _fit(f, a, b, deg, weight=None, n=4000, iters=60):
    x = _nodes(a, b, n)
    A = np.stack([x ** k for k in range(deg + 1)], 1)
    y = f(x)
    W = np.ones_like(x) if weight is None else weight(x)
    ww = np.ones_like(x)
    c = None
    for _ in range(iters):
        s = W * np.sqrt(ww)
        c, *_ = np.linalg.lstsq(A * s[:, None], y * s, rcond=None)
        e = np.abs((A @ c - y) * W)
        ww = ww * e
        ww /= ww.sum()
    return c


def _safe(fn, small):
    def g(z):
        out = np.empty_like(z)
        m = z > 1e-12
        out[m] = fn(z[m])
        out[~m] = small(z[~m])
        return out
    return g


def coefficients():
    """Dict name -> float32 coefficients, lowest degree first."""
    c = {}
    c["asin"] = _fit(_safe(lambda z: (np.arcsin(np.sqrt(z)) - np.sqrt(z)) / (z * np.sqrt(z)),
                           lambda z: 1 / 6 + 3 / 40 * z), 0, 0.25, 4, weight=lambda z: z)
    c["asin_oct"] = _fit(_safe(lambda z: (np.arcsin(np.sqrt(z)) - np.sqrt(z)) / (z * np.sqrt(z)),
                               lambda z: 1 / 6 + 3 / 40 * z), 0, 0.5, 5, weight=lambda z: z)
    c["atan"] = _fit(_safe(lambda z: np.arctan(np.sqrt(z)) / np.sqrt(z), lambda z: 1 + 0 * z), 0, 1, 8)
    c["exp"] = _fit(np.exp, -np.log(2) / 2, np.log(2) / 2, 6)
    lo, hi = np.sqrt(0.5) - 1, np.sqrt(2) - 1
    c["log"] = _fit(lambda f: np.where(np.abs(f) > 1e-12, np.log1p(f) / np.where(f == 0, 1, f), 1.0), lo, hi, 9)
    c["sin"] = _fit(_safe(lambda z: np.sin(np.pi / 2 * np.sqrt(z)) / np.sqrt(z), lambda z: np.pi / 2 + 0 * z), 0, 1, 5)
    c["cos"] = _fit(lambda z: np.cos(np.pi / 2 * np.sqrt(z)), 0, 1, 5)
    return {k: np.float32(v) for k, v in c.items()}


if __name__ == "__main__":
    for name, v in coefficients().items():
        print(f"{name:8s}", ", ".join("%.9ef" % x for x in v))
