"""Derive the polynomial coefficients of the binary32 FMA log and sin/cos
(logf_fma / gps_sincosf in gamesplat_amd/csrc/gs_math.h, or_logf_fma /
or_gps_sincosf in oracle/gs_oracle.c): the GPS point sampler's radius and
angle and the StochasticSplats fragment cutoff.

python oracle/fit_gps_math.py

Each polynomial is a minimax fit of the relative error, computed by Lawson's
algorithm (iteratively reweighted least squares) on 40001 Chebyshev nodes.
Coefficients are rounded to binary32 one at a time, lowest power first, the
remaining ones refitted after each rounding, so the rounding of the early
terms is compensated by the later ones.

  log(1 + f) = f + f^2 Q(f),  f in [sqrt(1/2) - 1, sqrt(2) - 1],
               Q(f) = -1/2 + q1 f + ... + q8 f^8
  sin r = r + r^3 S(r^2),  cos r = 1 - r^2 / 2 + r^4 C(r^2),
               |r| <= 1.02 pi / 4 (room for the rounding of the reduction),
               S and C of degree 2

tests/test_oracle_math.py pins the resulting functions against libm.
"""
import numpy as np


def lawson(A: np.ndarray, b: np.ndarray, iters: int = 300):
    u = np.full(len(b), 1.0 / len(b))
    for _ in range(iters):
        sw = np.sqrt(u)
        c, *_ = np.linalg.lstsq(A * sw[:, None], b * sw, rcond=None)
        e = np.abs(A @ c - b)
        u = u * e
        u /= u.sum()
    return c


def fit_rounded(x, target, powers, weight, fixed=()):
    """sum_k c_k x^powers[k] ~ target in the weighted max norm, with the
    coefficients rounded to float32 in turn; `fixed` are given leading ones."""
    coef = [float(np.float32(c)) for c in fixed]
    for i in range(len(coef), len(powers)):
        t = target - sum(c * x ** p for c, p in zip(coef, powers))
        A = np.stack([x ** p for p in powers[i:]], 1) * weight[:, None]
        coef.append(float(np.float32(lawson(A, t * weight)[0])))
    err = np.abs((target - sum(c * x ** p for c, p in zip(coef, powers))) * weight).max()
    return coef, err


def nodes(a: float, b: float, n: int = 40001) -> np.ndarray:
    return np.cos(np.linspace(0.0, np.pi, n)) * (b - a) / 2 + (a + b) / 2


def main():
    f = nodes(np.sqrt(0.5) - 1, np.sqrt(2.0) - 1)
    f = f[np.abs(f) > 1e-7]
    q_target = (np.log1p(f) - f) / f ** 2
    log_q, e_log = fit_rounded(f, q_target, list(range(9)), f ** 2 / np.abs(np.log1p(f)), fixed=(-0.5,))
    r = nodes(-1.02 * np.pi / 4, 1.02 * np.pi / 4)
    r = r[np.abs(r) > 1e-5]
    z = r * r
    sin_s, e_sin = fit_rounded(z, (np.sin(r) - r) / r ** 3, [0, 1, 2], np.abs(r ** 3 / np.sin(r)))
    cos_c, e_cos = fit_rounded(z, (np.cos(r) - 1 + z / 2) / z ** 2, [0, 1, 2], np.abs(z ** 2 / np.cos(r)))

    def show(name, v, e):
        print(f"{name} = {{{', '.join(f'{x!r}f' for x in v)}}};  // fit error {e / 2 ** -24:.3f} ulp")

    show("kLogFmaQ[9]", log_q, e_log)
    show("kGpsSinS[3]", sin_s, e_sin)
    show("kGpsCosC[3]", cos_c, e_cos)


if __name__ == "__main__":
    main()
