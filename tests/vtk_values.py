"""Adversarial reals for the "%.6f" formatting checks: exact ties at the
sixth decimal (k/128), values that carry into the integer part, signed zeros
and values that round to zero, denormals, integers beyond 2^64 and the
largest finite values, random magnitudes and random bit patterns."""

import numpy as np

SPECIAL = [0.0, -0.0, 1 / 128, 3 / 128, -5 / 128, 0.5, 0.9999995, -0.9999995, 999999.9999996,
           5e-7, -5e-7, 4.9e-7, 1e-7, -1e-7, 2.0 ** 63, 2.0 ** 64, 2.0 ** 64 - 2.0 ** 40,
           2.0 ** 100, 1e30, 123456.789, -9.81, 1000.0, 0.025, 7.0, 1e-30]


def adversarial(n, dtype, seed, nonfinite=False):
    dt = np.dtype(dtype)
    rng = np.random.default_rng(seed)
    fi = np.finfo(dt)
    special = list(SPECIAL) + [float(fi.max), -float(fi.max), float(fi.smallest_subnormal),
                               -float(fi.smallest_subnormal), float(fi.tiny)]
    if dt == np.float64:
        special += [1e300, -1e300, 2.0 ** 200 + 2.0 ** 150, 5e-324, 1e22, 1e23]
    k = rng.integers(-2000, 2000, 64)
    special += list(k / 128.0)  # ties and near-ties
    hi = 30 if dt == np.float32 else 300
    mags = 10.0 ** rng.uniform(-12, hi, n)
    rnd = np.where(rng.random(n) < 0.5, -mags, mags)
    bits = rng.integers(0, 2 ** (8 * dt.itemsize), n, dtype=np.uint64 if dt == np.float64
                        else np.uint32, endpoint=False).view(dt)
    bits = bits[np.isfinite(bits)]
    v = np.concatenate([np.array(special, np.float64).astype(dt), rnd.astype(dt), bits])
    if nonfinite:
        nan = np.array([np.nan], dt)
        v = np.concatenate([v, nan, -nan, np.array([np.inf, -np.inf], dt)])
    v = v[:n] if len(v) >= n else np.resize(v, n)
    return np.ascontiguousarray(v.astype(dt))
