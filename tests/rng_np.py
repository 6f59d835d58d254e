"""Independent numpy restatement of the library's counter-based generator
(lapack_amd/csrc/rng.cpp), used to pin it bit-for-bit.  Same algorithm:
splitmix64 mixing of (seed, index, attempt), 53-bit uniforms, Marsaglia's
polar method with an arithmetic-only logarithm."""
import numpy as np

M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def _mix64(z):
    z = (z + np.uint64(0x9E3779B97F4A7C15)) & M64
    z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & M64
    z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & M64
    return z ^ (z >> np.uint64(31))


def _draw(key, idx, attempt):
    with np.errstate(over="ignore"):
        return _mix64(key ^ _mix64(idx * np.uint64(0xD1B54A32D192ED03) + np.uint64(attempt)))


def _u01(r):
    return (r >> np.uint64(11)).astype(np.float64) * 2.0 ** -53


def _det_log(s):
    m, e = np.frexp(s)
    small = m < 0.70710678118654752440
    m = np.where(small, m * 2.0, m)
    e = np.where(small, e - 1, e).astype(np.float64)
    t = (m - 1.0) / (m + 1.0)
    t2 = t * t
    p = np.full_like(t, 1.0 / 21.0)
    for k in (19, 17, 15, 13, 11, 9, 7, 5, 3):
        p = p * t2 + 1.0 / k
    p = p * t2 + 1.0
    lnm = 2.0 * t * p
    ln2_hi = float.fromhex("0x1.62e42fefa3800p-1")
    ln2_lo = float.fromhex("0x1.ef35793c76730p-45")
    return e * ln2_hi + (lnm + e * ln2_lo)


def dlarnv(idist, seed, n, offset=0):
    with np.errstate(over="ignore"):
        key = _mix64(np.uint64(seed))
        idx = np.arange(n, dtype=np.uint64) + np.uint64(offset)
        if idist == 1:
            return _u01(_draw(key, idx, 0))
        if idist == 2:
            return 2.0 * _u01(_draw(key, idx, 0)) - 1.0
        out = np.empty(n)
        todo = np.ones(n, dtype=bool)
        a = 0
        while todo.any():
            ii = idx[todo]
            u = 2.0 * _u01(_draw(key, ii, a)) - 1.0
            v = 2.0 * _u01(_draw(key, ii, a + 1)) - 1.0
            s = u * u + v * v
            ok = (s > 0.0) & (s < 1.0)
            z = np.empty(len(ii))
            z[ok] = u[ok] * np.sqrt((-2.0 * _det_log(s[ok])) / s[ok])
            pos = np.nonzero(todo)[0]
            out[pos[ok]] = z[ok]
            todo[pos[ok]] = False
            a += 2
        return out
