"""CPU restatement of the two-pass Fourier stages of sht_kernels.hip
(fourier_analysis_kernel / fourier_synthesis_kernel and their generic forms),
loop for loop in float64, against the dense basis F[nlon][2 n_m - 1] that
dinosaur's RealSphericalHarmonics projects on (the oracle's Grid.f).

For every nlon of the Grid.T<N> table and a few others (odd, prime) it checks
  * the factorization nlon = n1 * n2 the plan makes (n1 the largest divisor
    <= sqrt(nlon)) and that the passes equal the dense contraction to 1e-11 (values of order 10),
  * that every LDS index the passes form stays inside the LDS size the host
    launches them with, for the latitude tile the plan picks.
The index formulas are restated from the kernel source: a bounds and algebra
check of that arithmetic, not a parity pin (the GPU tests pin the results)."""
import numpy as np
import pytest

from oracle import sht_oracle as so

THREADS, LDS_BYTES = 256, 64 * 1024
COMPILED = {(8, 8), (8, 12), (8, 16), (16, 16), (16, 20), (18, 20), (16, 32), (20, 32)}


def factor(nlon):
    n1 = max(d for d in range(1, int(nlon ** 0.5) + 1) if nlon % d == 0)
    return n1, nlon // n1


def lds_floats(fixed, kind, N, n1, n2, mm, tl):
    """analysis_lds / synthesis_lds of the host code, in floats."""
    if kind == "analysis":
        return 2 * N + 2 * n1 * (n2 // 2 + 1) * tl + (0 if fixed else N * tl)
    return 2 * N + 2 * N * tl + (0 if fixed else mm * tl)


def tiles(nlon, mm):
    n1, n2 = factor(nlon)
    fixed = (n1, n2) in COMPILED
    out = {}
    for kind in ("analysis", "synthesis"):
        tl = max(1, min(16, THREADS // max(n1, n2))) if fixed else 16
        while tl > 1 and 4 * lds_floats(fixed, kind, nlon, n1, n2, mm, tl) > LDS_BYTES:
            tl -= 1 if fixed else tl // 2
        out[kind] = tl
    return fixed, out


def analysis(x, n_m, used):
    """x [N][L] (line tile, latitude last) -> Y [2 n_m - 1][L]; `used` collects LDS indices."""
    N, L = x.shape
    n1, n2 = factor(N)
    nh = n2 // 2 + 1
    w = np.exp(-2j * np.pi * np.arange(N) / N)
    Z = np.zeros((nh, n1, L), complex)          # [k2][i1][lat]
    for i1 in range(n1):
        for k2 in range(nh):
            for i2 in range(n2):
                Z[k2, i1] += x[i1 + n1 * i2] * w[((i2 * k2) % n2) * n1]
                used.add(("tw", ((i2 * k2) % n2) * n1))
            used.add(("z", (k2 * n1 + i1)))
    Y = np.zeros((2 * n_m - 1, L))
    for k2 in range(n2):
        cj = k2 >= nh
        kz = n2 - k2 if cj else k2
        assert 0 <= kz < nh
        p = np.array([(np.conj(Z[kz, i1]) if cj else Z[kz, i1]) * w[(i1 * k2) % N] for i1 in range(n1)])
        for k1 in range(n1):
            if n2 * k1 >= n_m:
                break
            k = k2 + n2 * k1
            if k >= n_m:
                continue
            X = sum(p[i1] * w[((i1 * k1) % n1) * n2] for i1 in range(n1))
            if k == 0:
                Y[0] = X.real / np.sqrt(2 * np.pi)
            else:
                Y[2 * k - 1] = X.real / np.sqrt(np.pi)
                Y[2 * k] = -X.imag / np.sqrt(np.pi)
    return Y


def synthesis(Y, N, used):
    """Y [2 n_m - 1][L] -> x [N][L]."""
    mm, L = Y.shape
    n_m = (mm + 1) // 2
    n1, n2 = factor(N)
    w = np.exp(-2j * np.pi * np.arange(N) / N)
    W = np.zeros((n1, n2, L), complex)          # [i1][k2][lat]
    for k2 in range(n2):
        c = np.zeros((n1, L), complex)
        for k1 in range(n1):
            if n2 * k1 >= n_m:
                break
            k = k2 + n2 * k1
            if k < n_m:
                c[k1] = Y[0] / np.sqrt(2 * np.pi) if k == 0 else (Y[2 * k - 1] - 1j * Y[2 * k]) / np.sqrt(np.pi)
        for i1 in range(n1):
            s = sum(c[k1] * np.conj(w[((i1 * k1) % n1) * n2]) for k1 in range(n1))
            W[i1, k2] = s * np.conj(w[(i1 * k2) % N])
            used.add(("w", i1 * n2 + k2))
    x = np.zeros((N, L))
    for i1 in range(n1):
        for i2 in range(n2):
            x[i1 + n1 * i2] = sum((W[i1, k2] * np.conj(w[((i2 * k2) % n2) * n1])).real for k2 in range(n2))
    return x


GRIDS = sorted(so.TRUNCATIONS.values()) + [(22, 23, 66, 32), (30, 31, 97, 48), (5, 6, 9, 4), (3, 4, 5, 4)]


@pytest.mark.parametrize("grid", GRIDS)
def test_two_pass_fourier_stages(grid):
    n_m, n_l, nlon, nlat = grid
    mm = 2 * n_m - 1
    n1, n2 = factor(nlon)
    assert n1 * n2 == nlon and n1 <= n2
    fixed, tl = tiles(nlon, mm)
    assert fixed == (nlon in (64, 96, 128, 256, 320, 360, 512, 640)), (nlon, n1, n2)
    # the launch sizes fit 64 KiB of LDS and the passes' thread counts fit a workgroup
    for kind in ("analysis", "synthesis"):
        assert 4 * lds_floats(fixed, kind, nlon, n1, n2, mm, tl[kind]) <= LDS_BYTES, (kind, tl)
    if fixed:
        assert max(n1, n2) * tl["analysis"] <= THREADS
    if nlon > 400:   # the loops below are O(nlon * (n1 + n2)) in Python: algebra on the smaller grids
        return
    rng = np.random.default_rng(nlon)
    L = 3
    F = so.fourier_real_basis(n_m, nlon) if hasattr(so, "fourier_real_basis") else None
    lon = 2 * np.pi * np.arange(nlon) / nlon
    if F is None or F.shape != (nlon, mm):
        F = np.zeros((nlon, mm))
        F[:, 0] = 1 / np.sqrt(2 * np.pi)
        for m in range(1, n_m):
            F[:, 2 * m - 1] = np.cos(m * lon) / np.sqrt(np.pi)
            F[:, 2 * m] = np.sin(m * lon) / np.sqrt(np.pi)
    x = rng.standard_normal((nlon, L))
    used = set()
    np.testing.assert_allclose(analysis(x, n_m, used), F.T @ x, rtol=0, atol=1e-11)
    nh = n2 // 2 + 1
    assert all(i < nlon for kind, i in used if kind == "tw")
    assert all(i < nh * n1 for kind, i in used if kind == "z")
    y = rng.standard_normal((mm, L))
    used = set()
    np.testing.assert_allclose(synthesis(y, nlon, used), F @ y, rtol=0, atol=1e-11)
    assert all(i < nlon for kind, i in used if kind == "w")
