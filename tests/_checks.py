"""Shared checks: CSR comparison against the oracle and global invariants.

Tolerance (stated once, used by every parity test): row_ptr and col_ind
must be bit-identical to the oracle's; every value must satisfy
    |gpu - oracle| <= 1e-12 * absval
where absval is the oracle's sum over elements of the absolute elementary
products that make up the entry.  For an entry without cancellation absval
equals |oracle| and this is 1e-12 relative per entry; for an entry that is
a cancelling sum it is the textbook bound on reordered summation.
"""
import numpy as np

VAL_RTOL = 1e-12


def compare_csr(gpu, ref, rtol=VAL_RTOL):
    """gpu = (rp, ci, val); ref = (rp, ci, val, absval). Returns metrics."""
    grp, gci, gv = gpu
    rrp, rci, rv, ra = ref
    assert grp.shape == rrp.shape, (grp.shape, rrp.shape)
    assert np.array_equal(grp, rrp), "row_ptr differs from the oracle"
    assert np.array_equal(gci, rci), "col_ind differs from the oracle"
    assert gv.shape == rv.shape
    row_max = _row_max_per_entry(np.asarray(rrp, np.int64), np.abs(rv))
    # NaN entries (the blocks of a zero-area triangle: the reference's CST
    # formula has no determinant clamp) must sit at the oracle's positions.
    nan = np.isnan(rv)
    assert np.array_equal(np.isnan(gv), nan), "NaN values differ from the oracle's"
    if nan.any():
        gv, rv, ra, row_max = gv[~nan], rv[~nan], ra[~nan], row_max[~nan]
    err = np.abs(gv - rv)
    bound = rtol * ra
    bad = np.nonzero(err > bound)[0]
    assert bad.size == 0, (f"{bad.size} values outside 1e-12*absval; worst at {bad[np.argmax(err[bad] / np.maximum(bound[bad], 1e-300))]}")
    nz = rv != 0
    rel = err[nz] / np.abs(rv[nz]) if nz.any() else np.zeros(1)
    fro = np.linalg.norm(gv - rv) / max(np.linalg.norm(rv), 1e-300)
    # Entries past 1e-12 relative (allowed by the bound above only where the
    # entry is a cancelling sum): how many, and how small they are against
    # the largest entry of their row.
    over = err > rtol * np.abs(rv)
    n_over = int(over.sum())
    small = float((np.abs(rv[over]) / np.maximum(row_max[over], 1e-300)).max()) if n_over else 0.0
    return {"max_rel": float(rel.max()) if rel.size else 0.0, "frobenius_rel": float(fro),
            "max_err_over_bound": float((err / np.maximum(ra, 1e-300)).max()) if err.size else 0.0,
            "n_over_rel": n_over, "frac_over_rel": n_over / max(gv.size, 1),
            "over_rel_max_of_row_frac": small}


def _row_max_per_entry(rp, a):
    """max |a| of each entry's row, per entry (NaN ignored)."""
    counts = np.diff(rp)
    rmax = np.zeros(counts.size)
    ne = counts > 0
    if ne.any():
        rmax[ne] = np.fmax.reduceat(a, rp[:-1][ne])
    return np.repeat(rmax, counts)


def csr_matvec(rp, ci, v, x):
    rows = np.repeat(np.arange(rp.size - 1), np.diff(rp))
    return np.bincount(rows, weights=v * x[ci], minlength=rp.size - 1)


def csr_symmetry_error(rp, ci, v):
    """max |K - K^T| / max |K| computed on the CSR (pattern must be symmetric)."""
    n = rp.size - 1
    rows = np.repeat(np.arange(n, dtype=np.int64), np.diff(rp))
    key = rows * n + ci
    tkey = ci.astype(np.int64) * n + rows
    order = np.argsort(key)
    pos = np.searchsorted(key[order], tkey)
    pos = np.minimum(pos, key.size - 1)
    found = key[order][pos] == tkey
    assert found.all(), "pattern is not symmetric"
    vt = v[order][pos]
    return float(np.abs(v - vt).max() / max(np.abs(v).max(), 1e-300))


def rigid_modes(coords, dim):
    """Rigid-body displacement fields (dim=2: 3 modes, dim=3: 6 modes)."""
    X = coords.reshape(-1, dim)
    n = X.shape[0]
    modes = []
    for d in range(dim):
        u = np.zeros((n, dim)); u[:, d] = 1.0
        modes.append(u.reshape(-1))
    rots = [(0, 1)] if dim == 2 else [(0, 1), (1, 2), (0, 2)]
    for a, b in rots:
        u = np.zeros((n, dim)); u[:, a] = -X[:, b]; u[:, b] = X[:, a]
        modes.append(u.reshape(-1))
    return modes


def uniform_strain_field(coords, dim, eps):
    """u = eps_sym * x for a symmetric strain tensor eps (dim x dim)."""
    X = coords.reshape(-1, dim)
    return (X @ np.asarray(eps).T).reshape(-1)


def plane_stress_D(E, nu):
    f = E / (1 - nu * nu)
    return f * np.array([[1, nu, 0], [nu, 1, 0], [0, 0, (1 - nu) / 2]])


def solid_D(E, nu):
    lam = E * nu / ((1 + nu) * (1 - 2 * nu)); mu = E / (2 * (1 + nu))
    D = np.zeros((6, 6))
    D[:3, :3] = lam
    D[[0, 1, 2], [0, 1, 2]] += 2 * mu
    D[[3, 4, 5], [3, 4, 5]] = mu
    return D


def element_areas_2d(mesh):
    """Shoelace area of every Q4/T3 element of a 2D mesh."""
    X = mesh.coords.reshape(-1, 2)
    conn = mesh.conn.reshape(-1, mesh.npe)
    area = np.zeros(conn.shape[0])
    nn = np.where(conn[:, -1] >= 0, conn.shape[1], 3) if mesh.npe == 4 else np.full(conn.shape[0], 3)
    for i in range(conn.shape[1]):
        j = (i + 1) % conn.shape[1]
        a = conn[:, i]
        b = np.where((j < nn) & (conn[:, j % conn.shape[1]] >= 0), conn[:, j], conn[:, 0])
        valid = i < nn
        area += np.where(valid, X[a, 0] * X[b, 1] - X[b, 0] * X[a, 1], 0.0)
    return 0.5 * area


# ------------------------------------------------------------- Hex20 ---
HEX_NAT = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1], [-1, -1, 1], [1, -1, 1], [1, 1, 1],
                    [-1, 1, 1], [0, -1, -1], [1, 0, -1], [0, 1, -1], [-1, 0, -1], [0, -1, 1], [1, 0, 1],
                    [0, 1, 1], [-1, 0, 1], [-1, -1, 0], [1, -1, 0], [1, 1, 0], [-1, 1, 0]], float)


def hex20_dshape_numpy(p):
    """Serendipity derivatives from the product form (written independently)."""
    xi, eta, ze = p
    d = np.zeros((3, 20))
    for n, (a, b, c) in enumerate(HEX_NAT):
        if a != 0 and b != 0 and c != 0:
            fa, fb, fc = 1 + a * xi, 1 + b * eta, 1 + c * ze
            s = a * xi + b * eta + c * ze - 2
            d[0, n] = 0.125 * (a * fb * fc * s + fa * fb * fc * a)
            d[1, n] = 0.125 * (b * fa * fc * s + fa * fb * fc * b)
            d[2, n] = 0.125 * (c * fa * fb * s + fa * fb * fc * c)
        else:
            # 0.25 * (1 - q^2) * (1 + r r_n) * (1 + s s_n) with q the zero coordinate
            v = np.array([xi, eta, ze]); nat = np.array([a, b, c])
            z = int(np.nonzero(nat == 0)[0][0])
            f = [1 - v[z] ** 2 if k == z else 1 + nat[k] * v[k] for k in range(3)]
            df = [-2 * v[z] if k == z else nat[k] for k in range(3)]
            for k in range(3):
                prod = 0.25 * df[k]
                for l in range(3):
                    if l != k:
                        prod *= f[l]
                d[k, n] = prod
    return d



def hex20_gauss_volumes(coords, conn, ngp=3):
    """Sum over the ngp^3 Gauss points of w * det J for every Hex20 element
    (numpy, vectorised over elements): the volume a Gauss-integrated element
    sees, so that sum_e V_e eps^T D eps is the exact energy of a uniform
    strain on the assembled matrix."""
    X = np.asarray(coords, float).reshape(-1, 3)
    c = np.asarray(conn).reshape(-1, 20)
    g, w = np.polynomial.legendre.leggauss(ngp)
    vol = np.zeros(c.shape[0])
    Xe = X[c]                                                  # elements x 20 x 3
    for a, wa in zip(g, w):
        for b, wb in zip(g, w):
            for cc, wc in zip(g, w):
                dN = hex20_dshape_numpy((a, b, cc))            # 3 x 20
                J = np.einsum("kn,end->ekd", dN, Xe)           # elements x 3 x 3
                vol += np.linalg.det(J) * wa * wb * wc
    return vol
