"""Comparison helpers shared by the parity tests."""
import numpy as np


def singular_values(PCs):
    """Column norms of the scores U_k S_k, i.e. the singular values."""
    return np.linalg.norm(np.asarray(PCs, np.float64), axis=0)


def relative_sv_error(PCs, reference_sv):
    sv = singular_values(PCs)
    ref = np.asarray(reference_sv, np.float64)
    return np.abs(sv - ref) / ref


def well_separated(sv, rel_gap):
    """Indices whose singular value is at least rel_gap from its neighbours."""
    sv = np.asarray(sv, np.float64)
    out = []
    for i in range(len(sv)):
        lo = abs(sv[i] - sv[i - 1]) if i > 0 else np.inf
        hi = abs(sv[i] - sv[i + 1]) if i + 1 < len(sv) else np.inf
        if min(lo, hi) > rel_gap * sv[i]:
            out.append(i)
    return np.array(out, dtype=int)


def column_errors_up_to_sign(A, B):
    A = np.asarray(A, np.float64)
    B = np.asarray(B, np.float64)
    sign = np.sign(np.sum(A * B, axis=0))
    sign[sign == 0] = 1
    return np.linalg.norm(A * sign - B, axis=0) / np.linalg.norm(B, axis=0)


def cluster_projection_deficit(A, B, sv_ref, rel_window):
    """1 - |projection| of each normalised column of A onto the span of the
    columns of B whose singular values lie within rel_window of its own.

    Columns whose cluster touches the last column are skipped (the cluster
    may continue past k).  Returns the deficits of the checked columns.
    """
    A = np.asarray(A, np.float64)
    B = np.asarray(B, np.float64)
    A = A / np.linalg.norm(A, axis=0)
    Q, _ = np.linalg.qr(B / np.linalg.norm(B, axis=0))
    sv_ref = np.asarray(sv_ref, np.float64)
    k = len(sv_ref)
    out = []
    for i in range(k):
        members = np.flatnonzero(np.abs(sv_ref - sv_ref[i]) <= rel_window * sv_ref[i])
        if members.max() == k - 1 and i != k - 1 and len(members) > 1:
            continue
        if members.max() == k - 1 and len(members) > 1:
            continue
        Qm, _ = np.linalg.qr(B[:, members])
        proj = np.linalg.norm(Qm.T @ A[:, i])
        out.append(1.0 - proj)
    return np.array(out)


def principal_angles(A, B):
    """Principal angles (radians, ascending) between the column spans of A
    and B (both n x k): arccos of the singular values of Qa^T Qb."""
    Qa, _ = np.linalg.qr(np.asarray(A, np.float64))
    Qb, _ = np.linalg.qr(np.asarray(B, np.float64))
    c = np.linalg.svd(Qa.T @ Qb, compute_uv=False)
    return np.sort(np.arccos(np.clip(c, -1.0, 1.0)))


def knn_points(seed, n, d, centers, spread=0.5):
    """The inputs of the neighbour-search fixtures: brisc_amd.synth's
    seeded PC-space mixture, regenerated bit for bit from the seed."""
    from brisc_amd import synth
    return synth.pc_mixture(seed, n, d, centers, spread)


def lattice_points(seed, n, d, centres, half_width=4, spread=1.5):
    """Integer-valued FP32 points: ``centres`` random lattice centres in
    [-half_width, half_width]^d, each point a centre plus rounded Gaussian
    noise.  Squared distances, norms and dot products of such points are
    exact in FP32 in any summation order while their sums stay below 2^24,
    so k-means initialisations on them can be compared bit for bit."""
    rng = np.random.default_rng(seed)
    cen = rng.integers(-half_width, half_width + 1, (centres, d))
    X = cen[rng.integers(0, centres, n)] + \
        np.round(rng.standard_normal((n, d)) * spread)
    return np.ascontiguousarray(X.astype(np.float32))


def knn_sets_agree(X, nb_a, nb_b, rtol, rows=None):
    """Whether two neighbour lists of the same points (row i answering query
    ``rows[i]``, default i) differ only by near-ties: every neighbour in one
    list and not the other must lie within the slack
    ``2 rtol (||q||^2 + max ||c||^2)`` of the k-th exact (FP64) distance of
    both lists.  Returns (ok, fraction of identical rows, fraction of
    identical sets)."""
    X64 = np.asarray(X, np.float64)
    norms = (X64 * X64).sum(1)
    a = np.asarray(nb_a, np.int64)
    b = np.asarray(nb_b, np.int64)
    qrows = np.arange(len(a)) if rows is None else np.asarray(rows, np.int64)
    slack = 2.0 * rtol * (norms + norms.max())
    ok = True
    for i in np.flatnonzero((np.sort(a, 1) != np.sort(b, 1)).any(1)):
        q = qrows[i]
        da = ((X64[a[i]] - X64[q]) ** 2).sum(1)
        db = ((X64[b[i]] - X64[q]) ** 2).sum(1)
        kth = min(da.max(), db.max())
        diff = set(a[i].tolist()) ^ set(b[i].tolist())
        dd = ((X64[sorted(diff)] - X64[q]) ** 2).sum(1)
        if (np.abs(dd - kth) > slack[q]).any():
            ok = False
    same_rows = float((a == b).all(1).mean())
    same_sets = float((np.sort(a, 1) == np.sort(b, 1)).all(1).mean())
    return ok, same_rows, same_sets
