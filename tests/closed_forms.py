"""Closed-form exp(t*A)*v for the stencil workloads (independent of any Krylov code).

* 2-D Dirichlet Laplacian A = T (+) T, T = tridiag(1, -2, 1) of size N: the
  orthonormal DST-I diagonalises T, so exp(tA) v = S diag(e^{t(l_a + l_b)}) S v
  with l_k = -2 + 2 cos(pi (k+1) / (N+1)); O(n log n) through scipy.fft.dstn.
* 3-D convection-diffusion A = T_z (+) T_y (+) T_x: exp(tA) = E_z (x) E_y (x) E_x
  with three dense N x N exponentials (scipy.linalg.expm) applied as mode
  products.
"""
import numpy as np
import scipy.fft
import scipy.linalg


def laplacian_2d_expmv(N: int, t: float, v: np.ndarray) -> np.ndarray:
    V = np.asarray(v, dtype=np.float64).reshape(N, N)  # [y, x]
    lam = -2.0 + 2.0 * np.cos(np.pi * np.arange(1, N + 1) / (N + 1))
    W = scipy.fft.dstn(V, type=1, norm="ortho")
    W *= np.exp(t * (lam[:, None] + lam[None, :]))
    return scipy.fft.idstn(W, type=1, norm="ortho").ravel()


def laplacian_2d_expmv_block(N: int, t: float, V: np.ndarray) -> np.ndarray:
    """laplacian_2d_expmv of every column of V (n x k) at once, the DST-I over
    the two grid axes of an (N, N, k) array on all cores."""
    k = V.shape[1]
    W = scipy.fft.dstn(np.asarray(V, dtype=np.float64).reshape(N, N, k), type=1, norm="ortho",
                       axes=(0, 1), workers=-1)
    lam = -2.0 + 2.0 * np.cos(np.pi * np.arange(1, N + 1) / (N + 1))
    W *= np.exp(t * (lam[:, None] + lam[None, :]))[:, :, None]
    return scipy.fft.idstn(W, type=1, norm="ortho", axes=(0, 1), workers=-1).reshape(N * N, k)


def laplacian_2d_eigvec(N: int, ky: int, kx: int):
    """(eigenvector, eigenvalue) of the 2-D Laplacian for mode (ky, kx)."""
    j = np.arange(1, N + 1)
    sy = np.sqrt(2.0 / (N + 1)) * np.sin(np.pi * (ky + 1) * j / (N + 1))
    sx = np.sqrt(2.0 / (N + 1)) * np.sin(np.pi * (kx + 1) * j / (N + 1))
    lam = -4.0 + 2.0 * np.cos(np.pi * (ky + 1) / (N + 1)) + 2.0 * np.cos(np.pi * (kx + 1) / (N + 1))
    return np.outer(sy, sx).ravel(), lam


def kron3_expmv(Tz, Ty, Tx, t: float, v: np.ndarray) -> np.ndarray:
    N = Tx.shape[0]
    Ez, Ey, Ex = (scipy.linalg.expm(t * T) for T in (Tz, Ty, Tx))
    V = np.asarray(v, dtype=np.float64).reshape(N, N, N)  # [z, y, x]
    Y = V @ Ex.T                                     # along x
    Y = np.einsum("ab,zbx->zax", Ey, Y, optimize=True)  # along y
    Y = (Ez @ Y.reshape(N, N * N)).reshape(N, N, N)  # along z
    return Y.ravel()


def rel_err(y, ref) -> float:
    """||y - ref|| / ||ref|| (2-norm / Frobenius), scaled to avoid overflow."""
    y = np.asarray(y, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    s = np.max(np.abs(ref))
    if not np.isfinite(s) or s == 0.0:
        s = 1.0
    return float(np.linalg.norm((y - ref) / s) / np.linalg.norm(ref / s))


def diag2_m1_trace(lam1: float, lam2: float, a: float, b: float, t: float, tol: float,
                   max_restarts: int):
    """Step trace and result of the reference controller with m = 1 on
    A = diag(lam1, lam2), v = (a, b), from closed forms only.

    With one Krylov column, q1 = v / beta never changes direction: the lift is
    y = beta e^{dt h11} q1, so a restart only rescales beta.  With p = q1^2,
    h11 = lam1 p1 + lam2 p2 and h21 = |lam1 - lam2| sqrt(p1 p2), and the
    residual estimate |h21 [exp(dt H)e1]_1| beta of main.cu:2004-2020 is
    h21 e^{dt h11} beta.  The step rules (first step t / max_restarts, accept
    when err <= tol, grow by 1.5 * 0.8 when err < 0.1 tol, shrink by
    0.5 * 0.8 on reject, abort after 20 rejects, at most max_restarts accepted
    steps) are those of main.cu:1411-1518 / main_cpu.cpp:602-680.

    Returns (records, y, stats): records hold t, dt, err, accepts, rejects and
    the decision margin |err - threshold| / threshold of each attempt.
    """
    beta = float(np.hypot(a, b))
    q = np.array([a, b]) / beta
    p1, p2 = q[0] ** 2, q[1] ** 2
    h11 = lam1 * p1 + lam2 * p2
    h21 = abs(lam1 - lam2) * np.sqrt(p1 * p2)
    dt_max, dt_min = t, max(1e-12, t * 1e-6)
    dt = min(max(t / (max_restarts if max_restarts > 0 else 1), dt_min), dt_max)
    max_steps = max_restarts if max_restarts > 0 else 1000
    t_done, accepted, converged, aborted = 0.0, 0, 0, 0
    records, y = [], beta * q
    while t_done < t and accepted < max_steps:
        dt = min(dt, t - t_done)
        rejects = 0
        while True:
            growth = np.exp(dt * h11)
            err = abs(h21 * growth) * beta
            margin = abs(err - tol) / tol
            if err <= tol:
                margin = min(margin, abs(err - 0.1 * tol) / (0.1 * tol))
            records.append({"t": t_done, "dt": dt, "err": err, "accepts": accepted,
                            "rejects": rejects, "margin": margin})
            if err <= tol:
                t_done += dt
                accepted += 1
                beta *= growth
                y = beta * q
                if t_done >= t:
                    converged = 1
                    break
                if err < 0.1 * tol:
                    dt = min(dt * 1.5 * 0.8, dt_max)
                break
            rejects += 1
            if rejects >= 20:
                y = beta * growth * q  # the rejected lift is returned
                aborted, t_done = 1, t
                break
            dt = max(dt * 0.5 * 0.8, dt_min)
    stats = {"converged": converged, "restarts_done": accepted, "attempts": len(records),
             "aborted": aborted}
    return records, y, stats


def breakdown_trace(t: float, max_restarts: int):
    """Step trace of the reference controller when every attempt ends in a
    happy breakdown: err is 0 (main.cu:1467-1469), so every attempt is
    accepted and dt grows by 1.5 * 0.8 up to t, clipped to the time left
    (main.cu:1425-1518).  Returns a list of (t_done, dt) pairs."""
    dt_min = max(1e-12, t * 1e-6)
    dt = min(max(t / (max_restarts if max_restarts > 0 else 1), dt_min), t)
    max_steps = max_restarts if max_restarts > 0 else 1000
    t_done, out = 0.0, []
    while t_done < t and len(out) < max_steps:
        dt = min(dt, t - t_done)
        out.append((t_done, dt))
        t_done += dt
        if t_done >= t:
            break
        dt = min(dt * 1.5 * 0.8, t)
    return out
