"""A second, independent restatement of the reference controller with dense
numpy / scipy building blocks, used to pin the oracle's step traces for
m > 1 where no closed form exists.

It shares no code with the oracle: the Krylov basis comes from classical
Gram-Schmidt applied twice with numpy matrix products (the oracle runs MGS
twice in C), and exp(dt H) from scipy.linalg.expm (the oracle has its own
Pade).  The step rules are read from main_cpu.cpp:602-680 (main.cu:1411-1518)
a second time.  Agreement therefore checks the oracle's Arnoldi, its residual
estimate |h_{k+1,k}| |[exp(dt H_k) e1]_k| beta (main.cu:2004-2020) and its
controller together, up to rounding.
"""
import numpy as np
import scipy.linalg


def _arnoldi(A, q1, m):
    n = A.shape[0]
    Q = np.zeros((n, m + 1))
    H = np.zeros((m + 1, m))
    Q[:, 0] = q1
    k = 0
    for j in range(m):
        w = A @ Q[:, j]
        h = Q[:, :j + 1].T @ w
        w = w - Q[:, :j + 1] @ h
        h2 = Q[:, :j + 1].T @ w
        w = w - Q[:, :j + 1] @ h2
        H[:j + 1, j] = h + h2
        nrm = np.linalg.norm(w)
        H[j + 1, j] = nrm
        k = j + 1
        if not (nrm >= 1e-14):  # breakdown: norm below 1e-14 or not finite
            return Q, H, k, True
        Q[:, j + 1] = w / nrm
    return Q, H, k, False


def controller_trace(A, v, t, tol, m, max_restarts):
    """Returns (records, y, stats); records hold t, dt, err, m_used, accepts,
    rejects and the decision margin of each attempt."""
    A = np.asarray(A, dtype=np.float64)
    beta = float(np.linalg.norm(v))
    q1 = v / beta if beta > 1e-14 else np.asarray(v, dtype=np.float64)
    dt_max, dt_min = t, max(1e-12, t * 1e-6)
    dt = min(max(t / (max_restarts if max_restarts > 0 else 1), dt_min), dt_max)
    max_steps = max_restarts if max_restarts > 0 else 1000
    t_done, accepted, converged, aborted = 0.0, 0, 0, 0
    records, y = [], np.zeros_like(q1)
    while t_done < t and accepted < max_steps:
        dt = min(dt, t - t_done)
        rejects = 0
        Q, H, k, breakdown = _arnoldi(A, q1, m)
        while True:
            E = scipy.linalg.expm(dt * H[:k, :k])
            wH = E[:, 0]
            err = 0.0 if breakdown else abs(H[k, k - 1] * wH[k - 1]) * beta
            y = beta * (Q[:, :k] @ np.where(np.abs(wH) > 1e-15, wH, 0.0))
            margin = np.inf if breakdown else abs(err - tol) / tol
            if not breakdown and err <= tol:
                margin = min(margin, abs(err - 0.1 * tol) / (0.1 * tol))
            records.append({"t": t_done, "dt": dt, "err": err, "m_used": k, "accepts": accepted,
                            "rejects": rejects, "margin": margin})
            if err <= tol or breakdown:
                t_done += dt
                accepted += 1
                if t_done >= t:
                    converged = 1
                    break
                beta = float(np.linalg.norm(y))
                q1 = y / beta if beta > 1e-14 else y
                if err < 0.1 * tol:
                    dt = min(dt * 1.5 * 0.8, dt_max)
                break
            rejects += 1
            if rejects >= 20:
                aborted, t_done = 1, t
                break
            dt = max(dt * 0.5 * 0.8, dt_min)
    stats = {"converged": converged, "restarts_done": accepted, "attempts": len(records),
             "aborted": aborted}
    return records, y, stats
