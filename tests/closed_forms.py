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
