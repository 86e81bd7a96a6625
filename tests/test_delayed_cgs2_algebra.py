"""The step algebra of the delayed CGS2 Krylov-Schur expansion (krylov.hip,
lsqed_eps_solve on large sectors), restated in numpy on a random symmetric
matrix: with c = Q^T u, h = Q^T A u, nu = sqrt(u.u - c.c), phi = f.c,

    v_j = (u - Q c)/nu,  alpha_j = (u.Au - c.h)/nu^2 - phi,
    r   = Au/nu - ((phi + alpha_j)/nu) u - Q (h - (phi + alpha_j) c)/nu,

the basis stays orthonormal, A V = V T + beta v_m e_m^T holds, and T equals
the tridiagonal matrix of a fully reorthogonalised Lanczos run from the same
start vector -- also after a thick restart (f = the restart's arrow)."""
import numpy as np


def delayed_expand(A, V, T, k, m, arrow):
    beta = np.zeros(m + 1)
    for j in range(k, m):
        u = V[:, j].copy()
        y = A @ u
        Q = V[:, :j]
        c, h = Q.T @ u, Q.T @ y
        gam, dl = u @ u, u @ y
        nu2 = gam - c @ c
        nu = np.sqrt(nu2)
        assert nu >= 0.7071067811865476 * np.sqrt(gam)
        phi = c[j - 1] if j > k else float(arrow @ c[:k]) if k else 0.0
        alpha = (dl - c @ h) / nu2 - phi
        z = (h - (phi + alpha) * c) / nu
        V[:, j] = (u - Q @ c) / nu
        V[:, j + 1] = y / nu - ((phi + alpha) / nu) * u - Q @ z
        T[j, j] = alpha
        if j == k:
            T[:k, k] = T[k, :k] = arrow
        else:
            T[j, j - 1] = T[j - 1, j] = nu
            beta[j - 1] = nu
    u = V[:, m].copy()
    c = V[:, :m].T @ u
    beta[m - 1] = np.sqrt(u @ u - c @ c)
    V[:, m] = (u - V[:, :m] @ c) / beta[m - 1]
    return beta


def relation(A, V, T, b, m):
    e = np.zeros(m)
    e[-1] = 1.0
    return np.abs(A @ V[:, :m] - V[:, :m] @ T - b * np.outer(V[:, m], e)).max()


def test_delayed_step_matches_reorthogonalised_lanczos():
    rng = np.random.default_rng(5)
    n, m = 500, 20
    A = rng.standard_normal((n, n))
    A = (A + A.T) / 2
    V = np.zeros((n, m + 1))
    T = np.zeros((m, m))
    v0 = rng.standard_normal(n)
    V[:, 0] = v0 / np.linalg.norm(v0)
    beta = delayed_expand(A, V, T, 0, m, np.zeros(0))
    assert np.abs(V.T @ V - np.eye(m + 1)).max() < 1e-13
    assert relation(A, V, T, beta[m - 1], m) < 1e-12
    W = np.zeros((n, m + 1))
    W[:, 0] = V[:, 0]
    Tc = np.zeros((m, m))
    b = 0.0
    for j in range(m):
        w = A @ W[:, j]
        Tc[j, j] = W[:, j] @ w
        if j:
            Tc[j, j - 1] = Tc[j - 1, j] = b
        for _ in range(2):
            w -= W[:, :j + 1] @ (W[:, :j + 1].T @ w)
        b = np.linalg.norm(w)
        W[:, j + 1] = w / b
    assert np.abs(T - Tc).max() < 1e-12
    # thick restart: keep 6 Ritz vectors, the residual becomes column 6
    theta, S = np.linalg.eigh(T)
    keep = 6
    V2 = np.zeros_like(V)
    V2[:, :keep] = V[:, :m] @ S[:, :keep]
    V2[:, keep] = V[:, m]
    arrow = beta[m - 1] * S[m - 1, :keep]
    T2 = np.zeros((m, m))
    T2[np.arange(keep), np.arange(keep)] = theta[:keep]
    beta2 = delayed_expand(A, V2, T2, keep, m, arrow)
    assert np.abs(V2.T @ V2 - np.eye(m + 1)).max() < 1e-13
    assert relation(A, V2, T2, beta2[m - 1], m) < 1e-12
    # the Ritz values still approach the spectrum's lower end
    assert np.linalg.eigvalsh(T2)[0] - np.linalg.eigvalsh(A)[0] < np.linalg.eigvalsh(T)[0] - np.linalg.eigvalsh(A)[0]
