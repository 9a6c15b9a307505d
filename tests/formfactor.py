"""Independent textbook references used to pin the oracle (numpy only).

Nothing here follows the reference's formulas; each function is a separate
derivation used as a check:
  * polygon form factor by collapsed Gauss-Legendre quadrature (brute force)
  * triangle form factor by the Hermite-Genocchi divided-difference closed form
  * Jones -> Mueller by the Kronecker construction (Bohren & Huffman 1983, sec. 3.3)
  * Karczewski-Wolf e-m matrix from the vector Kirchhoff integrands
"""
import numpy as np


def tri_area(p):
    p = np.asarray(p, dtype=np.float64)
    return 0.5 * abs((p[1, 0] - p[0, 0]) * (p[2, 1] - p[0, 1]) - (p[2, 0] - p[0, 0]) * (p[1, 1] - p[0, 1]))


def ff_quadrature(q, p, order=96):
    """int_T exp(i q.r) dA over a 2-D triangle p (3,2), Duffy-collapsed Gauss-Legendre."""
    q = np.asarray(q, dtype=np.float64)
    p = np.asarray(p, dtype=np.float64)
    g, w = np.polynomial.legendre.leggauss(order)
    u = 0.5 * (g + 1); wu = 0.5 * w
    U, V = np.meshgrid(u, u, indexing="ij")
    W = np.outer(wu, wu) * (1 - U)
    s = U
    t = V * (1 - U)
    r = p[0] + s[..., None] * (p[1] - p[0]) + t[..., None] * (p[2] - p[0])
    jac = 2 * tri_area(p)
    ph = r @ q
    return jac * np.sum(W * np.exp(1j * ph))


def ff_divided_difference(q, p):
    """Triangle form factor: -2A * second divided difference of exp(i x) at a_j = q.r_j."""
    a = np.asarray(p, dtype=np.float64) @ np.asarray(q, dtype=np.float64)
    A = tri_area(p)
    s = 0j
    for j in range(3):
        den = 1.0
        for k in range(3):
            if k != j:
                den *= a[j] - a[k]
        s += np.exp(1j * a[j]) / den
    return -2 * A * s


def jones_to_mueller(J):
    """Bohren & Huffman Stokes (I, Q, U, V) with E = (E_par, E_perp)."""
    U = np.array([[1, 0, 0, 1], [1, 0, 0, -1], [0, 1, 1, 0], [0, 1j, -1j, 0]])
    M = U @ np.kron(J, np.conj(J)) @ np.linalg.inv(U)
    return M.real


def kw_matrix(k, K):
    """Karczewski-Wolf e-m polarisation matrix from the vector Kirchhoff integrands.

    Aperture normal +z, incident direction K (K_y = 0), incident basis
    (e_perp = y, e_par = K x y), outgoing basis (m, k x m) with
    m = unit(y - (y.k) k).  E-theory field: k x (z x E); M-theory field:
    -k x (k x (z x (K x E))); the e-m theory is their mean."""
    k = np.asarray(k, float); K = np.asarray(K, float)
    z = np.array([0.0, 0.0, 1.0]); y = np.array([0.0, 1.0, 0.0])
    m = y - y.dot(k) * k
    m /= np.linalg.norm(m)
    ev = np.cross(k, m)
    out = np.zeros((2, 2))
    for col, E in enumerate([y, np.cross(K, y)]):
        fe = np.cross(k, np.cross(z, E))
        fm = -np.cross(k, np.cross(k, np.cross(z, np.cross(K, E))))
        f = 0.5 * (fe + fm)
        out[0, col] = f.dot(m)
        out[1, col] = f.dot(ev)
    return out
