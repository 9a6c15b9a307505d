"""Derives the polynomial coefficients of the device sin/cos kernels in trig.h.

Method: weighted least squares on Chebyshev nodes in 60-digit arithmetic
(mpmath), minimising the relative error of the reduced-range kernel; the
result is close to minimax at these degrees.  Re-run to regenerate:

    python abt_amd/csrc/gen_trig_coeffs.py

Kernels fitted
  fp64:  sin(x) = x + x^3 * S(x^2),  cos(x) = 1 - x^2/2 + x^4 * C(x^2)   on |x| <= pi/4
  fp32:  sin(2 pi r) = r * s(r^2),   cos(2 pi r) = c(r^2)                on |r| <= 1/8
  fp32:  sin(2 pi r) = r * h(r^2),   cos(2 pi r) = k(r^2)                on |r| <= 1/4
"""
import mpmath as mp

mp.mp.dps = 60


def lsq(f, g, nterms, a, b, npts=400):
    """Fit f(x) ~ g(x) * sum_k c_k x^(2k) on [a, b] (relative error weight 1/f)."""
    xs = [(a + b) / 2 + (b - a) / 2 * mp.cos(mp.pi * (i + mp.mpf(0.5)) / npts) for i in range(npts)]
    rows, rhs = [], []
    for x in xs:
        w = 1 / abs(f(x)) if f(x) != 0 else mp.mpf(1)
        rows.append([w * g(x) * x ** (2 * k) for k in range(nterms)])
        rhs.append(w * f(x))
    A = mp.matrix(rows)
    y = mp.matrix(rhs)
    c = mp.lu_solve(A.T * A, A.T * y)
    return [c[k] for k in range(nterms)]


def main():
    q = mp.pi / 4
    # fp64 sin: (sin x - x) / x^3 = S(x^2), relative to sin x
    S = lsq(lambda x: mp.sin(x) - x, lambda x: x ** 3, 6, mp.mpf("1e-4"), q)
    C = lsq(lambda x: mp.cos(x) - 1 + x * x / 2, lambda x: x ** 4, 6, mp.mpf("1e-4"), q)
    s32 = lsq(lambda r: mp.sin(2 * mp.pi * r), lambda r: r, 5, mp.mpf("1e-6"), mp.mpf(1) / 8)
    c32 = lsq(lambda r: mp.cos(2 * mp.pi * r), lambda r: mp.mpf(1), 5, mp.mpf("1e-6"), mp.mpf(1) / 8)
    h32 = lsq(lambda r: mp.sin(2 * mp.pi * r), lambda r: r, 5, mp.mpf("1e-6"), mp.mpf(1) / 4)
    # cos on |r| <= 1/4 passes through 0 at the end: fit the absolute error (weight 1)
    k32 = lsq(lambda r: mp.cos(2 * mp.pi * r), lambda r: mp.mpf(1), 6, mp.mpf("1e-6"), mp.mpf(1) / 4)
    fmt = lambda v: mp.nstr(v, 20, min_fixed=-1, max_fixed=-1)
    print("// fp64 S(z), z = x^2, x^3 .. x^13")
    for i, v in enumerate(S):
        print(f"#define ABT_S{i + 1} {fmt(v)}")
    print("// fp64 C(z), x^4 .. x^14")
    for i, v in enumerate(C):
        print(f"#define ABT_C{i + 1} {fmt(v)}")
    print("// fp32 sin(2 pi r) = r * s(r^2)")
    for i, v in enumerate(s32):
        print(f"#define ABT_SR{i} {fmt(v)}f")
    print("// fp32 cos(2 pi r) = c(r^2)")
    for i, v in enumerate(c32):
        print(f"#define ABT_CR{i} {fmt(v)}f")
    print("// fp32 sin(2 pi r) = r * h(r^2), |r| <= 1/4")
    for i, v in enumerate(h32):
        print(f"#define ABT_HR{i} {fmt(v)}f")
    print("// fp32 cos(2 pi r) = k(r^2), |r| <= 1/4")
    for i, v in enumerate(k32):
        print(f"#define ABT_KR{i} {fmt(v)}f")
    errk = max(abs(sum(k32[i] * (r * r) ** i for i in range(6)) - mp.cos(2 * mp.pi * r))
               for r in [mp.mpf(k) / 4000 for k in range(0, 1001)])
    errh = max(abs(r * sum(h32[i] * (r * r) ** i for i in range(5)) - mp.sin(2 * mp.pi * r))
               for r in [mp.mpf(k) / 4000 for k in range(0, 1001)])
    print("// max abs err (exact arithmetic): sin(2 pi r)", mp.nstr(errh, 3), "cos(2 pi r)", mp.nstr(errk, 3))
    # report the fitted relative error in high precision
    err = lambda f, p: max(abs(p(x) / f(x) - 1) for x in [q * k / 997 for k in range(1, 998)])
    ps = lambda x: x + x ** 3 * sum(S[k] * x ** (2 * k) for k in range(6))
    pc = lambda x: 1 - x * x / 2 + x ** 4 * sum(C[k] * x ** (2 * k) for k in range(6))
    print("// max rel err (exact arithmetic): sin", mp.nstr(err(mp.sin, ps), 3),
          "cos", mp.nstr(err(mp.cos, pc), 3))


if __name__ == "__main__":
    main()
