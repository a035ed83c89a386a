"""Fit the polynomial coefficients used by gs_math.h.

Each kernel is a near-minimax polynomial obtained by relative-error weighted
least squares on Chebyshev nodes, in 50-digit arithmetic (mpmath), then rounded
to float32.  The printed block is pasted into gs_math.h; the accuracy of the
resulting float32 functions is checked exhaustively by tests/test_gs_math.py.

Methods (textbook, see J.-M. Muller, "Elementary Functions", ch. 3 and 11):
  exp  : Cody-Waite reduction x = n ln2 + r, |r| <= ln2/2,
         e^r = 1 + r + r^2 q(r),  q of degree 4.
  log  : x = 2^e m, m in [sqrt(1/2), sqrt(2)), f = m - 1,
         log1p(f) = f + f^2 Q(f),  Q of degree 8.
  atan : a in [0, 1], atan(a) = a + a^3 P(a^2),  P of degree 8.
  asin : x in [0, 1/2], asin(x) = x + x^3 S(x^2),  S of degree 5;
         x > 1/2 uses asin(x) = pi/2 - 2 asin(sqrt((1 - x) / 2)).
"""
import struct
import mpmath as mp

mp.mp.dps = 50


def fit(f, a, b, deg, npts=600):
    xs = [(a + b) / 2 + (b - a) / 2 * mp.cos(mp.pi * (i + 0.5) / npts) for i in range(npts)]
    A = mp.matrix(npts, deg + 1)
    y = mp.matrix(npts, 1)
    for i, x in enumerate(xs):
        fx = f(x)
        w = 1 / abs(fx)
        for k in range(deg + 1):
            A[i, k] = x ** k * w
        y[i] = fx * w
    c = mp.lu_solve(A.T * A, A.T * y)
    return [c[k] for k in range(deg + 1)]


def f32(v):
    return struct.unpack('f', struct.pack('f', float(v)))[0]


def emit(name, coeffs):
    out = []
    for k, c in enumerate(coeffs):
        out.append('#define %s_%d %sf  /* %s */' % (name, k, float.hex(f32(c)), mp.nstr(c, 12)))
    return out


def main():
    L = mp.log(2) / 2
    q = fit(lambda r: (mp.exp(r) - 1 - r) / r ** 2, -L, L, 4)
    Q = fit(lambda t: (mp.log1p(t) - t) / t ** 2, mp.sqrt(0.5) - 1, mp.sqrt(2) - 1, 8)
    P = fit(lambda u: (mp.atan(mp.sqrt(u)) - mp.sqrt(u)) / (u * mp.sqrt(u)), mp.mpf(1e-30), mp.mpf(1), 8)
    S = fit(lambda u: (mp.asin(mp.sqrt(u)) - mp.sqrt(u)) / (u * mp.sqrt(u)), mp.mpf(1e-30), mp.mpf(0.25), 5)
    lines = []
    for n, c in (('GS_EXP_Q', q), ('GS_LOG_Q', Q), ('GS_ATAN_P', P), ('GS_ASIN_S', S)):
        lines += emit(n, c)
    print('\n'.join(lines))


if __name__ == '__main__':
    main()
