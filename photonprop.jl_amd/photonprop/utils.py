"""Host numerics mirrored from PhotonProp.jl's Utils module (src/utils.jl)."""
import numpy as np

# Gauss-Legendre 50-point rule (src/utils.jl:8, FastGaussQuadrature.gausslegendre(50)).
GL50 = np.polynomial.legendre.leggauss(50)


def fast_linear_interp(x_eval, xs, ys):
    """Clamped linear interpolation on sorted xs (src/utils.jl:10-34)."""
    xs = np.asarray(xs, dtype=np.float64)
    ys = np.asarray(ys, dtype=np.float64)
    x = np.clip(np.asarray(x_eval, dtype=np.float64), xs[0], xs[-1])
    return np.interp(x, xs, ys)


def fast_linear_interp_knots(x, knots, lower, upper):
    """Clamped linear interpolation on knots equidistant over [lower, upper]
    (src/utils.jl:37-60)."""
    knots = np.asarray(knots, dtype=np.float64)
    x = np.clip(np.asarray(x, dtype=np.float64), lower, upper)
    n = knots.shape[0]
    along = (x - lower) / ((upper - lower) / (n - 1))
    fl = np.floor(along)
    k = fl.astype(np.int64)
    top = k >= n - 1
    k = np.minimum(k, n - 2)
    y = knots[k] + (knots[k + 1] - knots[k]) * (along - fl)
    return np.where(top, knots[-1], y)


def transform_integral_range(x, f, xrange):
    """Map a node on [-1, 1] to xrange and scale (src/utils.jl:63-69)."""
    ba_half = (xrange[1] - xrange[0]) / 2
    u = ba_half * x + (xrange[0] + xrange[1]) / 2
    return f(u) * ba_half


def integrate_gauss_quad(f, a, b, order=None):
    """Gauss-Legendre quadrature of f on [a, b] (src/utils.jl:71-78); f must
    accept numpy arrays."""
    nodes, weights = GL50 if order is None else np.polynomial.legendre.leggauss(order)
    return float(np.dot(weights, transform_integral_range(nodes, f, (a, b))))
