"""NumPy restatement of the reference's projected Gauss-Newton CG optimizer
(modules/litho-inversion/src/inversion/gncg_optimizer.cpp:13-194, settings
gncg_optimizer.h:41-45, tolerance optimizer.h:31-35).

TEST INFRASTRUCTURE ONLY: the checker for the host library's GNCGOptimizer.  It is
pinned by hand-computed cases in tests/test_host_gncg.py (the Newton step of a quadratic,
the indefinite-curvature CG branch, the rejected line search, the stalled step, an
active bound) before the product is compared with it.

Returns the final point, the objective at it, the iteration count, the stop message, the
inner CG iterations of each outer iteration and the objective / gradient call counts.
"""
from __future__ import annotations

import numpy as np


def project(x, lb, ub):
    """gncg_optimizer.cpp:13-22"""
    return np.maximum(lb, np.minimum(ub, x))


def projected_gradient_norm(x, g, lb, ub):
    """gncg_optimizer.cpp:24-37"""
    pg = np.where((x <= lb + 1e-12) & (g > 0), 0.0, g)
    pg = np.where((x >= ub - 1e-12) & (g < 0), 0.0, pg)
    return float(np.sqrt(np.sum(pg * pg)))


def cg_solve(hv, b, max_iter, tol):
    """gncg_optimizer.cpp:39-77: returns (x, iterations)"""
    x = np.zeros_like(b)
    r = b.copy()
    p = r.copy()
    rs_old = float(r @ r)
    if rs_old < 1e-30:
        return x, 0
    it = 0
    for _ in range(max_iter):
        hp = hv(p)
        it += 1
        p2 = float(p @ p)
        if p2 < 1e-30:
            break
        alpha = rs_old / p2
        denom = float(p @ hp)
        if denom > 1e-16:
            alpha = rs_old / denom
        x = x + alpha * p
        r = r - alpha * hp
        rs_new = float(r @ r)
        if rs_new < tol * tol:
            break
        p = r + (rs_new / rs_old) * p
        rs_old = rs_new
    return x, it


def gncg_minimize(fun, grad, x0, lb, ub, max_iter=500, cg_max_iter=50, cg_tol=1e-6,
                  fd_step=2.0, ftol=1e-6, line_search_max_iter=30):
    """gncg_optimizer.cpp:105-194 (line search :79-103)"""
    calls = {"f": 0, "g": 0}

    def f(x):
        calls["f"] += 1
        return float(fun(x))

    def g(x):
        calls["g"] += 1
        return np.asarray(grad(x), dtype=np.float64)

    lb = np.asarray(lb, dtype=np.float64)
    ub = np.asarray(ub, dtype=np.float64)
    x = project(np.asarray(x0, dtype=np.float64), lb, ub)
    pg0 = 0.0
    stall = 0
    iters = 0
    converged = False
    message = ""
    cg_iters = []
    for k in range(max_iter):
        iters = k + 1
        gx = g(x)
        pg = projected_gradient_norm(x, gx, lb, ub)
        if k == 0:
            pg0 = pg
        if k > 0 and pg < ftol * max(1.0, pg0):
            converged = True
            message = "Converged: projected gradient below tolerance"
            break

        def hv(v, x=x):
            vn = float(np.sqrt(v @ v))
            if vn < 1e-15:
                return np.zeros_like(v)
            h = fd_step / vn
            gp = g(project(x + h * v, lb, ub))
            gm = g(project(x - h * v, lb, ub))
            return (gp - gm) / (2.0 * h)

        d, it = cg_solve(hv, -gx, cg_max_iter, cg_tol)
        cg_iters.append(it)
        if float(d @ gx) >= 0.0:
            d = -gx
        f0 = f(x)
        alpha = 1.0
        x_new = None
        for _ in range(line_search_max_iter):
            xt = project(x + alpha * d, lb, ub)
            if f(xt) <= f0:
                x_new = xt
                break
            alpha *= 0.5
        if x_new is None:
            converged = True
            message = "Converged: no step taken"
            break
        if float(np.sqrt(np.sum((x_new - x) ** 2))) < 1e-8:
            stall += 1
            if stall >= 3:
                converged = True
                message = "Converged: step stalled"
                break
        else:
            stall = 0
        x = x_new
    final = f(x)
    if not converged and iters >= max_iter:
        message = "Maximum iterations reached"
    return dict(x=x, f=final, iterations=iters, message=message, cg_iterations=cg_iters,
                objective_calls=calls["f"], gradient_calls=calls["g"])
