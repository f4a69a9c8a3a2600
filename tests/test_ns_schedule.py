"""CPU checks of the sign-iteration schedule used by the mixed-precision
projection (madam_ns_schedule, no GPU needed).

The projection computes W_+ = W (I + sign(W)) / 2 with sign(W) from S_0 =
W / |W|_F by the odd quintic steps x -> a x + b x^3 + c x^5 in FP32, then one
FP64 step x -> x (3 - x^2) / 2.  On the eigenvalues this is a scalar map, so
the schedule is checked on a dense grid of scalars: every x in [1e-6, 1]
must reach 1 to FP64 precision, also when every FP32 step is perturbed by a
relative error far above FP32 rounding (a symmetric FP32 product of order
n <= 1024 has spectral error below ~1e-5 relative), and the iterates must
stay inside [-2.1, 2.1] (the schedule's intervals reach up to 2; the
kernel's guard rejects a final FP32 iterate above 1.5).  A schedule without
the stability margin (the minimax polynomials fitted on the exact
intervals) is the negative control: the same check must reject it.

Accuracy: a step must never squash a value that is already large towards 0
(a minimax polynomial for a wide interval does, at its interior minima):
FP32 rounding is absolute (~1e-7), so a squashed eigenvalue's eigenvector
mixes with its neighbours at relative size 1e-7 / (squashed value), and that
rotation survives into W_+.  The check: along every scalar trajectory,
x_{k+1} >= 0.3 x_k whenever x_k >= 0.1.  The widened minimax schedule
(stable, but with squashing minima) is its negative control, and an FP32
emulation of the whole projection on a spectrum built to hit those minima
checks the consequence against numpy's eigh.
"""
import numpy as np
import pytest

import madam_amd

# Minimax odd quintics fitted on the exact intervals [l_k, u_k] with no
# widening of the upper end: mathematically convergent, numerically unstable.
UNSTABLE = np.array([
    [6.5789436, -12.24106284, 7.66211266], [4.2570903, -3.16017155, 0.58647659],
    [4.25679561, -3.15982678, 0.58639591], [4.25610454, -3.15951454, 0.58640776],
    [4.25249924, -3.15681172, 0.58606352], [4.21653548, -3.10188261, 0.57502521],
    [4.17739833, -3.10488706, 0.58016048], [3.93924782, -2.93408607, 0.56002186],
    [3.23597078, -2.42721147, 0.50310531], [2.23586608, -1.61014368, 0.412479],
    [1.88551957, -1.26163833, 0.37616478], [1.87500592, -1.25001147, 0.37500556]],
    dtype=np.float32)


def run_schedule(coef, x0, pert):
    """FP32 steps with a relative perturbation pert (array or scalar) per
    step, then the FP64 correction.  Returns (final values, max |x| seen)."""
    x = x0.astype(np.float32)
    peak = float(np.nanmax(np.abs(x)))
    with np.errstate(over="ignore", invalid="ignore"):
        for k, (a, b, c) in enumerate(coef.astype(np.float32)):
            x2 = x * x
            x = (x * (a + x2 * (b + c * x2))) * np.float32(1.0 + pert[k] if np.ndim(pert) else 1.0 + pert)
            peak = max(peak, float(np.nanmax(np.abs(x))) if np.all(np.isfinite(x)) else np.inf)
        xd = x.astype(np.float64)
        xd = xd * (3.0 - xd * xd) / 2.0
    return xd, peak


def grid():
    return np.unique(np.concatenate([np.geomspace(1e-6, 1.0, 200001), np.linspace(1e-6, 1.0, 200001)]))


def schedule_ok(coef):
    xs = grid()
    rng = np.random.default_rng(7)
    perts = [0.0, 1e-5, -1e-5, 1e-4, -1e-4, rng.uniform(-1e-4, 1e-4, len(coef))]
    for p in perts:
        for sgn in (1.0, -1.0):
            xd, peak = run_schedule(coef, sgn * xs, p)
            if not np.all(np.isfinite(xd)) or peak > 2.1:
                return False
            if np.max(np.abs(xd - sgn)) > 1e-7:
                return False
    return True


# Minimax odd quintics fitted on the intervals widened by 2% at the top:
# stable, but squashing (maps x ~ 1.68 to ~2e-4 at its third step).
MINIMAX_WIDENED = np.array([
    [6.4774747, -11.6729765, 7.0469713], [4.172062, -2.9745662, 0.53019804],
    [4.1717052, -2.9740622, 0.53007025], [4.172711, -2.97731, 0.53112775],
    [4.1678057, -2.9713175, 0.52972764], [4.155663, -2.9661732, 0.52990985],
    [4.1052055, -2.9328153, 0.5263809], [3.9055731, -2.800026, 0.5122276],
    [3.2904637, -2.3776097, 0.4675132], [2.307551, -1.629218, 0.39117578],
    [1.8776932, -1.2338414, 0.35685068], [1.8565079, -1.2132474, 0.35673976]],
    dtype=np.float32)


def min_squash(coef):
    """min over trajectories and steps of x_{k+1} / x_k where x_k >= 0.1."""
    x = grid().astype(np.float64)
    worst = np.inf
    for a, b, c in coef.astype(np.float64):
        x2 = x * x
        xn = x * (a + x2 * (b + c * x2))
        big = x >= 0.1
        if np.any(big):
            worst = min(worst, float(np.min(xn[big] / x[big])))
        x = xn
    return worst


def emulate_projection(coef, W):
    """The kernel's arithmetic in numpy: FP32 products (lower triangle kept
    and mirrored), FP64 correction step and X_+ = W (I + S) / 2."""
    n = W.shape[0]
    S = (W / np.linalg.norm(W)).astype(np.float32)
    I32 = np.eye(n, dtype=np.float32)
    for a, b, c in coef.astype(np.float32):
        T = S @ S
        T = np.tril(T) + np.tril(T, -1).T
        U = T @ (b * I32 + c * T)
        U = np.tril(U) + np.tril(U, -1).T
        S = S @ (a * I32 + U)
        S = np.tril(S) + np.tril(S, -1).T
    S = S.astype(np.float64)
    S = S @ (1.5 * np.eye(n) - 0.5 * (S @ S))
    return W @ (0.5 * np.eye(n) + 0.5 * S)


def test_schedule_shape_and_oddness():
    coef = madam_amd.ns_schedule()
    assert coef.shape == (13, 3) and coef.dtype == np.float32
    # odd map: sign(-x) = -sign(x) holds by construction; the last steps
    # approach the classical stable quintic (15 x - 10 x^3 + 3 x^5) / 8
    # whose value at 1 is 1
    assert abs(float(coef[-1].sum()) - 1.0) < 1e-3


def test_schedule_converges_and_is_stable():
    assert schedule_ok(madam_amd.ns_schedule())


def test_unstable_schedule_is_rejected():
    assert not schedule_ok(UNSTABLE)


def test_small_eigenvalues_error_bounded():
    """Below 1e-6 the iteration need not reach 1; the projection error
    |lambda| (1 - |s|) / 2 must stay below |lambda| / 2 (s in [0, 1])."""
    coef = madam_amd.ns_schedule()
    xs = np.geomspace(1e-10, 1e-6, 2001)
    xd, peak = run_schedule(coef, xs, 0.0)
    assert peak <= 2.1
    assert np.all(xd >= 0.0) and np.all(xd <= 1.0 + 1e-12)


def test_schedule_never_squashes():
    assert min_squash(madam_amd.ns_schedule()) >= 0.3
    assert min_squash(MINIMAX_WIDENED) < 1e-2


def test_emulated_projection_accuracy():
    """Spectrum with eigenvalues at +0.166 and -0.0738 of |W|_F = 1 (values that
    the widened minimax schedule squashes to ~2e-4 at its third step) plus a
    spread down to 1e-4: the FP32 emulation must match eigh to 3e-6 |W|_F
    (Frobenius) with the schedule, and misses by >1e-5 with the control."""
    n = 100
    rng = np.random.default_rng(11)
    Qm, _ = np.linalg.qr(rng.standard_normal((n, n)))
    special = np.array([0.166, -0.0738, 0.1664, -0.0737, 0.3794, 0.01])
    rest = np.sign(rng.standard_normal(n - 6)) * np.geomspace(1e-4, 0.2, n - 6)
    rest *= np.sqrt(1.0 - np.sum(special ** 2)) / np.linalg.norm(rest)
    w = np.concatenate([special, rest])  # |w| = 1: S_0 = W keeps these values
    W = (Qm * w) @ Qm.T
    W = (W + W.T) / 2
    ev, V = np.linalg.eigh(W)
    Xe = (V * np.maximum(ev, 0)) @ V.T
    fro = np.linalg.norm(W)
    good = np.linalg.norm(emulate_projection(madam_amd.ns_schedule(), W) - Xe) / fro
    bad = np.linalg.norm(emulate_projection(MINIMAX_WIDENED, W) - Xe) / fro
    print("emulated projection error", good, "control", bad)
    assert good <= 3e-6
    assert bad > 1e-5
