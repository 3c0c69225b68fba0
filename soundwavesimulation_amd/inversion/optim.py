"""Host-side optimisers for FWI (NumPy, float64 model updates).

Three methods run a full inversion on the device objective of
:class:`~soundwavesimulation_amd.inversion.fwi.TorchLeastSquaresFWI`, with the
reference's signatures, step rules and ``history`` keys
(``fdtd2d/inversion/cg.py:15-101``, ``invert_ring.py:782-861``):

* :func:`polak_ribiere` -- nonlinear conjugate gradients (PR+, i.e. a
  negative beta resets to steepest descent) with an Armijo backtracking
  search. The first trial step of a fresh direction moves the model by
  ``step_fraction`` of its norm; later searches start from 1.5 times the last
  accepted step. When the conjugate direction finds no sufficient decrease,
  steepest descent is tried from a fresh step; when that fails as well the
  run stops where it is.
* :func:`fixed_step_gradient_descent` -- projected descent with one step.
* :func:`adam` -- projected Adam whose moments live on the mask only.

Every inner product, norm and moment is taken over the interior mask (the
model outside it is fixed by the projection). ``history`` holds one entry per
accepted model in ``misfit``, ``step_size`` and ``grad_norm``.
"""

from __future__ import annotations

import math

import numpy as np


def masked_dot(a, b, mask):
    """``<a, b>`` over the cells of ``mask``."""
    return float(np.dot(np.asarray(a)[mask], np.asarray(b)[mask]))


class _MaskSpace:
    """Vectors of the model restricted to the interior mask."""

    def __init__(self, mask):
        self.mask = np.asarray(mask, dtype=bool)

    def dot(self, a, b):
        return masked_dot(a, b, self.mask)

    def norm(self, a):
        return math.sqrt(max(self.dot(a, a), 0.0))


class _Log:
    """The reference's ``history`` dict, written one accepted model at a time."""

    def __init__(self, value, grad_norm):
        self.data = {"misfit": [float(value)], "step_size": [0.0],
                     "grad_norm": [float(grad_norm)]}

    def add(self, value, step, grad_norm):
        self.data["misfit"].append(float(value))
        self.data["step_size"].append(float(step))
        self.data["grad_norm"].append(float(grad_norm))


def armijo_backtrack(misfit, project, m, value, direction, slope, step, c1, shrink, tries):
    """First ``step * shrink**k`` (k < tries) whose projected trial model meets
    ``J(trial) <= J(m) + c1 * step * slope``; ``(step, trial, J)`` or None."""
    for _ in range(tries):
        trial = project(m + step * direction)
        trial_value = misfit(trial)
        if trial_value <= value + c1 * step * slope:
            return step, trial, trial_value
        step *= shrink
    return None


def polak_ribiere(m0, misfit, misfit_and_grad, project, mask, max_iter=15, armijo_c1=1e-4,
                  shrink=0.5, max_line_search=12, step_fraction=0.05, verbose=True):
    """Minimise J(m) with Polak-Ribiere CG; returns ``(m, history)``.

    ``misfit(m)`` returns J, ``misfit_and_grad(m)`` returns ``(J, g)``;
    ``project`` maps a model onto the admissible set.
    """
    space = _MaskSpace(mask)
    m = project(np.asarray(m0, dtype=float))
    value, grad = misfit_and_grad(m)
    log = _Log(value, space.norm(grad))
    direction = -grad
    last_step = None
    for it in range(max_iter):
        g2 = space.dot(grad, grad)
        if g2 <= 0.0:
            break
        m_norm = space.norm(m)
        found = None
        # the conjugate direction first (or steepest descent if it does not
        # descend), then steepest descent from a fresh step
        for fresh in (False, True):
            if fresh:
                direction = -grad
            slope = space.dot(grad, direction)
            if slope >= 0.0:
                direction, slope = -grad, -g2
            d_norm = space.norm(direction)
            if d_norm <= 0.0:
                break
            start = (step_fraction * m_norm / d_norm if fresh or last_step is None
                     else 1.5 * last_step)
            found = armijo_backtrack(misfit, project, m, value, direction, slope, start,
                                     armijo_c1, shrink, max_line_search)
            if found is not None:
                break
        if found is None:
            if verbose:
                print("  CG line search failed; stopping.")
            break
        last_step, m, _ = found
        new_value, new_grad = misfit_and_grad(m)
        beta = max(space.dot(new_grad, new_grad - grad) / (space.dot(grad, grad) + 1e-30), 0.0)
        direction = beta * direction - new_grad
        value, grad = new_value, new_grad
        log.add(value, last_step, space.norm(grad))
        if verbose:
            print(f"  CG {it + 1:3d}: J = {value:.6e}, step {last_step:.3e}, "
                  f"|g| = {space.norm(grad):.4e}")
    return m, log.data


def _projected_iterations(m0, misfit_and_grad, project, mask, max_iter, update, label,
                          stop_on_zero_gradient, verbose):
    """Shared loop of the fixed-rule methods: ``m <- project(m - update(k, g))``."""
    mask = np.asarray(mask, dtype=bool)
    m = project(np.asarray(m0, dtype=float))
    value, grad = misfit_and_grad(m)
    grad_norm = float(np.linalg.norm(grad[mask]))
    log = _Log(value, grad_norm)
    for k in range(1, max_iter + 1):
        if stop_on_zero_gradient and grad_norm <= 0.0:
            break
        delta, step = update(k, grad)
        m = project(m - delta)
        value, grad = misfit_and_grad(m)
        grad_norm = float(np.linalg.norm(grad[mask]))
        log.add(value, step, grad_norm)
        if verbose:
            print(f"  {label} {k:3d}: J = {value:.6e}, step {step:.3e}, |g| = {grad_norm:.4e}")
    return m, log.data


def fixed_step_gradient_descent(m0, misfit_and_grad, project, mask, step_size, max_iter=15,
                                verbose=True):
    """Projected gradient descent with one fixed step size."""
    step_size = float(step_size)
    return _projected_iterations(m0, misfit_and_grad, project, mask, max_iter,
                                 lambda k, g: (step_size * g, step_size), "GD", True, verbose)


class _AdamMoments:
    """First and second moments of the gradient on the mask cells."""

    def __init__(self, mask, learning_rate, beta1, beta2, epsilon):
        self.mask = np.asarray(mask, dtype=bool)
        n = int(self.mask.sum())
        self.first = np.zeros(n)
        self.second = np.zeros(n)
        self.lr, self.b1, self.b2, self.eps = float(learning_rate), beta1, beta2, epsilon

    def __call__(self, k, grad):
        g = grad[self.mask]
        self.first = self.b1 * self.first + (1.0 - self.b1) * g
        self.second = self.b2 * self.second + (1.0 - self.b2) * g**2
        m_hat = self.first / (1.0 - self.b1**k)
        v_hat = self.second / (1.0 - self.b2**k)
        delta = np.zeros(grad.shape)
        delta[self.mask] = self.lr * m_hat / (np.sqrt(v_hat) + self.eps)
        return delta, self.lr


def adam(m0, misfit_and_grad, project, mask, learning_rate, max_iter=15, beta1=0.9,
         beta2=0.999, epsilon=1e-30, verbose=True):
    """Projected Adam with moments kept on the mask only."""
    moments = _AdamMoments(mask, learning_rate, beta1, beta2, epsilon)
    return _projected_iterations(m0, misfit_and_grad, project, mask, max_iter, moments,
                                 "Adam", False, verbose)
