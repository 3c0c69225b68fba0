"""Host-side optimisers for FWI (NumPy, float64 model updates).

Restates the reference's optimisers so a full inversion runs on the device
objective of :class:`~soundwavesimulation_amd.inversion.fwi.TorchLeastSquaresFWI`:

* ``polak_ribiere`` -- nonlinear CG with Armijo backtracking; search directions
  live on the mask, a negative PR beta restarts as steepest descent, and a
  failed line search stops the run (``fdtd2d/inversion/cg.py:15-101``);
* ``fixed_step_gradient_descent`` (``invert_ring.py:782-814``);
* ``adam`` -- projected Adam on the mask (``invert_ring.py:817-861``).

All three keep a ``history`` dict with ``misfit``, ``step_size`` and
``grad_norm`` lists, as the reference does.
"""

from __future__ import annotations

import numpy as np


def masked_dot(a, b, mask):
    return float(np.dot(np.asarray(a)[mask], np.asarray(b)[mask]))


def _history(value, gnorm):
    return {"misfit": [float(value)], "step_size": [0.0], "grad_norm": [float(gnorm)]}


def polak_ribiere(m0, misfit, misfit_and_grad, project, mask, max_iter=15, armijo_c1=1e-4,
                  shrink=0.5, max_line_search=12, step_fraction=0.05, verbose=True):
    """Minimise J(m) with Polak-Ribiere CG and an Armijo line search."""
    m = project(np.asarray(m0, dtype=float))
    J, g = misfit_and_grad(m)
    history = _history(J, np.sqrt(max(masked_dot(g, g, mask), 0.0)))
    p = -g
    alpha_prev = None
    for it in range(max_iter):
        gnorm2 = masked_dot(g, g, mask)
        if gnorm2 <= 0.0:
            break
        mnorm = np.sqrt(max(masked_dot(m, m, mask), 0.0))
        accepted = False
        alpha = 0.0
        m_trial, J_trial = m, J
        for restart in (False, True):
            if restart:
                p = -g
            gdotp = masked_dot(g, p, mask)
            if gdotp >= 0.0:
                p = -g
                gdotp = -gnorm2
            pnorm = np.sqrt(max(masked_dot(p, p, mask), 0.0))
            if pnorm <= 0.0:
                break
            alpha = (step_fraction * mnorm / pnorm if (restart or alpha_prev is None)
                     else alpha_prev * 1.5)
            for _ in range(max_line_search):
                m_trial = project(m + alpha * p)
                J_trial = misfit(m_trial)
                if J_trial <= J + armijo_c1 * alpha * gdotp:
                    accepted = True
                    break
                alpha *= shrink
            if accepted:
                break
        if not accepted:
            if verbose:
                print("  CG line search failed; stopping.")
            break
        m = m_trial
        J_new, g_new = misfit_and_grad(m)
        beta = max(masked_dot(g_new, g_new - g, mask) / (masked_dot(g, g, mask) + 1e-30), 0.0)
        p = -g_new + beta * p
        g, J = g_new, J_new
        alpha_prev = alpha
        gnorm = np.sqrt(max(masked_dot(g, g, mask), 0.0))
        history["misfit"].append(float(J))
        history["step_size"].append(float(alpha))
        history["grad_norm"].append(gnorm)
        if verbose:
            print(f"  CG iter {it + 1:3d}  J = {J:.6e}  step = {alpha:.3e}  |g| = {gnorm:.4e}")
    return m, history


def fixed_step_gradient_descent(m0, misfit_and_grad, project, mask, step_size, max_iter=15,
                                verbose=True):
    """Projected gradient descent with one fixed step size."""
    m = project(np.asarray(m0, dtype=float))
    value, gradient = misfit_and_grad(m)
    grad_norm = float(np.linalg.norm(gradient[mask]))
    history = _history(value, grad_norm)
    for iteration in range(max_iter):
        if grad_norm <= 0.0:
            break
        m = project(m - step_size * gradient)
        value, gradient = misfit_and_grad(m)
        grad_norm = float(np.linalg.norm(gradient[mask]))
        history["misfit"].append(float(value))
        history["step_size"].append(float(step_size))
        history["grad_norm"].append(grad_norm)
        if verbose:
            print(f"  GD iter {iteration + 1:3d} J = {value:.6e} step = {step_size:.3e} "
                  f"|g| = {grad_norm:.4e}")
    return m, history


def adam(m0, misfit_and_grad, project, mask, learning_rate, max_iter=15, beta1=0.9,
         beta2=0.999, epsilon=1e-30, verbose=True):
    """Projected Adam with moments kept on the mask only."""
    m = project(np.asarray(m0, dtype=float))
    value, gradient = misfit_and_grad(m)
    first = np.zeros_like(m)
    second = np.zeros_like(m)
    history = _history(value, np.linalg.norm(gradient[mask]))
    for iteration in range(1, max_iter + 1):
        first[mask] = beta1 * first[mask] + (1.0 - beta1) * gradient[mask]
        second[mask] = beta2 * second[mask] + (1.0 - beta2) * gradient[mask] ** 2
        first_hat = first[mask] / (1.0 - beta1**iteration)
        second_hat = second[mask] / (1.0 - beta2**iteration)
        update = np.zeros_like(m)
        update[mask] = learning_rate * first_hat / (np.sqrt(second_hat) + epsilon)
        m = project(m - update)
        value, gradient = misfit_and_grad(m)
        grad_norm = float(np.linalg.norm(gradient[mask]))
        history["misfit"].append(float(value))
        history["step_size"].append(float(learning_rate))
        history["grad_norm"].append(grad_norm)
        if verbose:
            print(f"  Adam iter {iteration:3d} J = {value:.6e} lr = {learning_rate:.3e} "
                  f"|g| = {grad_norm:.4e}")
    return m, history
