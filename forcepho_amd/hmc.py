"""Leapfrog integration for Hamiltonian Monte Carlo on the patch model.

forcepho samples with littlemcmc's NUTS (forcepho/fitting.py:254-341), which
calls Posterior.lnprob_and_grad once per leapfrog step.  ``leapfrog`` is that
inner integrator: given U(q) = -lnp(q) and its gradient it advances
(q, p) by n_steps steps of size eps with a diagonal inverse mass matrix.
"""

import numpy as np

__all__ = ["leapfrog", "hamiltonian"]


def hamiltonian(U, p, inv_mass=1.0):
    return U + 0.5 * np.sum(inv_mass * p * p)


def leapfrog(q, p, potential_and_grad, eps, n_steps, inv_mass=1.0, record=False):
    """Integrate n_steps leapfrog steps.

    Parameters
    ----------
    q, p : float64 arrays, position and momentum
    potential_and_grad : q -> (U, dU/dq)
    eps : step size
    inv_mass : diagonal inverse mass (scalar or array)

    Returns (q, p, U, n_gradient_calls[, trajectory of (q, p, U)])
    """
    q = np.array(q, dtype=np.float64)
    p = np.array(p, dtype=np.float64)
    U, g = potential_and_grad(q)
    calls = 1
    traj = [(q.copy(), p.copy(), U)] if record else None
    p = p - 0.5 * eps * g
    for i in range(n_steps):
        q = q + eps * inv_mass * p
        U, g = potential_and_grad(q)
        calls += 1
        if i < n_steps - 1:
            p = p - eps * g
        if record:
            traj.append((q.copy(), p - 0.5 * eps * g if i < n_steps - 1 else None, U))
    p = p - 0.5 * eps * g
    if record:
        return q, p, U, calls, traj
    return q, p, U, calls
