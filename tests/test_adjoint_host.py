"""The step adjoint's mathematics (ode-0.12_amd/csrc/adj_step.hpp, built for
the host in double) against FP64 central differences through the CPU
restatement with a converged solver: gradients of the squared centre of mass
after N steps with respect to initial linear/angular velocities, mu and the
restitution coefficient."""
import numpy as np
import pytest

from adjoint.harness import contact_scene, converged_oracle, fd_grad, host_adjoint, jointed_scene, vehicle_scene

CASES = {
    "spheres_bounce": (dict(seed=4, n=4, kinds=(0,), bounce=0.5), 40),
    "boxes": (dict(seed=5, n=4, kinds=(1,)), 40),
    "mixed_pile": (dict(seed=6, n=6, kinds=(0, 1)), 60),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_host_adjoint_matches_central_differences(name):
    kw, n = CASES[name]
    sc = contact_scene(**kw)
    L, gin, gmu, gb, stats = host_adjoint(sc, n)
    assert stats[1] == 0, "dual-number collider disagreed with the recorded contacts"
    gmax = float(np.abs(gin[:, 7:13]).max())
    assert gmax > 0
    worst = 0.0
    for b in range(sc.nb):
        for field, col in (("body_lvel", 7), ("body_avel", 10)):
            k = (b + col) % 3
            fd = fd_grad(sc, n, field, (b, k))
            worst = max(worst, abs(fd - gin[b, col + k]) / gmax)
    assert worst <= 1e-4, worst
    fd_mu = fd_grad(sc, n, "mu")
    assert abs(fd_mu - gmu) <= 1e-4 * max(gmax, abs(fd_mu))
    if sc.bounce > 0:
        fd_b = fd_grad(sc, n, "bounce")
        assert abs(fd_b - gb) <= 1e-4 * max(gmax, abs(fd_b))


JOINT_CASES = {
    "chain_on_plane": (dict(seed=3, n_links=4), 30),
    "chain_falling": (dict(seed=8, n_links=5), 25),
}


@pytest.mark.parametrize("name", sorted(JOINT_CASES))
def test_host_adjoint_with_joints_matches_central_differences(name):
    """Ball joints and hinges (two-body and to the static world) with ground
    contacts and friction: every velocity component of every body against
    FP64 central differences, with the forward LCP converged (a chain lying
    on a plane is statically indeterminate, so a fixed iteration count
    leaves impulses that depend on the count)."""
    kw, n = JOINT_CASES[name]
    sc = jointed_scene(**kw)
    with converged_oracle(1e-12):
        L, gin, gmu, gb, stats = host_adjoint(sc, n)
        assert stats[1] == 0, "dual-number collider disagreed with the recorded contacts"
        gmax = float(np.abs(gin[:, 7:13]).max())
        assert gmax > 0
        worst = 0.0
        for b in range(sc.nb):
            for field, col in (("body_lvel", 7), ("body_avel", 10)):
                for k in range(3):
                    fd = fd_grad(sc, n, field, (b, k))
                    worst = max(worst, abs(fd - gin[b, col + k]) / gmax)
        assert worst <= 1e-4, worst
        fd_mu = fd_grad(sc, n, "mu")
        assert abs(fd_mu - gmu) <= 1e-4 * max(gmax, abs(fd_mu))


def test_host_adjoint_of_a_vehicle_matches_central_differences():
    """hinge2 joints: the suspension rows, the axis row, the steering limit
    row (stops set per step by the controls, lo = hi) and the wheel torques
    added through dJointAddHinge2Torques, over 50 steps in which the car
    lands: every velocity component of every body against FP64 central
    differences (measured 3.5e-9 of the largest gradient)."""
    n = 50
    sc = vehicle_scene(5, n)
    L, gin, gmu, gb, stats = host_adjoint(sc, n)
    assert stats[1] == 0
    gmax = float(np.abs(gin[:, 7:13]).max())
    assert gmax > 0 and float(np.abs(gin[:, 10:13]).max()) > 1e-3 * gmax
    worst = 0.0
    for b in range(sc.nb):
        for field, col in (("body_lvel", 7), ("body_avel", 10)):
            for k in range(3):
                fd = fd_grad(sc, n, field, (b, k))
                worst = max(worst, abs(fd - gin[b, col + k]) / gmax)
    assert worst <= 1e-4, worst
