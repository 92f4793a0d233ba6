import numpy as np

from odemi import scenes


def test_box_stack_config():
    s = scenes.box_stack()
    assert s.nb == 10 and s.ng == 11 and s.dt == 0.01 and s.iterations == 20
    assert np.allclose(s.body_pos[:, 2], 0.5 + 1e-3 + np.arange(10) * (1 + 1e-3))
    assert np.allclose(s.body_mass, 1.0)


def test_random_pile_config():
    s = scenes.random_pile(4096, seed=12345)
    assert s.nb == 4096 and s.ng == 4097 and s.dt == 0.005 and s.mu == 0.6
    assert (s.geom_type[1::2] == scenes.BOX).all() and (s.geom_type[2::2] == scenes.SPHERE).all()
    assert s.body_pos[:, 2].min() >= 0.5 and s.body_pos[:, 2].max() <= 40.5


def test_uneven_worlds_deterministic_and_in_range():
    a = scenes.uneven_worlds(32, seed=7)
    b = scenes.uneven_worlds(32, seed=7)
    assert [w.nb for w in a] == [w.nb for w in b]
    for w in a:
        links = w.nj + (w.nj > 0)
        assert 8 <= w.nb
        assert set(np.unique(w.joint_type)) <= {scenes.BALL, scenes.HINGE}


def test_vehicle_world():
    s = scenes.vehicle_world(np.random.default_rng(1), 10)
    assert s.nb == 5 and s.nj == 4 and s.ng == 1 + 5 + 32
    assert s.step_ctrl.shape == (10, 4, 4)
    assert np.isclose(s.body_mass[0], 1000.0)
