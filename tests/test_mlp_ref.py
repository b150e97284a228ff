"""The FP64 network reference that checks the GPU network kernels: its
analytic gradient against central finite differences, GELU against its
closed form at known points, and one Adam step against the textbook update."""
import numpy as np

import mlp_ref


def test_gelu_known_values():
    assert mlp_ref.gelu(0.0) == 0.0
    # tanh-form GELU approaches identity / zero far from 0
    assert abs(mlp_ref.gelu(10.0) - 10.0) < 1e-12 and abs(mlp_ref.gelu(-10.0)) < 1e-12
    x = np.linspace(-4, 4, 81)
    h = 1e-6
    fd = (mlp_ref.gelu(x + h) - mlp_ref.gelu(x - h)) / (2 * h)
    assert np.max(np.abs(fd - mlp_ref.gelu_grad(x))) < 1e-8


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    net = mlp_ref.Mlp(7, 6, 3, 4)
    p = rng.normal(0, 0.5, net.n_params())
    x = rng.normal(size=(5, 7))
    t = rng.normal(size=(5, 4))
    _, g = net.loss_grad(p, x, t)
    h = 1e-6
    fd = np.empty_like(p)
    for i in range(len(p)):
        e = np.zeros_like(p)
        e[i] = h
        fd[i] = (net.loss(p + e, x, t) - net.loss(p - e, x, t)) / (2 * h)
    assert np.max(np.abs(fd - g)) < 1e-8 * max(1.0, np.max(np.abs(g)))


def test_adam_step_is_the_textbook_update():
    p = np.array([1.0, -2.0])
    g = np.array([0.5, 0.25])
    p1, m, v = mlp_ref.adam_step(p, g, np.zeros(2), np.zeros(2), 1, lr=0.1)
    # first bias-corrected step moves every coordinate by lr * sign(g)
    assert np.allclose(p1, p - 0.1 * np.sign(g), atol=1e-7)
