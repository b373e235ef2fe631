"""The random linear sketches the C2 check compares (tests/sketch.py):

  - SplitMix64 against its published output sequence for seed 0
    (0xe220a8397b1dcdaf, 0x6e789e6aa1b965f4, 0x06c45d188009454f: the
    generator's state advances by the golden gamma, so output i is the mix of
    i+1 gammas);
  - weights exact 24-bit values in [-1, 1) with mean 0 and mean square 1/3;
  - a sketch is linear and equals the explicit weighted sum;
  - the relative error of 32 sketches tracks the relative L2 error of the
    field within a small factor, and a sign flip of one high mode (which low
    modes and shell spectra do not see) moves the sketches.
"""
import numpy as np

import sketch


def test_splitmix64_published_sequence():
    g = 0x9E3779B97F4A7C15
    x = np.array([0, g, (2 * g) % 2 ** 64], dtype=np.uint64)
    got = [int(v) for v in sketch.splitmix64(x)]
    assert got == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_weights_are_exact_uniform_values():
    w = sketch.weights(200_000, 4, seed=7)
    assert w.dtype == np.float32 and w.shape == (4, 200_000)
    assert w.min() >= -1.0 and w.max() < 1.0
    q = (w.astype(np.float64) + 1.0) * 2.0 ** 23
    assert np.all(q == np.rint(q)) and q.max() < 2 ** 24
    assert abs(float(w.mean())) < 5e-3
    assert abs(float((w.astype(np.float64) ** 2).mean()) - 1.0 / 3.0) < 5e-3
    # the same seed gives the same weights, another seed others
    assert np.array_equal(w, sketch.weights(200_000, 4, seed=7))
    assert not np.array_equal(w, sketch.weights(200_000, 4, seed=8))


def test_sketch_is_the_weighted_sum_and_linear():
    rng = np.random.default_rng(3)
    shape = (5, 8, 8)
    slots = int(np.prod(shape))
    w = sketch.weights(slots, 6, seed=1)
    a = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    b = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    direct = w.astype(np.float64) @ a.reshape(-1)
    assert np.allclose(sketch.sketch(w, a, chunk=37), direct, rtol=1e-13, atol=1e-13)
    assert np.allclose(sketch.sketch(w, 2 * a - b), 2 * sketch.sketch(w, a) - sketch.sketch(w, b),
                       rtol=1e-12, atol=1e-12)


def test_sketch_error_tracks_field_error():
    rng = np.random.default_rng(5)
    shape = (17, 32, 32)
    slots = int(np.prod(shape))
    w = sketch.weights(slots, 32, seed=7)
    u = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    e = 1e-6 * (rng.standard_normal(shape) + 1j * rng.standard_normal(shape))
    s0 = sketch.sketch(w, u)
    rel_field = np.linalg.norm(e) / np.linalg.norm(u)
    rel_sketch = np.linalg.norm(sketch.sketch(w, u + e) - s0) / np.linalg.norm(s0)
    assert 0.5 < rel_sketch / rel_field < 2.0
    # one high mode with its sign flipped
    v = u.copy()
    v[16, 17, 15] = -v[16, 17, 15]
    moved = np.linalg.norm(sketch.sketch(w, v) - s0) / np.linalg.norm(s0)
    expect = 2 * abs(u[16, 17, 15]) / np.linalg.norm(u)
    assert moved > 0.2 * expect
