"""Pin the local-PSF FFT plan restatement to a direct sum over kernel offsets (FP64)."""

import itertools

import numpy as np
import pytest

from oracle import local_fft_np as L


def case(seed, shape=(13, 11, 4), nodes=(3, 3), kshape=(5, 3, 3), plane=False):
    rng = np.random.default_rng(seed)
    X, Y, Z = shape
    kz = Z if plane else kshape[2]
    kernels = rng.random((nodes[0], nodes[1], kshape[0], kshape[1], kz))
    xs = np.linspace(0, X, nodes[0])
    ys = np.linspace(0, Y, nodes[1])
    w = (L.axis_weights(xs, np.arange(X) + 0.5), L.axis_weights(ys, np.arange(Y) + 0.5))
    return kernels, w, rng.random(shape)


@pytest.mark.parametrize("plane,mode", list(itertools.product((False, True),
                                                              ("detection", "excitation", "native"))))
def test_tiled_fft_equals_direct_sum(plane, mode):
    det, corr = mode == "detection", mode == "excitation"
    kernels, w, src = case(7, plane=plane)
    a = L.local_fft(src.shape, kernels, w, src, plane=plane, detection=det, correlate=corr)
    b = L.local_direct(src.shape, kernels, w, src, plane=plane, detection=det, correlate=corr)
    assert a.shape == (b.shape)
    np.testing.assert_allclose(a, b, rtol=1e-11, atol=1e-11)


def test_axis_weights_partition_of_unity_and_nodes():
    nodes = np.array([0.0, 4.0, 10.0])
    pos = np.linspace(0.0, 10.0, 41)
    w = L.axis_weights(nodes, pos)
    np.testing.assert_allclose(w.sum(0), 1.0)
    assert (w >= 0).all()
    for k, x in enumerate(nodes):
        col = np.argmin(abs(pos - x))
        assert w[k, col] == pytest.approx(1.0)


def test_even_kernel_offsets_and_single_node_support():
    # even kernel sides move the centre (K//2 versus K-1-K//2); a node whose weights vanish
    # contributes nothing
    kernels, w, src = case(3, shape=(10, 9, 3), nodes=(2, 3), kshape=(4, 2, 2))
    w = (w[0], w[1] * (np.arange(3)[:, None] != 1))
    for det, corr in ((True, False), (False, True), (False, False)):
        a = L.local_fft(src.shape, kernels, w, src, plane=False, detection=det, correlate=corr)
        b = L.local_direct(src.shape, kernels, w, src, plane=False, detection=det, correlate=corr)
        np.testing.assert_allclose(a, b, rtol=1e-11, atol=1e-11)


def _uniform(op, kshape=(5, 3, 5), shape=(12, 10, 5), seed=30, depths=None, native=False):
    from oracle.rana_np import Operator
    rng = np.random.default_rng(seed)
    k = rng.random(kshape)
    kern = np.broadcast_to(k, (3, 2) + kshape).copy()
    kl = [kern] if depths is None else [kern, kern.copy()]
    o = Operator(kl, [0.0, 3.0, 6.0], [0.0, 5.0], (0.5, 0.5, 1.0), op, depths_um=depths,
                 native=native)
    return o, k, rng.random(shape)


@pytest.mark.parametrize("op,native", [("detection", False), ("excitation", False),
                                       ("excitation", True)])
def test_operator_with_one_kernel_is_matlab_convn(op, native):
    import oracle as O
    o, k, vol = _uniform(op, native=native)
    kk = np.flip(k) if (op == "excitation" and not native) else k
    np.testing.assert_allclose(o.static(vol), np.maximum(O.zstack(vol, kk), 0), rtol=1e-11,
                               atol=1e-12)
    # a plane projection pairs kernel z with source z: detection reverses the kernel depth
    # first (a convolution in z), excitation correlates, native pairs them directly, which is
    # the centre plane of a convolution with the kernel flipped in z
    c = k.shape[2] // 2
    ks = k[:, :, ::-1] if native else kk
    np.testing.assert_allclose(o.scan_slab(vol), np.maximum(O.zstack(vol, ks)[:, :, c], 0),
                               rtol=1e-11, atol=1e-12)


def test_depth_bank_of_equal_fields_is_the_field():
    o1, _, vol = _uniform("excitation")
    o2, _, _ = _uniform("excitation", depths=[0.0, 10.0])
    np.testing.assert_allclose(o2.static(vol), o1.static(vol), rtol=1e-12, atol=1e-13)
    np.testing.assert_allclose(o2.scan_slab(vol, z_start=1), o1.scan_slab(vol, z_start=1),
                               rtol=1e-12, atol=1e-13)
