"""Warm-start kernels (BF16 MFMA convolutions, gradients, Adam, features) on the
MI355X against the convolution oracle (oracle/conv_oracle.py).

Tolerances: FP32 outputs of the convolutions are exact sums of exact BF16
products, accumulated in FP32 in another order than the FP64 oracle: 1e-5
relative to the output scale.  BF16 outputs: one BF16 rounding (2^-8 relative).
"""
import ctypes

import numpy as np
import pytest

import conv_oracle as co
import qmin

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def dev_bf16(a):
    return torch.from_numpy(co.to_bf16_bits(a).view(np.int16).copy()).cuda()


def host_bf16(t):
    return co.from_bf16_bits(t.cpu().numpy().view(np.uint16))


def ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def check(rc):
    assert rc == 0, qmin.library().qm_last_error().decode()


@pytest.mark.parametrize("dims", [(8, 8, 6), (7, 5, 4)])
@pytest.mark.parametrize("cin,cout,dil,relu", [(16, 32, 1, 0), (32, 64, 2, 1), (64, 64, 4, 1), (64, 16, 3, 1),
                                               (16, 16, 1, 1)])
def test_conv_forward_matches_oracle(dims, cin, cout, dil, relu):
    lib = qmin.library()
    rng = np.random.default_rng(cin * 7 + cout + dil)
    n = dims[0] * dims[1] * dims[2]
    x = co.bf16_round(rng.normal(size=(n, cin)))
    w = co.bf16_round(rng.normal(size=(27, cout, cin)) * 0.2)
    b = rng.normal(size=cout).astype(np.float32).astype(np.float64)
    ref = co.forward(x, w, b, dims, dil, relu_in=bool(relu))
    xd, wd = dev_bf16(x), dev_bf16(w)
    bd = torch.from_numpy(b.astype(np.float32)).cuda()
    yf = torch.empty(n * cout, dtype=torch.float32, device="cuda")
    check(lib.qm_ws_conv_forward(ptr(xd), cin, ptr(wd), ptr(bd), cout, None, None, ptr(yf), *dims, dil, relu,
                                 stream()))
    torch.cuda.synchronize()
    got = yf.cpu().numpy().reshape(n, cout)
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-5 * np.abs(ref).max())
    yb = torch.empty(n * cout, dtype=torch.int16, device="cuda")
    check(lib.qm_ws_conv_forward(ptr(xd), cin, ptr(wd), ptr(bd), cout, None, ptr(yb), None, *dims, dil, relu,
                                 stream()))
    torch.cuda.synchronize()
    gotb = host_bf16(yb).reshape(n, cout)
    np.testing.assert_allclose(gotb, ref, rtol=2 ** -8, atol=1e-5 * np.abs(ref).max())


@pytest.mark.parametrize("cin,cout,dil", [(32, 64, 2), (64, 16, 8), (16, 32, 1)])
def test_conv_gradients_match_oracle(cin, cout, dil):
    lib = qmin.library()
    dims = (8, 6, 10)
    n = dims[0] * dims[1] * dims[2]
    rng = np.random.default_rng(cin + cout + 11)
    x = co.bf16_round(rng.normal(size=(n, cin)))        # previous layer's pre-activation
    w = co.bf16_round(rng.normal(size=(27, cout, cin)) * 0.2)
    g = co.bf16_round(rng.normal(size=(n, cout)))
    # weight gradient with the input ReLU
    ref_dw = co.weight_grad(g, x, dims, dil, relu_in=True)
    chunk = 64
    chunks = (n + chunk - 1) // chunk
    part = torch.empty(chunks * 27 * cout * cin, dtype=torch.float32, device="cuda")
    dw = torch.empty(27 * cout * cin, dtype=torch.float32, device="cuda")
    gd, xd = dev_bf16(g), dev_bf16(x)
    check(lib.qm_ws_conv_wgrad(ptr(gd), cout, ptr(xd), cin, 1, *dims, dil, ptr(part), chunk, ptr(dw), stream()))
    # input gradient through the previous layer's ReLU, with the mirrored transposed weights
    wf = torch.from_numpy(w.astype(np.float32).reshape(-1)).cuda()
    wbf = torch.empty(27 * cout * cin, dtype=torch.int16, device="cuda")
    wbd = torch.empty(27 * cout * cin, dtype=torch.int16, device="cuda")
    check(lib.qm_ws_weights_to_bf16(ptr(wf), wf.numel(), ptr(wbf), ptr(wbd), cout, cin, stream()))
    gin = torch.empty(n * cin, dtype=torch.float32, device="cuda")
    check(lib.qm_ws_conv_forward(ptr(gd), cout, ptr(wbd), None, cin, ptr(xd), None, ptr(gin), *dims, dil, 0,
                                 stream()))
    db = torch.empty(cout, dtype=torch.float32, device="cuda")
    bpart = torch.empty(7 * cout, dtype=torch.float32, device="cuda")   # 7 blocks: a ragged last chunk
    check(lib.qm_ws_bias_grad(ptr(gd), n, cout, ptr(bpart), 7, ptr(db), stream()))
    torch.cuda.synchronize()
    np.testing.assert_allclose(dw.cpu().numpy().reshape(27, cout, cin), ref_dw, rtol=0,
                               atol=1e-5 * np.abs(ref_dw).max())
    ref_dx = co.input_grad(g, w, dims, dil, mask=x)
    np.testing.assert_allclose(gin.cpu().numpy().reshape(n, cin), ref_dx, rtol=0, atol=1e-5 * np.abs(ref_dx).max())
    np.testing.assert_allclose(db.cpu().numpy(), g.sum(axis=0), rtol=1e-5, atol=1e-5)
    assert np.array_equal(wbf.cpu().numpy().view(np.uint16), co.to_bf16_bits(w).reshape(-1))


def test_adam_and_loss_kernels():
    lib = qmin.library()
    rng = np.random.default_rng(5)
    cnt = 27 * 16 * 16
    w = rng.normal(size=cnt).astype(np.float32)
    g = rng.normal(size=cnt).astype(np.float32)
    wt, gt = torch.from_numpy(w.copy()).cuda(), torch.from_numpy(g).cuda()
    m, v = torch.zeros(cnt, device="cuda"), torch.zeros(cnt, device="cuda")
    check(lib.qm_ws_adam(ptr(wt), ptr(gt), ptr(m), ptr(v), cnt, 1e-2, 0.9, 0.999, 1e-8, 1, None, None, 16, 16,
                         stream()))
    torch.cuda.synchronize()
    # first Adam step: m = 0.1 g, v = 0.001 g^2, bias-corrected update = lr g / (|g| + eps)
    np.testing.assert_allclose(wt.cpu().numpy(), w - 1e-2 * g / (np.abs(g) + 1e-8), rtol=1e-5, atol=1e-6)
    n = 1000
    y = rng.normal(size=(n, 16)).astype(np.float32)
    t = rng.normal(size=(n, 5)).astype(np.float32)
    types = rng.integers(-1, 2, size=n).astype(np.int8)
    gbuf = torch.empty(n * 16, dtype=torch.int16, device="cuda")
    part = torch.empty(8, dtype=torch.float64, device="cuda")
    yt, tt, ty = torch.from_numpy(y).cuda(), torch.from_numpy(t).cuda(), torch.from_numpy(types).cuda()
    off = np.array([0.1, -0.2, 0.05, 0.0, 0.3])
    check(lib.qm_ws_mse_grad(ptr(yt), 16, ptr(tt), ptr(ty), n, 0.01, off.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                             ptr(gbuf), ptr(part), 8, stream()))
    torch.cuda.synchronize()
    lc = types <= 0
    y5 = y[:, :5] + off.astype(np.float32)
    ref_loss = float(((y5 - t) ** 2)[lc].sum())
    assert part.cpu().numpy().sum() == pytest.approx(ref_loss, rel=1e-5)
    ref_g = np.zeros((n, 16))
    ref_g[:, :5] = 2 * (y5 - t) * 0.01
    ref_g[~lc] = 0
    np.testing.assert_allclose(host_bf16(gbuf).reshape(n, 16), ref_g, rtol=2 ** -8, atol=1e-12)


def test_features_and_apply_round_trip():
    import pyoracle as po
    A, B, C = po.rescaled_phase()
    S0 = po.uniaxial_s0(A, B, C)
    lat = qmin.Lattice(16, 12, 10)
    lat.random_init(4, S0)
    lat.wall(2, 0, qmin.DEGENERATE_PLANAR, 3.0, S0)
    lat.sphere((8.0, 6.0, 5.0), 2.5, qmin.HOMEOTROPIC, 7.0, S0)
    from qmin import warmstart as ws
    d = np.array([1.0, 0.0, 0.0])
    f = host_bf16(ws.lattice_features(lat, d)).reshape(lat.n, 16)
    types = lat.types()
    assert np.array_equal(f[:, :3].sum(axis=1), np.ones(lat.n))
    assert np.array_equal(f[:, 0] == 1, types == 0) and np.array_equal(f[:, 2] == 1, types > 0)
    np.testing.assert_allclose(f[:, 10], co.bf16_round(np.full(lat.n, 2.0 / 3.0)))
    assert np.array_equal(f[:, 8] == 1, (types == 1))      # the wall (object 1) is planar
    q0 = lat.q()
    y = torch.zeros(lat.n * 16, dtype=torch.float32, device="cuda")
    yq = np.tile(np.arange(1, 17, dtype=np.float32) * 0.01, (lat.n, 1))
    y.copy_(torch.from_numpy(yq.reshape(-1)))
    torch.cuda.synchronize()
    off = np.array([0.5, 0.0, -0.25, 0.125, 0.0])
    check(qmin.library().qm_ws_apply(lat.h, ptr(y), 16, off.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
    q1 = lat.q()
    lc = types <= 0
    expect = (yq[0, :5] + off.astype(np.float32)).astype(np.float32)
    np.testing.assert_allclose(q1[lc], np.tile(expect, (lc.sum(), 1)), rtol=1e-7)
    assert np.array_equal(q1[~lc], q0[~lc])
    assert np.abs(lat.v()).max() == 0


def test_energy_term_gradient_is_the_force():
    """qm_ws_energy_grad, the gradient of the loss's energy term: with the
    prediction written into the lattice, the first five of 16 BF16 channels of g
    get -scale F at liquid-crystal sites (F from the oracle at the same Q, so
    one BF16 rounding), object sites and the other channels keep their values,
    and a small step of the prediction along -g lowers the energy."""
    import pyoracle as po
    from qmin import warmstart as ws
    A, B, C = po.rescaled_phase()
    S0 = po.uniaxial_s0(A, B, C)
    shape = (20, 18, 16)
    lat = qmin.Lattice(*shape)
    lat.random_init(8, S0)
    lat.sphere((10.0, 9.0, 8.0), 4.0, qmin.HOMEOTROPIC, 5.3, S0)
    f = qmin.LandauDeGennesLC(A, B, C, 2.32, 0.5, 0.3, semantics=qmin.SEMANTICS_REF_GPU)
    n = lat.n
    rng = np.random.default_rng(3)
    yq = (0.05 * rng.standard_normal((n, 16))).astype(np.float32)
    y = torch.from_numpy(yq.reshape(-1).copy()).cuda()
    off = ws.uniform_q((1.0, 1.0, 1.0), S0)
    offp = off.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    lib = qmin.library()
    torch.cuda.synchronize()
    check(lib.qm_ws_apply(lat.h, ptr(y), 16, offp))
    q = lat.q()
    o = po.OracleSystem(*shape)
    o.random_init(8, S0)
    o.sphere((10.0, 9.0, 8.0), 4.0, 0, 5.3, S0)
    o.set_q(q)
    o.set_params(A, B, C, 2.32, 0.5, 0.3, multi=True, semantics=1)
    fref = o.forces()
    g0 = (0.01 * rng.standard_normal((n, 16))).astype(np.float32)
    g = dev_bf16(g0.reshape(-1))
    scale = 0.37
    check(lib.qm_ws_energy_grad(f.h, lat.h, ctypes.c_float(scale), ptr(g), 16, stream()))
    got = host_bf16(g).reshape(n, 16)
    base = co.bf16_round(g0.reshape(-1)).reshape(n, 16)
    lc = lat.types() <= 0
    want = co.bf16_round((base[:, :5] - np.float32(scale) * fref.astype(np.float32)).astype(np.float32).reshape(-1))
    want = want.reshape(n, 5)
    ulp = np.abs(want) * 2.0 ** -7 + 1e-30
    assert np.all(np.abs(got[lc, :5] - want[lc]) <= ulp[lc])
    assert np.array_equal(got[~lc], base[~lc])
    assert np.array_equal(got[:, 5:], base[:, 5:])
    # descent: the prediction moved along +F (-g without the base) lowers the energy
    e0, _ = f.compute_energy(lat)
    y2 = yq.copy()
    y2[:, :5] += (1e-3 * fref).astype(np.float32)
    y.copy_(torch.from_numpy(y2.reshape(-1)))
    torch.cuda.synchronize()
    check(lib.qm_ws_apply(lat.h, ptr(y), 16, offp))
    e1, _ = f.compute_energy(lat)
    assert e1 < e0


def test_training_reduces_the_loss():
    from qmin import warmstart as ws
    A, B, C, S0 = qmin.workloads.phase_constants()
    force = qmin.LandauDeGennesLC(A, B, C, 2.32, semantics=qmin.SEMANTICS_REF_GPU)
    specs = ws.training_specs(6, size=32, seed=3)
    samples = ws.make_dataset(qmin, specs, force, fc=1e-6, max_iter=400, batch=6)
    net = ws.WarmStartNet(seed=1)
    hist = net.train(samples, epochs=12, lr=3e-3)
    assert np.isfinite(hist).all()
    assert hist[-1] < 0.5 * hist[0], hist


def test_trained_network_beats_random_start():
    """The committed weights on a held-out 64^3 geometry (planar walls,
    homeotropic sphere): FIRE from the prediction reaches the tolerance in far
    fewer iterations than from the seeded random start (a uniform start along the
    same director does as well; the network adds corrections near objects)."""
    import os
    from qmin import warmstart as ws
    A, B, C, S0 = qmin.workloads.phase_constants()
    force = qmin.LandauDeGennesLC(A, B, C, 2.32, 0.5, 0.3, semantics=qmin.SEMANTICS_REF_GPU)
    net = ws.WarmStartNet()
    net.load(os.path.join(os.path.dirname(qmin.__file__), "..", "warmstart", "weights_64.npz"))
    its = {}
    for mode in ("random", "network"):
        lat = qmin.Lattice(64, 64, 64)
        lat.random_init(2718, S0)
        lat.wall(2, 0, qmin.DEGENERATE_PLANAR, 5.3, S0)
        lat.wall(2, 63, qmin.DEGENERATE_PLANAR, 5.3, S0)
        lat.sphere((32.0, 32.0, 32.0), 8.0, qmin.HOMEOTROPIC, 5.3, S0)
        if mode == "network":
            net.predict(lat, (1.0, 0.0, 0.0), S0)
        st = qmin.FIRE(fc=1e-8, max_iter=100000).minimize(force, lat)
        assert st.converged
        its[mode] = st.iterations
    assert its["network"] * 3 < its["random"], its


def _heldout_iterations(seed, mode, L=128):
    """FIRE iterations to forceMax < 1e-8 on the held-out periodic geometry
    warmstart.heldout_geometry(L, seed) (FP32 state, FP64 reductions) from the
    seeded random start, a uniform director along a body diagonal, or the
    committed network's prediction along that diagonal; and the final E/N."""
    import os
    from qmin import warmstart as ws
    A, B, C, S0 = qmin.workloads.phase_constants()
    force = qmin.LandauDeGennesLC(A, B, C, 2.32, 0.5, 0.3, semantics=qmin.SEMANTICS_REF_GPU)
    d = ws.diagonal_director()
    lat = qmin.Lattice(L, L, L, precision=qmin.QM_FP32)
    if mode == "random":
        lat.random_init(4321, S0)
    else:
        lat.uniform_init(d, S0)
    for c, r, t, W in ws.heldout_geometry(L, seed):
        lat.sphere(c, r, t, W, S0)
    if mode == "network":
        net = ws.WarmStartNet()
        net.load(os.path.join(os.path.dirname(qmin.__file__), "..", "warmstart", "weights_64.npz"))
        net.predict(lat, d, S0)
    st = qmin.FIRE(dt=5e-4, dt_max=2e-2, fc=1e-8, max_iter=200000).minimize(force, lat)
    assert st.converged
    e, _ = force.compute_energy(lat)
    return st.iterations, e / lat.n_active()


def test_network_start_beats_random_start_on_heldout_spheres():
    """Held-out 128^3 periodic lattice with five spheres of either anchoring
    (seed 12): FIRE from the committed network's prediction needs under a third
    of the iterations of the seeded random start (measured 4,198 against
    56,613)."""
    print("held-out 128^3, 5 spheres: network against random start", flush=True)
    net_its, _ = _heldout_iterations(12, "network")
    rnd_its, _ = _heldout_iterations(12, "random")
    assert 3 * net_its < rnd_its, (net_its, rnd_its)


def test_network_start_beats_uniform_start_on_heldout_spheres():
    """Held-out 128^3 periodic lattice with six spheres (seed 13): the network's
    corrections near the objects save FIRE iterations over a uniform director
    along the same diagonal (measured 19,513 against 32,112); the two smooth
    starts end in nearby minima (E/N 1.6e-6 relative apart)."""
    print("held-out 128^3, 6 spheres: network against uniform start", flush=True)
    net_its, e_net = _heldout_iterations(13, "network")
    uni_its, e_uni = _heldout_iterations(13, "uniform")
    assert net_its < 0.8 * uni_its, (net_its, uni_its)
    assert abs(e_net - e_uni) <= 1e-5 * abs(e_uni)

