"""Bead-image model from a radial Z-LUT (QTrkLUTModel*): forward, adjoint, Poisson MLE, Fisher/CRLB.

Pins: the forward pass equals the oracle's GenerateImageFromLUT restatement
bit for bit; the Jacobian and the adjoint agree with central finite
differences (h = 1e-3) of the FP64 oracle; the 5x5 Fisher matrix agrees with
the FP64 CPU value built from finite-difference derivatives, and its (x, y, z)
block with the reference's SampleFisherMatrix convention; the CRLB is the
square root of the diagonal of the inverse; the fit recovers noise-free
parameters, and on Poisson data its errors scale with the CRLB.
"""
import numpy as np
import pytest

from qtrk_amd import api, synth

W = H = 48
MAXR = 24.0


@pytest.fixture(scope="module")
def lut():
    return synth.model_lut(200, 48, MAXR)


def oracle_mu(o, lut, th):
    g = o.generate_image_from_lut(W, H, lut, 0.0, MAXR, float(th[0]), float(th[1]), float(th[2]))
    return th[4] + th[3] * g.astype(np.float64)


def fd_jacobian(o64, lut, th, h=1e-3):
    J = np.zeros((5, H, W))
    for k in range(3):
        tp, tm = th.astype(np.float64).copy(), th.astype(np.float64).copy()
        tp[k] += h
        tm[k] -= h
        J[k] = (oracle_mu(o64, lut, tp) - oracle_mu(o64, lut, tm)) / (2 * h)
    J[3] = o64.generate_image_from_lut(W, H, lut, 0.0, MAXR, *map(float, th[:3]))
    J[4] = 1.0
    return J


def test_oracle_fisher_closed_form(oracle64):
    """SampleFisherMatrix restatement on a LUT constant in r and linear in z: g = a + b z on every
    pixel, so F_zz = (maxValue / mean g) * N b^2 / (1e-5 + g) and every x/y element is 0."""
    planes, steps, a, b = 40, 24, 0.5, 2.0 ** -7  # rows exactly representable in float32
    L = (a + b * np.arange(planes, dtype=np.float64))[:, None] * np.ones((1, steps))
    z, mv = 17.3, 1000.0
    F = oracle64.fisher_ref(16, 16, L.astype(np.float32), 0.0, 12.0, (8.0, 8.0, z), (1e-3, 1e-3, 1e-3), mv)
    g = a + b * z
    eps = float(np.float32(1e-5))  # the reference adds the float constant 1e-5f
    assert F[2, 2] == pytest.approx(mv / g * 256 * b * b / (eps + g), rel=1e-9)
    assert np.all(np.abs(F[:2]) < 1e-9) and np.all(np.abs(F[:, :2]) < 1e-9)


@pytest.mark.gpu
def test_forward_bit_exact(oracle, lut, gpu):
    import torch
    rng = np.random.default_rng(11)
    th = synth.model_params(24, rng, W)
    th[0, :3] = (24.0, 24.0, 100.0)      # centred, on a plane
    th[1, 2] = 0.4                       # below the first spline plane
    th[2, 2] = 198.7                     # above the last
    m = api.LUTModel(lut, 0.0, MAXR, W, H)
    try:
        mu = m.forward(torch.from_numpy(th)).cpu().numpy()
        for i in range(len(th)):
            g = oracle.generate_image_from_lut(W, H, lut, 0.0, MAXR, *map(float, th[i, :3]))
            ref = np.float32(th[i, 4]) + np.float32(th[i, 3]) * g
            assert np.array_equal(mu[i], ref), f"ROI {i}: max diff {np.max(np.abs(mu[i] - ref))}"
    finally:
        m.free()


@pytest.mark.gpu
@pytest.mark.parametrize("ovs,spline", [(1, True), (3, True), (8, True), (4, False)])
def test_forward_bit_exact_sub_sampling_and_linear(oracle, lut, gpu, ovs, spline):
    """Other sub-sample counts, and linear interpolation between planes."""
    import torch
    rng = np.random.default_rng(12 + ovs)
    th = synth.model_params(8, rng, W)
    th[0, 2] = 198.7
    m = api.LUTModel(lut, 0.0, MAXR, W, H, ovs=ovs, spline=spline)
    try:
        mu = m.forward(torch.from_numpy(th)).cpu().numpy()
        for i in range(len(th)):
            g = oracle.generate_image_from_lut(W, H, lut, 0.0, MAXR, *map(float, th[i, :3]), spline=spline, ovs=ovs)
            ref = np.float32(th[i, 4]) + np.float32(th[i, 3]) * g
            assert np.array_equal(mu[i], ref), f"ROI {i}: max diff {np.max(np.abs(mu[i] - ref))}"
    finally:
        m.free()


@pytest.mark.gpu
def test_jacobian_and_adjoint_vs_finite_differences(oracle64, lut, gpu):
    import torch
    rng = np.random.default_rng(12)
    th = synth.model_params(6, rng, W)
    m = api.LUTModel(lut, 0.0, MAXR, W, H)
    try:
        mu, J = m.forward(torch.from_numpy(th), jacobian=True)
        J = J.cpu().numpy().astype(np.float64)
        gmu = torch.from_numpy(rng.standard_normal((len(th), H, W)).astype(np.float32))
        grad = m.backward(torch.from_numpy(th), gmu).cpu().numpy()
        for i in range(len(th)):
            # h = 1e-3 differences straddle kinks of the linear-in-r interpolation (up to ~0.3% in
            # x/y); h = 1e-6 in FP64 resolves the analytic derivative to its float precision, which
            # the cancelling x/y-weighted sub-sample sums limit to a few 1e-4 of the norm
            Jf = fd_jacobian(oracle64, lut, th[i])
            Jt = fd_jacobian(oracle64, lut, th[i], h=1e-6)
            for k in range(5):
                rel = np.linalg.norm(J[i, k] - Jf[k]) / np.linalg.norm(Jf[k])
                rel_t = np.linalg.norm(J[i, k] - Jt[k]) / np.linalg.norm(Jt[k])
                assert rel < 1e-2 and rel_t < 5e-4, (i, k, rel, rel_t)
            # adjoint = J^T gmu, and the directional derivative of <gmu, mu(theta)>
            gf = np.einsum("khw,hw->k", Jt, gmu[i].numpy().astype(np.float64))
            assert np.allclose(grad[i], np.einsum("khw,hw->k", J[i], gmu[i].numpy()), rtol=1e-4, atol=1e-3)
            assert np.allclose(grad[i], gf, rtol=1e-3, atol=1e-4 * np.abs(Jt).sum((1, 2)).max())
    finally:
        m.free()


@pytest.mark.gpu
def test_autograd_uses_adjoint(lut, gpu):
    import torch
    rng = np.random.default_rng(13)
    th = torch.from_numpy(synth.model_params(3, rng, W)).cuda().requires_grad_(True)
    m = api.LUTModel(lut, 0.0, MAXR, W, H)
    try:
        w = torch.from_numpy(rng.standard_normal((3, H, W)).astype(np.float32)).cuda()
        loss = (m.render(th) * w).sum()
        loss.backward()
        assert torch.allclose(th.grad, m.backward(th.detach(), w), rtol=1e-6, atol=1e-6)
    finally:
        m.free()


@pytest.mark.gpu
def test_fisher_crlb_vs_fp64(oracle64, lut, gpu):
    import torch
    rng = np.random.default_rng(14)
    th = synth.model_params(5, rng, W)
    m = api.LUTModel(lut, 0.0, MAXR, W, H)
    try:
        data = m.forward(torch.from_numpy(th))
        th1, F, crlb, _ = m.fit(data, torch.from_numpy(th), iterations=0)
        F, crlb = F.cpu().numpy(), crlb.cpu().numpy()
        assert np.array_equal(th1.cpu().numpy(), th)
        for i in range(len(th)):
            Jf = fd_jacobian(oracle64, lut, th[i], h=1e-6).reshape(5, -1)
            mu = oracle_mu(oracle64, lut, th[i]).reshape(-1)
            Fc = (Jf / mu) @ Jf.T
            assert np.allclose(F[i], Fc, rtol=1e-4, atol=1e-4 * np.sqrt(np.outer(np.diag(Fc), np.diag(Fc)))), i
            assert np.allclose(crlb[i], np.sqrt(np.diag(np.linalg.inv(F[i]))), rtol=1e-6)
            print(f"ROI {i}: CRLB x {crlb[i, 0]:.4f} px  y {crlb[i, 1]:.4f} px  z {crlb[i, 2]:.4f} planes  "
                  f"I {crlb[i, 3]:.1f}  bg {crlb[i, 4]:.3f}")
        # (x, y, z) block against SampleFisherMatrix: mu = I g with I = maxValue / mean(g), no background
        mv = 3000.0
        t = th[0].copy()
        g = oracle64.generate_image_from_lut(W, H, lut, 0.0, MAXR, *map(float, t[:3]))
        t[3], t[4] = mv / g.mean(), 0.0
        _, F0, _, _ = m.fit(m.forward(torch.from_numpy(t[None])), torch.from_numpy(t[None]), iterations=0)
        # SampleFisherMatrix adds 1e-5 to the image, which matters at this LUT's dark ring minima;
        # its maxValue / mean(mu) scaling makes the result independent of the LUT's scale otherwise,
        # so the reference sees the LUT x 1000, where the 1e-5 term is negligible
        Fr = oracle64.fisher_ref(W, H, lut * 1000, 0.0, MAXR, tuple(map(float, t[:3])), (1e-3, 1e-3, 1e-3), mv)
        F0 = F0.cpu().numpy()[0, :3, :3]
        scale = np.sqrt(np.outer(np.diag(Fr), np.diag(Fr)))
        # the reference differentiates with h = 1e-3, which straddles the r-interpolation kinks (~0.3%)
        assert np.allclose(F0, Fr, rtol=1e-2, atol=1e-2 * scale), (F0, Fr)
    finally:
        m.free()


@pytest.mark.gpu
def test_fit_noise_free_recovers_truth(lut, gpu):
    import torch
    rng = np.random.default_rng(15)
    th = synth.model_params(64, rng, W)
    m = api.LUTModel(lut, 0.0, MAXR, W, H)
    try:
        data = m.forward(torch.from_numpy(th))
        t0 = th.copy()
        t0[:, :2] += rng.uniform(-0.4, 0.4, (64, 2)).astype(np.float32)
        t0[:, 2] += rng.uniform(-1.5, 1.5, 64).astype(np.float32)
        t0[:, 3] *= 1.15
        t0[:, 4] *= 0.8
        fit, _, _, _ = m.fit(data, torch.from_numpy(t0), iterations=10)
        fit = fit.cpu().numpy()
        assert np.max(np.abs(fit[:, :2] - th[:, :2])) < 1e-3
        assert np.max(np.abs(fit[:, 2] - th[:, 2])) < 1e-2
        assert np.max(np.abs(fit[:, 3] / th[:, 3] - 1)) < 1e-3
        assert np.max(np.abs(fit[:, 4] - th[:, 4])) < 1e-2
    finally:
        m.free()


@pytest.mark.gpu
@pytest.mark.parametrize("ovs", [1, 3, 8])
def test_fit_noise_free_other_sub_sampling(lut, gpu, ovs):
    """The fit's packed row table at the smallest, an odd and the largest sub-sample count it takes."""
    import torch
    rng = np.random.default_rng(40 + ovs)
    th = synth.model_params(32, rng, W)
    m = api.LUTModel(lut, 0.0, MAXR, W, H, ovs=ovs)
    try:
        data = m.forward(torch.from_numpy(th))
        t0 = th.copy()
        t0[:, :2] += rng.uniform(-0.3, 0.3, (32, 2)).astype(np.float32)
        t0[:, 2] += rng.uniform(-1.0, 1.0, 32).astype(np.float32)
        fit, _, _, _ = m.fit(data, torch.from_numpy(t0), iterations=10)
        fit = fit.cpu().numpy()
        assert np.max(np.abs(fit[:, :2] - th[:, :2])) < 1e-3
        assert np.max(np.abs(fit[:, 2] - th[:, 2])) < 1e-2
    finally:
        m.free()


@pytest.mark.gpu
@pytest.mark.parametrize("size,ovs", [(48, 4), (100, 3)])
def test_fit_lanes_per_roi_agree(lut, gpu, size, ovs):
    """The fit with 64 to 1024 lanes per ROI (QTrkLUTModelSetFitLanes): the lanes only split the pixel
    sums differently, so on Poisson data every lane count lands on the 256-lane fit to float rounding,
    from a given start and from the least-squares start (I <= 0)."""
    import torch
    rng = np.random.default_rng(77 + size)
    n = 48
    th = synth.model_params(n, rng, size)
    m = api.LUTModel(lut, 0.0, MAXR * size / W, size, size, ovs=ovs)
    try:
        mu = m.forward(torch.from_numpy(th)).cpu().numpy()
        data = torch.from_numpy(rng.poisson(mu).astype(np.float32))
        t0 = th.copy()
        t0[:, :2] += rng.uniform(-0.3, 0.3, (n, 2)).astype(np.float32)
        t0[:, 2] += rng.uniform(-1.0, 1.0, n).astype(np.float32)
        t0[::3, 3] = -1.0
        out = {}
        wide = (512, 1024)
        for lanes in (64, 256) + wide:
            m.set_fit_lanes(lanes)
            fit, _, crlb, nll = m.fit(data, torch.from_numpy(t0), iterations=8)
            out[lanes] = (fit.cpu().numpy(), crlb.cpu().numpy(), nll.cpu().numpy())
        ref = out[256]
        for lanes, (fit, crlb, nll) in out.items():
            d = np.abs(fit - ref[0])
            print(f"{size}px lanes {lanes}: max |dx| {d[:, 0].max():.2e} |dz| {d[:, 2].max():.2e} "
                  f"rel dI {np.max(d[:, 3] / ref[0][:, 3]):.2e}")
            assert np.max(d[:, :2]) < 1e-3
            assert np.max(d[:, 2]) < 1e-3
            assert np.max(d[:, 3] / ref[0][:, 3]) < 1e-3
            np.testing.assert_allclose(crlb, ref[1], rtol=1e-3)
            np.testing.assert_allclose(nll, ref[2], rtol=1e-5)
        # the widest fit still recovers noise-free parameters
        m.set_fit_lanes(wide[-1])
        clean = m.forward(torch.from_numpy(th))
        fit, _, _, _ = m.fit(clean, torch.from_numpy(t0), iterations=10)
        fit = fit.cpu().numpy()
        assert np.max(np.abs(fit[:, :2] - th[:, :2])) < 1e-3
        assert np.max(np.abs(fit[:, 2] - th[:, 2])) < 1e-2
    finally:
        m.free()


@pytest.mark.gpu
def test_fit_lanes_rejects_unsupported(lut, gpu):
    for ovs, bad, good in ((7, 512, 256), (5, 1024, 512), (4, 300, 1024)):
        m = api.LUTModel(lut, 0.0, MAXR, W, H, ovs=ovs)
        try:
            with pytest.raises(api.QTrkError, match="lanes per ROI"):
                m.set_fit_lanes(bad)
            m.set_fit_lanes(good)
        finally:
            m.free()


@pytest.mark.gpu
def test_fit_rejects_more_than_8_sub_samples(lut, gpu):
    import torch
    th = synth.model_params(4, np.random.default_rng(3), W)
    m = api.LUTModel(lut, 0.0, MAXR, W, H, ovs=9)
    try:
        data = m.forward(torch.from_numpy(th))          # rendering takes any count
        with pytest.raises(api.QTrkError, match="at most 8 sub-samples"):
            m.fit(data, torch.from_numpy(th))
    finally:
        m.free()


@pytest.mark.gpu
def test_fit_precision_matches_crlb(lut, gpu):
    """The differentiation workload's precision bar: 16,384 Poisson ROIs of 48x48 (x, y +-3 px, z
    20-180, I 2,000-20,000, bg 5-50), 10 iterations from the tracker's COM + QI + Z-LUT estimate.
    The errors divided by each ROI's CRLB have a spread within 5% of 1 for all of x, y, z, I and bg
    (the sample standard deviation's own standard error is 1/sqrt(2n) = 0.0055 here), and no bias
    beyond 0.05.  Fits whose start fell outside the likelihood's capture range (an error above 6
    CRLB in a parameter) are counted, at most 0.2% of the ROIs, and left out of the spread."""
    import torch
    rng = np.random.default_rng(16)
    n = 16384
    th = synth.model_params(n, rng, W)
    m = api.LUTModel(lut, 0.0, MAXR, W, H)
    try:
        gen = torch.Generator(device="cuda")
        gen.manual_seed(16)
        data = torch.poisson(m.forward(torch.from_numpy(th)), generator=gen)
        t0 = synth.mle_start(m, lut, data)
        fit, _, crlb, _ = m.fit(data, t0, iterations=10)
        fit, crlb = fit.cpu().numpy(), crlb.cpu().numpy()
        ok = np.all(np.isfinite(crlb), axis=1) & np.all(crlb > 0, axis=1)
        e = (fit[ok] - th[ok]) / crlb[ok]
        inl = np.all(np.abs(e) <= 6, axis=1)
        sd, mu = e[inl].std(axis=0), e[inl].mean(axis=0)
        print(f"{int((~ok).sum())} without a CRLB, {int((~inl).sum())} beyond 6 CRLB of {n}; "
              "std of error / CRLB (x, y, z, I, bg):", np.round(sd, 4), " mean:", np.round(mu, 4), flush=True)
        assert ok.mean() >= 0.999 and inl.mean() >= 0.998
        assert np.all(np.abs(sd - 1) < 0.05), sd
        assert np.all(np.abs(mu) < 0.05), mu
    finally:
        m.free()


@pytest.mark.gpu
def test_model_rejects_bad_input(lut, gpu):
    with pytest.raises(api.QTrkError, match="too few planes"):
        api.LUTModel(lut[:3], 0.0, MAXR, W, H)
    with pytest.raises(api.QTrkError, match="maxradius"):
        api.LUTModel(lut, 5.0, 5.0, W, H)
