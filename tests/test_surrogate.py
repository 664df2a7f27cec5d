"""The surrogate's kernels and network (surrogate.py, fno_kernels.hip).

- The GEMM against numpy on bf16-rounded operands (products are exact in
  FP32, so only the summation order differs): shapes off the 64x64x32 tile
  grid, transposed strides, batches, accumulation; its fused epilogue (bias,
  residual, GELU, stored pre-activation, bf16 copy); the GELU backward
  against torch's autograd; the Adam kernel against torch.optim.Adam.
- Station layout, trace symmetries and the reduced-time alignment of the
  seismogram head.
- Gradients of every parameter of the network, products on the matrix
  cores in BF16, against the same network in FP32 on the CPU (torch
  autograd with plain matmul), within the BF16 tolerance stated below.
- A small training run on FD data from this engine: the trained network
  beats the mechanism-blind baseline (the training set's mean
  log10(PGV/M0) map plus log10 M0), and the gate's fallback answers come
  from the FD engine.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
import surrogate  # noqa: E402


def _bf16(a):
    return torch.from_numpy(a).to(torch.bfloat16)


def test_encoding_and_target_shapes():
    base = {"ni": 20, "nj": 16, "dh": 100.0}
    src = surrogate.source_rows(np.array([[500.0, 600.0, 1500.0, 30.0, 60.0, 90.0, 15.0]]))
    x = surrogate.encode(base, src)
    assert x.shape == (1, 20, 16, surrogate.NCH)
    assert x[0, 5, 6, 4] == pytest.approx(1.0)          # the epicentral bump peaks at the source
    pgv = np.full((1, 3, 20, 16), 1e-3, np.float32)
    t = surrogate.target(pgv, src)
    np.testing.assert_allclose(t, -3.0 - 15.0, atol=1e-5)   # log10(PGV / M0)


def test_dihedral_transforms_are_orthogonal_similarities():
    # each element maps the moment tensor to R M R^T: eigenvalues are kept;
    # positions stay on the grid; the 8 images are distinct, and swapping or
    # mirroring twice is the identity
    base = {"ni": 40, "nj": 40, "dh": 100.0}
    rng = np.random.default_rng(0)
    p = np.array([[1200.0, 1700.0, 2000.0, *rng.uniform([0, 10, -180], [360, 90, 180]), 15.0]])
    src = surrogate.source_rows(p)
    pgv = rng.random((1, 3, 40, 40)).astype(np.float32)

    def mat(r):
        xx, yy, zz, xy, xz, yz = r[3:9]
        return np.array([[xx, xy, xz], [xy, yy, yz], [xz, yz, zz]])

    ev = np.linalg.eigvalsh(mat(src[0]))
    images = set()
    for g in surrogate.D4:
        s2, p2 = surrogate.transform(base, src, pgv, g)
        np.testing.assert_allclose(np.linalg.eigvalsh(mat(s2[0])), ev, atol=1e-12)
        assert p2.shape == pgv.shape and np.isclose(np.sort(p2.ravel()), np.sort(pgv.ravel())).all()
        images.add((round(s2[0, 0]), round(s2[0, 1])) + tuple(np.round(s2[0, 3:9], 9)))
        if g in ((1, 0, 0), (0, 1, 0), (0, 0, 1)):
            s3, p3 = surrogate.transform(base, s2, p2, g)
            np.testing.assert_allclose(s3, src)
            np.testing.assert_array_equal(p3, pgv)
    assert len(images) == 8


@pytest.mark.parametrize("B,H,W,C,m", [(2, 24, 20, 5, 6), (3, 80, 80, 4, 16), (1, 17, 30, 3, 8)])
def test_truncated_dft_matrices_match_fft(B, H, W, C, m):
    """The spectral layer's transform matrices, applied in float64 in the
    layouts the MFMA path uses, give torch.fft's kept modes and, after a
    random complex mixing, its irfft2 of them."""
    torch.manual_seed(B + H + W)
    v = torch.randn(B, H, W, C, dtype=torch.float64)
    fw, gh, gi, fi = [torch.from_numpy(a) for a in surrogate.dft_arrays(H, W, m)]
    left = lambda a, x: torch.einsum("mk,zkn->zmn", a, x)
    y1 = left(fw, v.reshape(B * H, W, C))
    y2 = left(gh, y1.reshape(B, 2 * H, m * C))
    x = y2.reshape(B, 2, 2 * m, m, C).permute(2, 3, 0, 1, 4).reshape(2 * m * m, B, 2 * C)
    ft = torch.fft.rfft2(v, dim=(1, 2))
    sel = torch.cat([ft[:, :m, :m], ft[:, H - m:, :m]], dim=1).permute(1, 2, 0, 3).reshape(2 * m * m, B, C)
    want = torch.cat([sel.real, sel.imag], dim=2)
    assert float((x - want).abs().max()) <= 1e-12 * float(want.abs().max())
    wc = torch.randn(2 * m * m, C, C, dtype=torch.complex128)
    wr = torch.cat([torch.cat([wc.real, wc.imag], 2), torch.cat([-wc.imag, wc.real], 2)], 1)
    z = torch.bmm(x, wr).reshape(2 * m, m, B, 2, C).permute(2, 0, 3, 1, 4).reshape(B, 4 * m, m * C)
    out = left(fi, left(gi, z).reshape(B * H, 2 * m, C)).reshape(B, H, W, C)
    y = torch.einsum("pbc,pcd->pbd", sel, wc).reshape(2 * m, m, B, C).permute(2, 0, 1, 3)
    o = torch.zeros(B, H, W // 2 + 1, C, dtype=torch.complex128)
    o[:, :m, :m], o[:, H - m:, :m] = y[:, :m], y[:, m:]
    ref = torch.fft.irfft2(o, s=(H, W), dim=(1, 2))
    assert float((out - ref).abs().max()) <= 1e-12 * float(ref.abs().max())


def test_dft_matrices_reject_modes_that_do_not_fit():
    with pytest.raises(surrogate.SurrogateError):
        surrogate.dft_arrays(20, 40, 11)     # 2m > H
    with pytest.raises(surrogate.SurrogateError):
        surrogate.dft_arrays(40, 20, 11)     # m > W/2
    fw, gh, gi, fi = surrogate.dft_arrays(40, 20, 10)
    assert fw.shape == (20, 20) and gh.shape == (40, 80) and gi.shape == (80, 40) and fi.shape == (20, 20)


def test_library_missing_raises(monkeypatch):
    monkeypatch.setattr(surrogate, "_LIB", None)
    monkeypatch.setenv("CGFD_FNO_LIB", "/nonexistent/libcgfd_fno.so")
    with pytest.raises(surrogate.SurrogateError):
        surrogate.load_library()


@pytest.mark.gpu
@pytest.mark.parametrize("M,N,K,batch,ta,tb,off", [
    (64, 64, 32, 1, False, False, 0), (70, 33, 45, 3, False, False, 0),
    (130, 96, 200, 2, True, False, 0), (17, 129, 64, 4, False, True, 0),
    (1000, 48, 7, 1, True, True, 0),
    # strides in multiples of 8: 16-byte staging along k (A), along n (B),
    # along m (transposed A) and along k (transposed B), with ragged edges
    (72, 80, 40, 2, False, False, 0), (136, 24, 48, 3, True, True, 0), (9, 200, 1000, 2, False, False, 0),
    # the same shapes from a misaligned base: single-element staging
    (72, 80, 40, 2, False, False, 3), (136, 24, 48, 3, True, True, 5)])
def test_gemm_matches_numpy(M, N, K, batch, ta, tb, off):
    rng = np.random.default_rng(M + N + K)
    a = rng.standard_normal((batch, M, K)).astype(np.float32)
    b = rng.standard_normal((batch, K, N)).astype(np.float32)

    def dev(x):  # on the GPU, off elements into a larger buffer
        t = _bf16(x).cuda()
        if not off:
            return t
        buf = torch.zeros(t.numel() + off, dtype=t.dtype, device="cuda")
        buf[off:] = t.reshape(-1)
        return buf[off:].view(t.shape)

    A = dev(np.ascontiguousarray(a.transpose(0, 2, 1)) if ta else a)
    B = dev(np.ascontiguousarray(b.transpose(0, 2, 1)) if tb else b)
    A = A.transpose(1, 2) if ta else A
    B = B.transpose(1, 2) if tb else B
    out = torch.full((batch, M, N), 0.5, device="cuda")
    surrogate.gemm(A, B, out)
    ar = A.float().cpu().numpy().astype(np.float64)
    br = B.float().cpu().numpy().astype(np.float64)
    want = ar @ br
    got = out.cpu().numpy()
    assert np.linalg.norm(got - want) <= 1e-5 * np.linalg.norm(want)
    surrogate.gemm(A, B, out, accumulate=True)
    assert np.linalg.norm(out.cpu().numpy() - 2 * want) <= 1e-5 * np.linalg.norm(want)


@pytest.mark.gpu
def test_gemm_rejects_bad_operands():
    a = torch.zeros(8, 8, dtype=torch.bfloat16, device="cuda")
    b = torch.zeros(8, 6, dtype=torch.bfloat16, device="cuda")
    with pytest.raises(surrogate.SurrogateError):
        surrogate.gemm(a, b, torch.zeros(8, 8, device="cuda"))           # shape mismatch
    with pytest.raises(surrogate.SurrogateError):
        surrogate.gemm(a.float(), b, torch.zeros(8, 6, device="cuda"))   # not bf16
    with pytest.raises(surrogate.SurrogateError):
        surrogate.gemm(a, b, torch.zeros(6, 8, device="cuda").t())       # output columns not unit-stride


@pytest.mark.gpu
def test_gradients_match_fp32_cpu_reference():
    # BF16 operands carry 8 significant bits: gradients of each parameter
    # agree with the FP32 CPU network to a few percent in relative L2
    net = surrogate.FNO2d(12, 3, width=32, modes=6, layers=4, proj=64, seed=3)
    ref = surrogate.FNO2d(12, 3, width=32, modes=6, layers=4, proj=64, seed=3)
    ref.load_state_dict(net.state_dict())
    net = net.cuda()
    rng = np.random.default_rng(1)
    x = torch.from_numpy(rng.standard_normal((2, 24, 20, 12)).astype(np.float32))
    y = torch.from_numpy(rng.standard_normal((2, 24, 20, 3)).astype(np.float32))
    loss = torch.mean((net(x.cuda(), backend="mfma") - y.cuda()) ** 2)
    loss.backward()
    loss_ref = torch.mean((ref(x, backend="torch") - y) ** 2)
    loss_ref.backward()
    assert abs(float(loss.detach()) - float(loss_ref.detach())) <= 2e-2 * float(loss_ref.detach())
    for (name, p), (_, q) in zip(net.named_parameters(), ref.named_parameters()):
        g, h = p.grad.cpu(), q.grad
        err = float(torch.linalg.vector_norm(g - h) / torch.linalg.vector_norm(h))
        assert err < 5e-2, (name, err)


@pytest.mark.gpu
def test_trained_network_beats_baseline_and_gate_falls_back_to_fd():
    base = surrogate.base_medium(ni=40, nj=40, nk=32, nt=260, dt=0.006, npml=8)
    ptrain = surrogate.draw_sources(96, base, seed=1)
    ptest = surrogate.draw_sources(16, base, seed=2)
    print("FD data for the surrogate test", flush=True)
    ytrain = surrogate.run_fd(base, ptrain, group=32)
    ytest = surrogate.run_fd(base, ptest, group=16)
    s = surrogate.Surrogate(base, width=24, modes=8, layers=4)
    losses = s.fit(ptrain, ytrain, epochs=60, batch=8, lr=3e-3)
    assert losses[59] < 0.3 * losses[0]
    pred, ind = s.predict(ptest)
    t = surrogate.target(ytrain, surrogate.source_rows(ptrain)).mean(0)   # mechanism-blind baseline
    base_pred = np.stack([10.0 ** t.transpose(2, 0, 1) * 10.0 ** p[6] for p in ptest])
    for metric in (surrogate.pgv_error, surrogate.pgv_pointwise_error):
        e, b = np.median(metric(pred, ytest)), np.median(metric(base_pred, ytest))
        print(f"{metric.__name__}: surrogate {e:.3f}, baseline {b:.3f}", flush=True)
        assert e < 0.85 * b
    # a threshold below every indicator sends everything to FD: exact answers
    s.threshold = -1.0
    got, ok = s.answer(ptest[:3], fd_group=3)
    assert not ok.any()
    np.testing.assert_array_equal(got, ytest[:3])


@pytest.mark.gpu
def test_fd_engine_is_symmetric_under_the_transforms():
    # the FD answer for a transformed source is the transformed answer: the
    # x/y swap exactly (x and y share the forward/backward sign of each
    # stage), the mirror of both axes to 0.5 %; the training subgroup SYM is
    # built from these
    base = surrogate.base_medium(ni=40, nj=40, nk=32, nt=200, dt=0.006, npml=8)
    src = surrogate.source_rows(surrogate.draw_sources(1, base, seed=5))
    gs = [g for g in surrogate.SYM if g != (0, 0, 0)] + [(1, 1, 0)]
    pairs = [surrogate.transform(base, src, None, g)[0] for g in gs]
    ys = surrogate.run_fd(base, np.concatenate([src] + pairs), group=len(gs) + 1, rows=True)
    for k, g in enumerate(gs):
        _, want = surrogate.transform(base, src, ys[:1], g)
        err = float(surrogate.pgv_error(ys[k + 1:k + 2], want)[0])
        print(f"symmetry {g}: {err:.4f}", flush=True)
        if g == (1, 0, 0):
            assert err < 1e-6
        elif g in surrogate.SYM:
            assert err < 5e-3, (g, err)


@pytest.mark.gpu
def test_symmetry_averaged_prediction_is_equivariant():
    # averaging over the images of SYM makes the predictor commute with SYM:
    # the prediction for a transformed source is the transformed prediction
    base = surrogate.base_medium(ni=40, nj=40, nk=32, nt=100, dt=0.006, npml=8)
    s = surrogate.Surrogate(base, width=16, modes=6, layers=2)
    s.mu, s.sd, s.tta = -20.0, 1.0, True
    src = surrogate.source_rows(surrogate.draw_sources(3, base, seed=9))
    pred, _ = s.predict(src, rows=True)
    for g in surrogate.SYM[1:]:
        sg, want = surrogate.transform(base, src, pred, g)
        got, _ = s.predict(sg, rows=True)
        assert np.abs(got / want - 1).max() < 1e-4, g


@pytest.mark.gpu
@pytest.mark.parametrize("M,N,K,batch,off,accumulate", [
    (64, 64, 32, 1, 0, False), (70, 33, 45, 3, 0, True), (130, 96, 200, 2, 0, False),
    (1000, 48, 7, 1, 0, True), (72, 80, 40, 2, 3, False)])
def test_gemm_epilogue_matches_numpy(M, N, K, batch, off, accumulate):
    """pre = (C +) A B + bias + D and C = GELU(pre) against float64 numpy on
    the same bf16 operands (GELU from the error function, as torch's), and
    the bf16 copy of C equal to torch's rounding of it."""
    from math import erf, sqrt

    rng = np.random.default_rng(M * N + K)
    a = rng.standard_normal((batch, M, K)).astype(np.float32)
    b = rng.standard_normal((batch, K, N)).astype(np.float32)
    bias = rng.standard_normal(N).astype(np.float32)
    d = rng.standard_normal((batch, M, N)).astype(np.float32)
    c0 = rng.standard_normal((batch, M, N)).astype(np.float32)

    def dev(t):
        if not off:
            return t.cuda()
        buf = torch.zeros(t.numel() + off, dtype=t.dtype, device="cuda")
        buf[off:] = t.reshape(-1).cuda()
        return buf[off:].view(t.shape)

    A, B = dev(_bf16(a)), dev(_bf16(b))
    out = torch.from_numpy(c0).cuda()
    pre = torch.empty_like(out)
    outb = torch.empty(out.shape, dtype=torch.bfloat16, device="cuda")
    surrogate.gemm(A, B, out, accumulate=accumulate, bias=torch.from_numpy(bias).cuda(),
                   addend=torch.from_numpy(d).cuda(), pre=pre, act=surrogate.ACT_GELU, out_bf16=outb)
    assert torch.equal(outb, out.to(torch.bfloat16))  # the bf16 copy: nearest even, as torch rounds
    ar, br = A.float().cpu().numpy().astype(np.float64), B.float().cpu().numpy().astype(np.float64)
    want_pre = ar @ br + bias + d + (c0 if accumulate else 0.0)
    gelu = np.vectorize(lambda v: 0.5 * v * (1.0 + erf(v / sqrt(2.0))))
    scale = np.linalg.norm(want_pre)
    assert np.linalg.norm(pre.cpu().numpy() - want_pre) <= 1e-5 * scale
    assert np.linalg.norm(out.cpu().numpy() - gelu(want_pre)) <= 1e-5 * scale


@pytest.mark.gpu
def test_gelu_backward_matches_torch():
    rng = np.random.default_rng(3)
    x = torch.from_numpy(rng.standard_normal(100003).astype(np.float32) * 3).cuda()
    gy = torch.from_numpy(rng.standard_normal(100003).astype(np.float32)).cuda()
    xr = x.clone().requires_grad_(True)
    torch.nn.functional.gelu(xr).backward(gy)
    gx, gxb = surrogate.gelu_backward(gy, x)
    want = xr.grad
    assert float((gx - want).abs().max()) <= 2e-6 * float(want.abs().max())
    assert torch.equal(gxb, gx.to(torch.bfloat16))     # round to nearest even, as torch
    g2, _ = surrogate.gelu_backward(gy, x, gx=gy.clone(), want_bf16=False)
    assert torch.equal(g2, gx)


@pytest.mark.gpu
def test_adam_matches_torch_adam():
    """Five steps with a changing learning rate over tensors of several sizes
    (one spanning many 8192-element chunks): the project's kernel against
    torch.optim.Adam on the same gradients."""
    rng = np.random.default_rng(4)
    shapes = [(64,), (13, 7), (3, 100, 50), (40000,)]
    ours = [torch.nn.Parameter(torch.from_numpy(rng.standard_normal(s).astype(np.float32)).cuda()) for s in shapes]
    ref = [torch.nn.Parameter(p.detach().clone()) for p in ours]
    opt = surrogate.Adam(ours, lr=3e-3)
    topt = torch.optim.Adam(ref, lr=3e-3)
    for step in range(5):
        lr = surrogate.cosine_lr(3e-3, step, 5)
        grads = [torch.from_numpy(rng.standard_normal(s).astype(np.float32)).cuda() for s in shapes]
        opt.zero_grad()
        for p, g in zip(ours, grads):
            assert float(p.grad.abs().max()) == 0.0  # cleared by the previous step's kernel
            p.grad += g
        for p, g in zip(ref, grads):
            p.grad = g.clone()
        for grp in topt.param_groups:
            grp["lr"] = lr
        opt.step(lr)
        topt.step()
    for p, q in zip(ours, ref):
        assert float((p.detach() - q.detach()).abs().max()) <= 1e-6 * float(q.detach().abs().max()) + 1e-9


def test_station_layout_is_closed_under_the_symmetries():
    base = surrogate.base_medium(ni=192, nj=192, nk=24, nt=8, window=128, stations=32)
    st = {tuple(p) for p in base["stations"]}
    assert len(st) == 32 and base["window"] == (32, 32, 128)
    assert all(38 <= i < 154 and 38 <= j < 154 for i, j in st)
    for i, j in st:
        assert (j, i) in st and (191 - i, 191 - j) in st
    tr = np.random.default_rng(0).standard_normal((2, 32, 3, 5)).astype(np.float32)
    for g in surrogate.SYM:
        t2 = surrogate.transform_traces(base, tr, g)
        np.testing.assert_array_equal(surrogate.transform_traces(base, t2, g), tr)
    # the swap moves station (i, j)'s Vx to station (j, i)'s Vy
    t2 = surrogate.transform_traces(base, tr, (1, 0, 0))
    k = 0
    i, j = base["stations"][k]
    kk = [tuple(p) for p in base["stations"]].index((j, i))
    np.testing.assert_array_equal(t2[:, kk, 1], tr[:, k, 0])


@pytest.mark.gpu
def test_fd_traces_are_symmetric_under_the_transforms():
    # station traces of a transformed source equal the transformed traces:
    # the swap exactly, the mirror of both axes to 1 %
    base = surrogate.base_medium(ni=48, nj=48, nk=32, nt=200, dt=0.006, npml=8, window=40, stations=8)
    src = surrogate.source_rows(surrogate.draw_sources(1, base, seed=6))
    gs = [g for g in surrogate.SYM if g != (0, 0, 0)]
    pairs = [surrogate.transform(base, src, None, g)[0] for g in gs]
    ys, trs = surrogate.run_fd(base, np.concatenate([src] + pairs), group=4, rows=True, traces=True)
    assert trs.shape == (4, 8, 3, 200) and np.abs(trs).max() > 0
    for k, g in enumerate(gs):
        want = surrogate.transform_traces(base, trs[:1], g)
        err = float(surrogate.trace_error(trs[k + 1:k + 2], want)[0])
        print(f"trace symmetry {g}: {err:.4f}", flush=True)
        assert err < (1e-6 if g == (1, 0, 0) else 1e-2), (g, err)


def test_reduced_time_alignment_round_trip_and_symmetry():
    """align_traces shifts each record to start at the P arrival bound
    r / V_ALIGN and scales it by r / 1 km; the inverse restores every sample
    from that bound on (zeros before it).  Distances, hence shifts, are
    unchanged by the symmetry images, so aligned traces transform as the
    raw ones do."""
    base = surrogate.base_medium(ni=192, nj=192, nk=24, nt=200, window=128, stations=32, depth=(5, 20))
    src = surrogate.source_rows(surrogate.draw_sources(3, base, seed=1))
    tr = np.random.default_rng(0).standard_normal((3, 32, 3, 200))
    shift, r = surrogate._station_geometry(base, src)
    back = surrogate.align_traces(base, src, surrogate.align_traces(base, src, tr), inverse=True)
    for n in range(3):
        for q in range(32):
            k = shift[n, q]
            np.testing.assert_allclose(back[n, q, :, k:], tr[n, q, :, k:], rtol=1e-12)
            assert (back[n, q, :, :k] == 0).all()
            assert k == min(int(np.floor(r[n, q] / surrogate.V_ALIGN / base["dt"])), base["nt"] - 1)
    for g in surrogate.SYM:
        sg, _ = surrogate.transform(base, src, None, g)
        a = surrogate.transform_traces(base, surrogate.align_traces(base, src, tr), g)
        b = surrogate.align_traces(base, sg, surrogate.transform_traces(base, tr, g))
        np.testing.assert_allclose(a, b, rtol=1e-12)
