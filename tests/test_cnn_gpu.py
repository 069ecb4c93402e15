"""BF16 MFMA CNN kernels against the CPU restatement (oracle/cnn.py): GEMM in
every operand layout and epilogue, im2col / col2im, the whole forward and
backward pass, the pose loss, Adam, and the network input pipeline."""

import numpy as np
import pytest
import torch

from cryojax_amd import _lib, cnn, ops
from oracle import cnn as C
from tests.helpers import draw, packed_params

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def bf16_dev(x):
    """float array -> (bf16-rounded float64 host copy, int16 device bit patterns)."""
    r = C.bf16_round(np.asarray(x, dtype=np.float32))
    bits = (r.view(np.uint32) >> 16).astype(np.uint16).view(np.int16)
    return r.astype(np.float64), torch.from_numpy(bits.copy()).to(DEV)


def host(t16):
    return cnn.from_bf16(t16).cpu().numpy().astype(np.float64)


def rel(a, b):
    return np.linalg.norm(np.ravel(a - b)) / max(np.linalg.norm(np.ravel(b)), 1e-30)


@pytest.mark.parametrize("a_k,b_k", [(True, True), (True, False), (False, True), (False, False)])
@pytest.mark.parametrize("m,n,k", [(72, 40, 72), (128, 96, 512), (8, 8, 8), (200, 136, 40), (70, 36, 72)])
def test_gemm_layouts(rng, a_k, b_k, m, n, k):
    if (not a_k and m % 8) or (not b_k and n % 8):
        pytest.skip("a row-contiguous operand needs its contiguous extent to be a multiple of 8")
    print("case: gemm", a_k, b_k, m, n, k, flush=True)
    A, a16 = bf16_dev(rng.standard_normal((m, k)) if a_k else rng.standard_normal((k, m)))
    B, b16 = bf16_dev(rng.standard_normal((n, k)) if b_k else rng.standard_normal((k, n)))
    Am = A if a_k else A.T
    Bm = B.T if b_k else B
    ref = Am @ Bm
    c = torch.zeros((m, n), dtype=torch.float32, device=DEV)
    cnn.gemm(a16, k if a_k else 1, 1 if a_k else m, b16, 1 if b_k else n, k if b_k else 1, m, n, k, c32=c, ldc32=n)
    assert rel(c.cpu().numpy(), ref) <= 1e-5


def test_gemm_epilogues_and_split_k(rng):
    m, n, k = 96, 72, 4096
    A, a16 = bf16_dev(rng.standard_normal((m, k)))
    B, b16 = bf16_dev(rng.standard_normal((n, k)))
    bias = rng.standard_normal(n).astype(np.float32)
    Mk, m16 = bf16_dev(rng.standard_normal((m, n)))
    ref = A @ B.T + bias
    c32 = torch.empty((m, n), dtype=torch.float32, device=DEV)
    c16 = torch.empty((m, n), dtype=torch.int16, device=DEV)
    cnn.gemm(a16, k, 1, b16, 1, k, m, n, k, c32=c32, ldc32=n, c16=c16, ldc16=n,
             bias=torch.from_numpy(bias).to(DEV), mask=m16, ld_mask=n, relu=True)
    want = np.where(Mk > 0, np.maximum(ref, 0), 0)
    assert rel(c32.cpu().numpy(), want) <= 1e-5
    np.testing.assert_array_equal(host(c16), C.bf16_round(c32.cpu().numpy()).astype(np.float64))
    acc = torch.full((m, n), 1.5, dtype=torch.float32, device=DEV)
    cnn.gemm(a16, k, 1, b16, 1, k, m, n, k, c32=acc, ldc32=n, accumulate=True, split_k=7, alpha=0.5)
    assert rel(acc.cpu().numpy(), 1.5 + 0.5 * (A @ B.T)) <= 1e-5


def test_gemm_rejects_bad_layouts(rng):
    _, a16 = bf16_dev(rng.standard_normal((16, 16)))
    c = torch.empty((16, 16), dtype=torch.float32, device=DEV)
    with pytest.raises(ValueError):
        cnn.gemm(a16, 16, 1, a16, 1, 16, 16, 16, 12, c32=c, ldc32=16)  # k % 8
    with pytest.raises(ValueError):
        cnn.gemm(a16, 16, 2, a16, 1, 16, 16, 16, 16, c32=c, ldc32=16)  # no unit stride
    with pytest.raises(ValueError):
        cnn.gemm(a16, 16, 1, a16, 1, 16, 16, 16, 16, c32=c, ldc32=16, split_k=2)  # split without accumulate


@pytest.mark.parametrize("c,kp,h", [(1, 16, 64), (1, 24, 16), (3, 32, 16)])
def test_im2col_padded_columns(rng, c, kp, h):
    """c = 1 with kp = 16 (the first layer's own kernel) at the network's 64 x 64 input, and widths
    beyond kpad(c) through the generic kernel: taps in order (dy, dx), zero columns after 9 c."""
    lib = _lib.load()
    n = 5
    X, x16 = bf16_dev(rng.standard_normal((n, h, h, c)))
    col = torch.empty((n * (h // 2) ** 2, kp), dtype=torch.int16, device=DEV)
    st = torch.cuda.current_stream().cuda_stream
    _lib.check(lib.cryomi_im2col_s2(x16.data_ptr(), n, h, h, c, kp, col.data_ptr(), st), "im2col")
    np.testing.assert_array_equal(host(col), C.im2col(X, kp))


# c % 8 == 0 takes the 16-byte channel-group path, other c the per-element one
@pytest.mark.parametrize("c", [1, 3, 32, 64])
def test_im2col_col2im(rng, c):
    lib = _lib.load()
    n, h = 3, 16
    kp = cnn.kpad(c)
    X, x16 = bf16_dev(rng.standard_normal((n, h, h, c)))
    col = torch.empty((n * (h // 2) ** 2, kp), dtype=torch.int16, device=DEV)
    st = torch.cuda.current_stream().cuda_stream
    _lib.check(lib.cryomi_im2col_s2(x16.data_ptr(), n, h, h, c, kp, col.data_ptr(), st), "im2col")
    np.testing.assert_array_equal(host(col), C.im2col(X, kp))
    D = rng.standard_normal((n * (h // 2) ** 2, kp)).astype(np.float32)
    dx = torch.empty((n, h, h, c), dtype=torch.int16, device=DEV)
    _lib.check(lib.cryomi_col2im_s2(torch.from_numpy(D).to(DEV).data_ptr(), x16.data_ptr(), n, h, h, c, kp,
                                    dx.data_ptr(), st), "col2im")
    want = C.bf16_round((C.col2im(D.astype(np.float64), (n, h, h, c)) * (X > 0)).astype(np.float32))
    assert np.max(np.abs(host(dx) - want)) <= 2**-7 * np.max(np.abs(want))
    # no ReLU mask (x = NULL)
    _lib.check(lib.cryomi_col2im_s2(torch.from_numpy(D).to(DEV).data_ptr(), None, n, h, h, c, kp, dx.data_ptr(), st),
               "col2im")
    want = C.bf16_round(C.col2im(D.astype(np.float64), (n, h, h, c)).astype(np.float32))
    assert np.max(np.abs(host(dx) - want)) <= 2**-7 * np.max(np.abs(want))


@pytest.mark.parametrize("m,n,pad", [(1000, 96, 8), (64, 32, 8), (5, 512, 8), (200000, 32, 8), (3, 8, 0),
                                     (40, 2304, 8), (1000, 30, 8), (500, 96, 3)])
def test_column_sums(rng, m, n, pad):
    """Widths that are multiples of 8 on 16-byte rows take the 16-byte-per-lane kernel (with
    column windows of 2048 beyond that width); the others take the lane-per-column kernel."""
    lib = _lib.load()
    ld = n + pad
    Y, y16 = bf16_dev(rng.standard_normal((m, ld)))
    grad = torch.full((n,), 0.5, dtype=torch.float32, device=DEV)
    st = torch.cuda.current_stream().cuda_stream
    _lib.check(lib.cryomi_column_sums_bf16(y16.data_ptr(), m, n, ld, grad.data_ptr(), st), "column_sums")
    want = 0.5 + Y[:, :n].astype(np.float64).sum(axis=0)
    # FP32 partial sums: error bounded relative to the column's sum of magnitudes
    bound = 2e-6 * np.abs(Y[:, :n]).sum(axis=0) + 1e-5
    assert np.all(np.abs(grad.cpu().numpy() - want) <= bound)


@pytest.fixture(scope="module")
def net_and_batch():
    rng = np.random.default_rng(11)
    params = C.init_params(5)
    for k in params:
        if k.endswith(".b"):
            params[k] = (rng.standard_normal(params[k].shape) * 0.05).astype(np.float32)
    net = cnn.PoseNet(DEV, params=params)
    n = 16
    X, x16 = bf16_dev(rng.standard_normal((n, 64, 64)))
    return net, params, X, x16, n, rng


def test_forward_matches_oracle(net_and_batch):
    net, params, X, x16, n, _ = net_and_batch
    out = net.forward(x16, n).cpu().numpy()
    ref, _ = C.forward(params, X, bf16=True)
    err = rel(out, ref)
    print(f"forward rel L2 {err:.2e}")
    # the oracle rounds to bf16 where the kernels do; what remains is FP32 summation order, which
    # flips an occasional bf16 rounding (one 2^-8 step) in a layer's output
    assert err <= 2e-3


def test_backward_matches_oracle(net_and_batch):
    net, params, X, x16, n, rng = net_and_batch
    net.forward(x16, n)
    D, d16 = bf16_dev(rng.standard_normal((n, cnn.N_OUT)) * 0.1)
    net.backward(d16)
    _, cache = C.forward(params, X, bf16=True)
    g = C.backward(params, cache, D, bf16=True)
    for key in net.offsets:
        got = net.view(key, "g").cpu().numpy()
        err = rel(got, g[key])
        print(f"{key}: rel L2 {err:.2e}")
        assert err <= 2e-3, key


@pytest.mark.parametrize("n", [40, 64, 200, 512])
def test_pose_loss_matches_oracle(rng, n):
    """Whole and partial workgroups of 64 images (the loss stages rows through LDS)."""
    out = rng.standard_normal((n, cnn.N_OUT)).astype(np.float32)
    rot = C.orthonormalize(rng.standard_normal((n, 9))).reshape(n, 9).astype(np.float32)
    off = rng.standard_normal((n, 2)).astype(np.float32)
    net = cnn.PoseNet(DEV)
    dout, terms = net.loss(torch.from_numpy(out).to(DEV), torch.from_numpy(rot).to(DEV),
                           torch.from_numpy(off).to(DEV), eps=0.05, off_weight=0.3)
    loss, g, w = C.pose_loss(out.astype(np.float64), rot.astype(np.float64), off.astype(np.float64), 0.05, 0.3)
    t = terms.cpu().numpy()
    assert np.array_equal(t[:, 2].astype(int), w)
    assert abs(t[:, 0].mean() - loss) <= 1e-5 * abs(loss)
    np.testing.assert_allclose(host(dout), C.bf16_round(g.astype(np.float32)), rtol=2**-7, atol=1e-7)


def test_adam_matches_oracle(rng):
    net = cnn.PoseNet(DEV, seed=3)
    p0 = net.p.cpu().numpy().astype(np.float64)
    m = np.zeros_like(p0)
    v = np.zeros_like(p0)
    p = p0
    for t in range(1, 4):
        g = rng.standard_normal(p0.shape).astype(np.float32)
        net.g.copy_(torch.from_numpy(g))
        net.adam_step(lr=1e-3)
        p, m, v = C.adam(p, m, v, g.astype(np.float64), t, lr=1e-3)
    np.testing.assert_allclose(net.p.cpu().numpy(), p, rtol=1e-5, atol=1e-6)
    np.testing.assert_array_equal(host(net.p16), C.bf16_round(net.p.cpu().numpy()))


def test_crop_phase_flip_irfft_matches_numpy(rng):
    """cryomi_crop_spectrum + cryomi_irfft2 = numpy Fourier crop (and CTF sign flip) of the image."""
    from oracle import imaging as O

    lib = _lib.load()
    n, box, ob = 3, 128, 64
    img = rng.standard_normal((n, box, box))
    y = np.fft.rfftn(np.fft.ifftshift(img, axes=(-2, -1)), axes=(-2, -1))
    d = draw(n, 4)
    prm = packed_params(d, 1.5)
    st = torch.cuda.current_stream().cuda_stream
    ydev = torch.from_numpy(y).to(DEV)
    pdev = ops.params_to_device(prm, DEV)
    out = torch.empty((n, ob, ob // 2 + 1), dtype=torch.complex64, device=DEV)
    for flags in (0, _lib.CROP_PHASE_FLIP):
        _lib.check(lib.cryomi_crop_spectrum(ydev.data_ptr(), 1, n, box, ob, pdev.data_ptr(), None, flags,
                                            out.data_ptr(), st), "crop")
        got = ops.irfft2(out, ob).cpu().numpy()
        rows = np.r_[0:ob // 2, box - ob // 2:box]
        crop = y[:, rows, :ob // 2 + 1]
        if flags:
            freq = O.safe_multiply_by_constant(O.make_frequency_grid((box, box), outputs_rfftfreqs=True), 1 / 1.5)
            for i in range(n):
                ctf = O.astigmatic_ctf(freq, 300.0, d["defocus"][i], d["astigmatism"][i], d["astigmatism_angle"][i],
                                       2.7, 0.1)
                crop[i] *= np.where(ctf[rows, :ob // 2 + 1] < 0, -1.0, 1.0)
        want = np.fft.fftshift(np.fft.irfftn(crop, s=(ob, ob), axes=(-2, -1)), axes=(-2, -1))
        err = rel(got, want)
        print(f"flags {flags}: rel L2 {err:.2e}")
        assert err <= 1e-5


@pytest.mark.parametrize("box,ob", [(128, 64), (256, 64), (96, 32)])
def test_crop_hartley_is_fourier_crop_downsample(rng, box, ob):
    """CRYOMI_CROP_HARTLEY gives rfftn of the reference's fourier_crop_downsample (the oracle's
    FP64 restatement, pinned on its own) from the observed rfft half-plane; with the phase flip,
    each bin times the sign of the CTF at its frequency; and irfft2 of it the downsampled image."""
    from oracle import downsample as D
    from oracle import imaging as O

    lib = _lib.load()
    n = 3
    img = rng.standard_normal((n, box, box))
    y = np.fft.rfftn(np.fft.ifftshift(img, axes=(-2, -1)), axes=(-2, -1))
    d = draw(n, 5)
    prm = packed_params(d, 1.0)
    st = torch.cuda.current_stream().cuda_stream
    ydev = torch.from_numpy(y).to(DEV)
    pdev = ops.params_to_device(prm, DEV)
    out = torch.empty((n, ob, ob // 2 + 1), dtype=torch.complex64, device=DEV)
    want = np.stack([D.fourier_crop_downsample(img[i], box / ob, outputs_real_space=False) for i in range(n)])
    want_img = np.stack([D.fourier_crop_downsample(img[i], box / ob) for i in range(n)])
    _lib.check(lib.cryomi_crop_spectrum(ydev.data_ptr(), 1, n, box, ob, pdev.data_ptr(), None, _lib.CROP_HARTLEY,
                                        out.data_ptr(), st), "crop")
    got = out.cpu().numpy()
    err = rel(got, want)
    print(f"box {box} -> {ob}: spectrum rel L2 {err:.2e}")
    assert err <= 1e-6
    img_err = rel(ops.irfft2(out, ob).cpu().numpy(), want_img)
    assert img_err <= 1e-5
    # the unpaired edge really differs from a plain crop (the test would not see a dropped fold)
    rows = np.r_[0:ob // 2, box - ob // 2:box]
    plain = y[:, rows, :ob // 2 + 1]
    assert rel(plain, want) > 1e-3
    _lib.check(lib.cryomi_crop_spectrum(ydev.data_ptr(), 1, n, box, ob, pdev.data_ptr(), None,
                                        _lib.CROP_HARTLEY | _lib.CROP_PHASE_FLIP, out.data_ptr(), st), "crop")
    freq = O.safe_multiply_by_constant(O.make_frequency_grid((box, box), outputs_rfftfreqs=True), 1.0)
    flipped = want.copy()
    for i in range(n):
        ctf = O.astigmatic_ctf(freq, 300.0, d["defocus"][i], d["astigmatism"][i], d["astigmatism_angle"][i], 2.7, 0.1)
        flipped[i] *= np.where(ctf[rows, :ob // 2 + 1] < 0, -1.0, 1.0)
    assert rel(out.cpu().numpy(), flipped) <= 1e-6


def test_band_simulation_and_band_crop(rng):
    """cryomi_simulate_band equals the matching rows / columns of the full Fourier output bit for
    bit, and the band-layout crop equals the full-layout crop (bit for bit without normalisation;
    with it, on a band-limited spectrum, to rounding)."""
    from tests.helpers import fourier_grid, make_volume

    lib = _lib.load()
    box, band, n = 128, 64, 5
    grid = torch.from_numpy(fourier_grid(make_volume(box, 300, 40, 3)).astype(np.complex64)).to(DEV)
    d = draw(n, 9)
    pdev = ops.params_to_device(packed_params(d, 1.0), DEV)
    for extra in (0, _lib.GATHER_FILL):
        flags = _lib.APPLY_CTF | _lib.APPLY_TRANSLATION | _lib.OUTPUT_FOURIER | extra
        full = ops.simulate_images(grid, pdev, flags)
        bnd = ops.simulate_band(grid, pdev, flags, band)
        rows = np.r_[box - band // 2:box, 0:band // 2 + 1]
        assert torch.equal(bnd, full[:, torch.from_numpy(rows).to(DEV), :band // 2 + 1])
    st = torch.cuda.current_stream().cuda_stream
    a = torch.empty((n, band, band // 2 + 1), dtype=torch.complex64, device=DEV)
    b = torch.empty_like(a)
    for fl in (_lib.CROP_HARTLEY | _lib.CROP_PHASE_FLIP, _lib.CROP_PHASE_FLIP, 0):
        _lib.check(lib.cryomi_crop_spectrum(full.data_ptr(), 0, n, box, band, pdev.data_ptr(), None, fl, a.data_ptr(),
                                            st), "crop")
        _lib.check(lib.cryomi_crop_spectrum(bnd.data_ptr(), 0, n, box, band, pdev.data_ptr(), None,
                                            fl | _lib.CROP_BAND_INPUT, b.data_ptr(), st), "crop")
        assert torch.equal(a, b)
    # normalisation: the band layout normalises the band-limited image
    lim = full.clone()
    ky = np.fft.fftfreq(box, 1.0 / box)
    outside = torch.from_numpy((np.abs(ky) > band // 2)).to(DEV)
    lim[:, outside, :] = 0
    lim[:, :, band // 2 + 1:] = 0
    snr = torch.full((n,), 0.3, dtype=torch.float32, device=DEV)
    fl = _lib.CROP_HARTLEY | _lib.CROP_NORMALIZE
    _lib.check(lib.cryomi_crop_spectrum(lim.data_ptr(), 0, n, box, band, None, snr.data_ptr(), fl, a.data_ptr(), st),
               "crop")
    _lib.check(lib.cryomi_crop_spectrum(bnd.data_ptr(), 0, n, box, band, None, snr.data_ptr(),
                                        fl | _lib.CROP_BAND_INPUT, b.data_ptr(), st), "crop")
    assert rel(b.cpu().numpy(), a.cpu().numpy()) <= 1e-5


def test_crop_normalize_scales_to_snr(rng):
    lib = _lib.load()
    n, box = 4, 64
    img = rng.standard_normal((n, box, box)) * 3.0 + 2.0
    y = np.fft.rfftn(np.fft.ifftshift(img, axes=(-2, -1)), axes=(-2, -1)).astype(np.complex64)
    snr = np.array([0.1, 0.5, 1.0, 2.0], dtype=np.float32)
    out = torch.empty((n, box, box // 2 + 1), dtype=torch.complex64, device=DEV)
    st = torch.cuda.current_stream().cuda_stream
    _lib.check(lib.cryomi_crop_spectrum(torch.from_numpy(y).to(DEV).data_ptr(), 0, n, box, box, None,
                                        torch.from_numpy(snr).to(DEV).data_ptr(), _lib.CROP_NORMALIZE,
                                        out.data_ptr(), st), "crop")
    got = ops.irfft2(out, box).cpu().numpy()
    want = (img - img.mean(axis=(1, 2), keepdims=True)) / img.std(axis=(1, 2), keepdims=True)
    want = want * np.sqrt(snr)[:, None, None]
    assert rel(got, want) <= 1e-5


@pytest.mark.parametrize("s", [64, 30, 72])
def test_noise_normalize(rng, s):
    """Images of at most 4096 pixels keep their noisy values in registers between the two passes;
    larger ones (72 x 72) regenerate the noise."""
    lib = _lib.load()
    n = 6
    x = (rng.standard_normal((n, s, s)) * 2 + 1).astype(np.float32)
    xd = torch.from_numpy(x).to(DEV)
    out = torch.empty((n, s, s), dtype=torch.int16, device=DEV)
    st = torch.cuda.current_stream().cuda_stream
    _lib.check(lib.cryomi_noise_normalize(xd.data_ptr(), n, s, 0.0, 1, out.data_ptr(), st), "noise")
    want = (x - x.mean(axis=(1, 2), keepdims=True)) / x.std(axis=(1, 2), keepdims=True)
    assert np.max(np.abs(host(out) - want)) <= 2**-7 * np.abs(want).max()
    zeros = torch.zeros((64, s, s), dtype=torch.float32, device=DEV)
    a = torch.empty((64, s, s), dtype=torch.int16, device=DEV)
    b = torch.empty_like(a)
    _lib.check(lib.cryomi_noise_normalize(zeros.data_ptr(), 64, s, 1.0, 7, a.data_ptr(), st), "noise")
    _lib.check(lib.cryomi_noise_normalize(zeros.data_ptr(), 64, s, 1.0, 8, b.data_ptr(), st), "noise")
    za, zb = host(a).ravel(), host(b).ravel()
    assert abs(np.mean(za ** 4) - 3.0) < 0.05  # Gaussian kurtosis
    assert abs(np.corrcoef(za, zb)[0, 1]) < 0.01  # seeds give independent streams
    assert abs(np.corrcoef(za[:-1], za[1:])[0, 1]) < 0.01  # white
    c = torch.empty_like(a)
    _lib.check(lib.cryomi_noise_normalize(zeros.data_ptr(), 64, s, 1.0, 7, c.data_ptr(), st), "noise")
    assert torch.equal(a, c)


@pytest.mark.parametrize("n,h,c,co", [(3, 16, 32, 64), (2, 8, 128, 128), (5, 32, 8, 32), (1, 4, 64, 8)])
def test_implicit_convolution_ops(rng, n, h, c, co):
    """cryomi_conv3x3_s2_bf16: the forward equals explicit patches + the dense GEMM bit for bit
    (same kernel, same k order), and all three ops match the FP64 products of the oracle's
    patches (weight grad: split-k partial slices summed in order, with and without the bias
    column and in both write modes; data grad: the adjoint with the ReLU mask, one bf16 rounding)."""
    lib = _lib.load()
    st = torch.cuda.current_stream().cuda_stream
    kp = 9 * c
    npix = n * (h // 2) ** 2
    X, x16 = bf16_dev(rng.standard_normal((n, h, h, c)))
    W, w16 = bf16_dev(rng.standard_normal((co, kp)) / np.sqrt(kp))
    bias = (rng.standard_normal(co) * 0.1).astype(np.float32)
    bdev = torch.from_numpy(bias).to(DEV)
    col = torch.empty((npix, kp), dtype=torch.int16, device=DEV)
    _lib.check(lib.cryomi_im2col_s2(x16.data_ptr(), n, h, h, c, kp, col.data_ptr(), st), "im2col")
    y_exp = torch.empty((npix, co), dtype=torch.int16, device=DEV)
    cnn.gemm(col, kp, 1, w16, 1, kp, npix, co, kp, c16=y_exp, ldc16=co, bias=bdev, relu=True)
    y_imp = torch.empty_like(y_exp)
    y32 = torch.empty((npix, co), dtype=torch.float32, device=DEV)
    cnn.conv(_lib.CONV_FORWARD, n, h, h, c, co, x=x16, weights=w16, bias=bdev, c16=y_imp, relu=True)
    cnn.conv(_lib.CONV_FORWARD, n, h, h, c, co, x=x16, weights=w16, bias=bdev, c32=y32, relu=True)
    assert torch.equal(y_imp, y_exp)
    P = C.im2col(X, kp)
    ref = np.maximum(P @ W.T + bias, 0.0)
    assert rel(y32.cpu().numpy(), ref) <= 1e-5
    D, d16 = bf16_dev(rng.standard_normal((npix, co)))
    ws = torch.empty((cnn.conv_workspace_floats(c, co, 3),), dtype=torch.float32, device=DEV)
    gw = torch.full((co, kp), 0.25, dtype=torch.float32, device=DEV)
    gb = torch.full((co,), -0.5, dtype=torch.float32, device=DEV)
    cnn.conv(_lib.CONV_WEIGHT_GRAD, n, h, h, c, co, x=x16, dy=d16, c32=gw, bias_grad=gb, split_k=3, workspace=ws,
             accumulate=True)
    assert rel(gw.cpu().numpy(), 0.25 + D.T @ P) <= 1e-5
    assert rel(gb.cpu().numpy(), -0.5 + D.sum(axis=0)) <= 1e-5
    # overwrite mode, one layer, no bias column: the same products
    gw1 = torch.full((co, kp), 7.0, dtype=torch.float32, device=DEV)
    cnn.conv(_lib.CONV_WEIGHT_GRAD, n, h, h, c, co, x=x16, dy=d16, c32=gw1, split_k=1, workspace=ws)
    assert rel(gw1.cpu().numpy(), D.T @ P) <= 1e-5
    gw2 = torch.empty_like(gw1)  # deterministic: the same call twice gives the same bits
    cnn.conv(_lib.CONV_WEIGHT_GRAD, n, h, h, c, co, x=x16, dy=d16, c32=gw2, split_k=1, workspace=ws)
    assert torch.equal(gw1, gw2)
    with pytest.raises(ValueError, match="workspace"):
        cnn.conv(_lib.CONV_WEIGHT_GRAD, n, h, h, c, co, x=x16, dy=d16, c32=gw, bias_grad=gb, split_k=3,
                 workspace=ws[:co * kp])
    dx = torch.empty((n, h, h, c), dtype=torch.int16, device=DEV)
    cnn.conv(_lib.CONV_DATA_GRAD, n, h, h, c, co, weights=w16, dy=d16, mask=x16, c16=dx)
    want = C.bf16_round((C.col2im(D @ W, (n, h, h, c)) * (X > 0)).astype(np.float32))
    got = host(dx)
    assert np.max(np.abs(got - want)) <= 2**-7 * np.max(np.abs(want))
    assert rel(got, want) <= 4e-3
