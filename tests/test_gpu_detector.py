"""GPU parity of the detector chain (dlx_detector_apply / dlx_detector_vjp),
the detector layers, LayeredDetector, and the source-to-detector models
(BinarySource, Scene, Telescope, Dither) against the FP64 oracle.

Error metric: peak-normalised max |x - ref| / max |ref|.  Bounds: 2e-6 for the
detector chain and its VJP (float32 sums of at most a few hundred terms),
1e-5 for anything that contains a PSF (the PSF bound of the hot path).
"""

import numpy as np
import pytest

from oracle import dlux_oracle as orc

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

CHAIN_TOL = 2e-6
PSF_TOL = 1e-5


def _err(x, ref):
    x = x.detach().cpu().numpy() if isinstance(x, torch.Tensor) else np.asarray(x)
    return float(np.max(np.abs(x - ref)) / np.max(np.abs(ref)))


@pytest.fixture(scope="module")
def dl():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import dlux_amd
    return dlux_amd


def _threshold_in_gap(values, q):
    """A threshold near quantile q at least 1e-4 (relative) from every value,
    so float32 and float64 put each pixel on the same side of it."""
    v = np.sort(values.ravel())
    scale = np.max(np.abs(v))
    i = int(q * (len(v) - 1))
    for j in range(i, len(v) - 1):
        if v[j + 1] - v[j] > 2e-4 * scale:
            return 0.5 * (v[j] + v[j + 1])
    raise AssertionError("no gap found")


def _chains(rng):
    """(name, n, batch, oracle chain) cases; thresholds put in gaps of the data."""
    cases = []
    cases.append(("jitter", 64, 3, [("convolve", orc.jitter_kernel(0.8, 9, 3))]))
    r64, r16 = rng.uniform(0.5, 1.5, (64, 64)), rng.uniform(0.5, 1.5, (16, 16))
    cases.append(("full", 64, 2, [("add", 0.5), ("pixel_response", r64),
                                  ("convolve", rng.normal(size=(4, 4))), ("saturate", None),
                                  ("downsample", 4), ("pixel_response", r16), ("add", -0.1)]))
    cases.append(("odd", 45, 1, [("convolve", rng.uniform(0, 1, (7, 7))), ("downsample", 3),
                                 ("saturate", None)]))
    cases.append(("wide-kernel", 20, 2, [("convolve", rng.normal(size=(33, 33)))]))
    cases.append(("max-kernel", 70, 1, [("convolve", rng.normal(size=(64, 64)))]))
    many = []
    for k in range(11):
        many.append(("add", 0.01 * (k + 1)))
        many.append(("pixel_response", rng.uniform(0.9, 1.1, (32, 32))))
    cases.append(("ew-overflow", 32, 2, many))
    cases.append(("down-down", 96, 2, [("downsample", 2), ("add", 1.0), ("downsample", 3),
                                       ("pixel_response", rng.uniform(0.5, 1.5, (16, 16)))]))
    cases.append(("identity", 17, 2, []))
    cases.append(("large", 1024, 2, [("convolve", orc.jitter_kernel(0.6, 9, 3)),
                                     ("downsample", 4), ("saturate", None)]))
    # convolve + block sum fused with 15 x 60 tiles; then a factor too big to fuse
    cases.append(("conv-down5", 80, 2, [("convolve", rng.normal(size=(6, 6))), ("add", 0.2),
                                        ("downsample", 5), ("pixel_response",
                                                            rng.uniform(0.5, 1.5, (16, 16)))]))
    cases.append(("conv-down32", 96, 1, [("convolve", rng.normal(size=(3, 3))),
                                         ("downsample", 32)]))
    cases.append(("conv-down16-odd-tiles", 144, 1, [("convolve", rng.uniform(0, 1, (9, 9))),
                                                    ("downsample", 16), ("saturate", None)]))
    # separable kernels: unequal factors, alone, fused with a block sum, at the size limit
    cases.append(("sep", 70, 2, [("pixel_response", rng.uniform(0.5, 1.5, (70, 70))),
                                 ("convolve_separable", (rng.normal(size=7), rng.normal(size=7))),
                                 ("add", 0.3)]))
    cases.append(("sep-down3", 99, 2, [("convolve_separable", (rng.uniform(0, 1, 10),
                                                               rng.uniform(0, 1, 10))),
                                       ("saturate", None), ("downsample", 3)]))
    cases.append(("sep-max", 40, 1, [("convolve_separable", (rng.normal(size=64),
                                                             rng.normal(size=64)))]))
    # the register-window limit (13 taps) and one past it, back to back
    cases.append(("k13-k14", 50, 2, [("convolve_separable", (rng.normal(size=13),
                                                             rng.normal(size=13))),
                                     ("convolve", rng.normal(size=(14, 14))),
                                     ("convolve", rng.normal(size=(13, 13)))]))
    # the streaming separable kernel (<= 13 taps; block sums of 1, 2, 4): the
    # jitter -> response -> downsample 4 -> saturation -> constant chain over
    # several segments per band; ragged bands and steps; an even kernel under a
    # load-side map; the tap limit with 2 x 2 blocks (their own generator, so
    # the draws of the cases above are unchanged)
    rng = np.random.default_rng(7)
    f9 = orc.jitter_kernel(0.5, 9, 3).sum(0)   # the kernel is outer(f9, f9)
    cases.append(("stream-jitter-down4", 1024, 2,
                  [("convolve_separable", (f9, f9)),
                   ("pixel_response", rng.uniform(0.9, 1.1, (1024, 1024))), ("downsample", 4),
                   ("saturate", None), ("add", 0.01)]))
    cases.append(("stream-down2-ragged", 98, 3,
                  [("convolve_separable", (rng.normal(size=7), rng.normal(size=7))),
                   ("downsample", 2), ("add", 0.2)]))
    cases.append(("stream-even-pre-map", 100, 2,
                  [("pixel_response", rng.uniform(0.5, 1.5, (100, 100))), ("add", 0.1),
                   ("convolve_separable", (rng.uniform(0, 1, 6), rng.uniform(0, 1, 6))),
                   ("downsample", 4), ("pixel_response", rng.uniform(0.5, 1.5, (25, 25)))]))
    cases.append(("stream-k13-down2", 130, 1,
                  [("convolve_separable", (rng.normal(size=13), rng.normal(size=13))),
                   ("saturate", None), ("downsample", 2)]))
    # the column-walking kernel (batches of >= 512 band-segment-image waves):
    # the jitter chain over 32 frames of 512^2; a load-side map, which sends
    # 128^2 frames to the tile kernels; 1028^2 (a fifth 256-column band of one
    # column group)
    cases.append(("col-512-jitter-down4", 512, 32,
                  [("convolve_separable", (f9, f9)),
                   ("pixel_response", rng.uniform(0.9, 1.1, (512, 512))), ("downsample", 4),
                   ("saturate", None), ("add", 0.01)]))
    cases.append(("col-128-pre-map", 128, 128,
                  [("pixel_response", rng.uniform(0.5, 1.5, (128, 128))),
                   ("convolve_separable", (rng.uniform(0, 1, 7), rng.uniform(0, 1, 7))),
                   ("downsample", 2), ("add", 0.05)]))
    cases.append(("col-1028-down4", 1028, 8,
                  [("convolve_separable", (rng.normal(size=5), rng.normal(size=5))),
                   ("downsample", 4), ("saturate", None)]))
    return cases


def _resolve(chain, x):
    """Fill each saturation threshold with a value in a gap of the data
    reaching it (per image batch, taken over the whole batch)."""
    out = []
    for kind, arg in chain:
        if kind == "saturate" and arg is None:
            pre = np.stack([orc.detector(xi, out) for xi in x])
            arg = float(np.float32(_threshold_in_gap(pre, 0.8)))
        out.append((kind, arg))
    return out


def _host_op(op):
    """A chain op with its device arrays copied to the host (for the oracle)."""
    kind, arg = op

    def host(a):
        return a.double().cpu().numpy() if isinstance(a, torch.Tensor) else a
    return (kind, tuple(host(a) for a in arg) if isinstance(arg, tuple) else host(arg))


def _oracle(x, chain):
    return np.stack([orc.detector(xi, chain) for xi in x])


@pytest.mark.parametrize("case", range(23))
def test_detector_chain_and_vjp_vs_oracle(dl, case):
    rng = np.random.default_rng(40 + case)
    name, n, batch, chain = _chains(rng)[case]
    print(f"[detector] {name} n={n} batch={batch} ops={len(chain)}", flush=True)
    x = rng.uniform(0.0, 1.0, (batch, n, n))
    chain = _resolve(chain, x)
    ref = _oracle(x, chain)
    got = dl.runtime.detector(torch.as_tensor(x, dtype=torch.float32).cuda(), chain)
    assert tuple(got.shape) == ref.shape
    err = _err(got, ref)
    print(f"  forward err {err:.2e}")
    assert err < CHAIN_TOL
    g = rng.normal(size=ref.shape)
    vjp = dl.runtime.detector_vjp(torch.as_tensor(x, dtype=torch.float32).cuda(), chain,
                                  torch.as_tensor(g, dtype=torch.float32).cuda())
    vref = np.stack([orc.detector_vjp(xi, chain, gi) for xi, gi in zip(x, g)])
    err = _err(vjp, vref)
    print(f"  vjp err {err:.2e}")
    assert err < CHAIN_TOL


# Chains the column-walking kernel takes (separable, odd K <= 13, n % 4 == 0,
# no op before the convolution, at most one map after it and none after the
# block sum, at least 512 waves of 256 columns x 32 rows): block sums of 4, 2
# and 1 (the convolution alone), with and without the streamed map, the 1- and
# 13-tap limits, bands that end inside the image (128, 300, 1028) and halos
# past both edges, ragged segments (300).  One image at a time the same chains
# run on the tile kernels, whose arithmetic the column kernel repeats in the
# same order: the results must be bit-identical, and match the oracle.
COLUMN_CASES = [
    ("jitter-map-down4", 512, 32, 9, [("pixel_response", "map"), ("downsample", 4),
                                      ("saturate", None), ("add", 0.01)]),
    ("k5-down4-ragged-band", 1028, 8, 5, [("downsample", 4), ("saturate", None)]),
    ("k7-down2-half-band", 128, 128, 7, [("downsample", 2), ("add", 0.05)]),
    ("k13-map-alone", 256, 64, 13, [("pixel_response", "map"), ("add", -0.2)]),
    ("k1-down2", 64, 256, 1, [("downsample", 2)]),
    ("k3-map-down4-ragged-seg", 300, 40, 3, [("add", 0.3), ("pixel_response", "map"),
                                             ("downsample", 4), ("add", 1.0)]),
    ("k11-alone-saturate", 200, 100, 11, [("saturate", None), ("add", 0.5)]),
    ("k5-map-down4-partial-band", 320, 48, 5, [("pixel_response", "map"), ("downsample", 4),
                                               ("saturate", None)]),
]


@pytest.mark.parametrize("case", range(len(COLUMN_CASES)))
def test_column_kernel_bit_identical_to_tile_kernels(dl, case):
    name, n, batch, K, tail = COLUMN_CASES[case]
    print(f"[detector column] {name} n={n} batch={batch} K={K}", flush=True)
    rng = np.random.default_rng(90 + case)
    x = rng.uniform(0.0, 1.0, (batch, n, n)).astype(np.float32)
    chain = [("convolve_separable", (rng.uniform(0, 1, K), rng.uniform(0, 1, K)))]
    side = n
    for kind, arg in tail:
        if kind == "pixel_response":
            arg = rng.uniform(0.8, 1.2, (side, side))
        elif kind == "downsample":
            side //= arg
        chain.append((kind, arg))
    chain = _resolve(chain, x[:2].astype(np.float64))
    dev = [(k, tuple(torch.as_tensor(np.asarray(b, np.float32)).cuda() for b in a)
            if isinstance(a, tuple) else torch.as_tensor(np.asarray(a, np.float32)).cuda()
            if isinstance(a, np.ndarray) else a) for k, a in chain]
    xt = torch.as_tensor(x).cuda()
    got = dl.runtime.detector(xt, dev)
    one = torch.stack([dl.runtime.detector(xt[b], dev) for b in range(batch)])
    torch.cuda.synchronize()
    assert torch.equal(got, one), f"{(got != one).sum().item()} pixels differ"
    ref = _oracle(x[:2].astype(np.float64), chain)
    err = _err(got[:2], ref)
    print(f"  err vs oracle {err:.2e}")
    assert err < CHAIN_TOL


# The image cotangent of those chains: where the convolution's ops need no
# values for their derivative (adds, one map) and the side divides by 4 KD, one
# forward recompute of the block sums and the adjoint column walk
# (det_sep_col_vjp_kernel); one image at a time, the multi-pass adjoint (saved
# pass inputs, elementwise derivatives, block replication, correlation).  The
# products are the same in the same order: bit-identical, and the oracle's.
@pytest.mark.parametrize("case", range(len(COLUMN_CASES)))
def test_column_vjp_bit_identical_to_multipass(dl, case):
    name, n, batch, K, tail = COLUMN_CASES[case]
    print(f"[detector column vjp] {name} n={n} batch={batch} K={K}", flush=True)
    rng = np.random.default_rng(120 + case)
    x = rng.uniform(0.0, 1.0, (batch, n, n)).astype(np.float32)
    chain = [("convolve_separable", (rng.uniform(0, 1, K), rng.uniform(0, 1, K)))]
    side = n
    for kind, arg in tail:
        if kind == "pixel_response":
            arg = rng.uniform(0.8, 1.2, (side, side))
        elif kind == "downsample":
            side //= arg
        chain.append((kind, arg))
    chain = _resolve(chain, x[:2].astype(np.float64))
    dev = [(k, tuple(torch.as_tensor(np.asarray(b, np.float32)).cuda() for b in a)
            if isinstance(a, tuple) else torch.as_tensor(np.asarray(a, np.float32)).cuda()
            if isinstance(a, np.ndarray) else a) for k, a in chain]
    g = rng.normal(size=(batch, side, side)).astype(np.float32)
    xt, gt = torch.as_tensor(x).cuda(), torch.as_tensor(g).cuda()
    got = dl.runtime.detector_vjp(xt, dev, gt)
    one = torch.stack([dl.runtime.detector_vjp(xt[b], dev, gt[b]) for b in range(batch)])
    torch.cuda.synchronize()
    assert torch.equal(got, one), f"{(got != one).sum().item()} pixels differ"
    ref = np.stack([orc.detector_vjp(x[b].astype(np.float64), chain, g[b].astype(np.float64))
                    for b in range(2)])
    err = _err(got[:2], ref)
    print(f"  err vs oracle {err:.2e}")
    assert err < CHAIN_TOL


def test_detector_autograd_and_fusion_invariance(dl):
    """torch autograd through LayeredDetector uses dlx_detector_vjp; the fused
    chain equals the layers applied one at a time."""
    print("[detector] autograd / fusion", flush=True)
    from dlux_amd import layers
    rng = np.random.default_rng(3)
    n = 96
    x = torch.as_tensor(rng.uniform(0, 1, (n, n)), dtype=torch.float32).cuda()
    resp = rng.uniform(0.8, 1.2, (n, n)).astype(np.float32)
    det = dl.LayeredDetector([layers.ApplyPixelResponse(resp), layers.ApplyJitter(0.7),
                              layers.Downsample(3), layers.AddConstant(2.0)])
    psf = dl.PSF(x, 0.1)
    fused = det.model(psf, return_psf=True)
    step = psf
    for layer in det.layers.values():
        step = layer(step)
    assert torch.allclose(fused.data, step.data, rtol=1e-6, atol=1e-6)
    assert np.isclose(float(fused.pixel_scale), 0.3) and np.isclose(float(step.pixel_scale), 0.3)
    xg = x.clone().requires_grad_(True)
    w = torch.as_tensor(rng.normal(size=(32, 32)), dtype=torch.float32).cuda()
    (grad,) = torch.autograd.grad((det(dl.PSF(xg, 0.1)) * w).sum(), xg)
    chain = [_host_op(layer._det_op()) for layer in det.layers.values()]
    ref = orc.detector_vjp(x.double().cpu().numpy(), chain, w.double().cpu().numpy())
    assert _err(grad, ref) < CHAIN_TOL


def test_detector_rejects_mismatched_response(dl):
    print("[detector] errors", flush=True)
    from dlux_amd import layers
    psf = dl.PSF(torch.zeros((64, 64), device="cuda"), 1.0)
    det = dl.LayeredDetector([layers.Downsample(2), layers.ApplyPixelResponse(np.ones((64, 64)))])
    with pytest.raises(RuntimeError, match="64x64 but the image is 32x32"):
        det(psf)
    with pytest.raises(RuntimeError, match="not divisible by 5"):
        layers.Downsample(5)(psf)


def _optics_and_oracle(dl):
    from dlux_amd import workloads
    cfg = workloads.MONO
    optics = workloads.build_system(cfg)
    mask = orc.circular_aperture(cfg.n_in, 1.0, oversample=1)
    basis = orc.zernike_basis(cfg.js, orc.pixel_coords(cfg.n_in, 1.0), 1.0)

    def star(wls, weights, pos):
        return orc.poly_psf(mask, basis, cfg.coefficients(), wls, weights, 1.0, cfg.n_out,
                            cfg.p_out, offset=pos)
    return cfg, optics, star


@pytest.mark.parametrize("precision", ["fp32", "bf16x3"])
def test_telescope_binary_scene_with_detector_vs_oracle(dl, precision):
    """Scene(BinarySource, PointSources on another wavelength grid) through the
    optics and a jitter + pixel-response + downsample detector; under bf16x3 the
    whole scene is one batched launch."""
    print(f"[telescope] scene + detector, {precision}", flush=True)
    from dlux_amd import layers, runtime
    cfg, optics, star = _optics_and_oracle(dl)
    p = cfg.p_out
    wl_a = np.array([0.95e-6, 1.05e-6])
    wl_b = np.array([0.9e-6, 1.0e-6, 1.1e-6])
    binary = dl.BinarySource(wl_a, position=[1.5 * p, -0.5 * p], mean_flux=2.0,
                             separation=6.3 * p, position_angle=0.7, contrast=3.0,
                             weights=np.array([[1.0, 2.0], [2.0, 1.0]]))
    pts = dl.PointSources(wl_b, position=np.array([[-4.2 * p, 3.1 * p], [2.2 * p, 2.9 * p]]),
                          flux=[0.5, 1.5])
    resp = np.random.default_rng(8).uniform(0.9, 1.1, (cfg.n_out, cfg.n_out))
    det = dl.LayeredDetector([layers.ApplyJitter(0.9, 5, 3), layers.ApplyPixelResponse(resp),
                              layers.Downsample(2)])
    tel = dl.Telescope(optics, [("binary", binary), ("pts", pts)], det)
    with runtime.precision(precision):
        out = tel.model(return_psf=True)
    # oracle: every star, then the detector chain
    f = 2.0 * 2 * np.array([3.0, 1.0]) / 4.0
    sep = np.array([np.sin(0.7), np.cos(0.7)]) * 6.3 * p / 2
    centre = np.array([1.5 * p, -0.5 * p])
    wa = np.array([[1.0, 2.0], [2.0, 1.0]]) / 3.0
    img = star(wl_a, wa[0] * f[0], centre + sep) + star(wl_a, wa[1] * f[1], centre - sep)
    for pos, flux in zip(pts.position, [0.5, 1.5]):
        img = img + star(wl_b, np.ones(3) / 3 * flux, pos)
    ref = orc.detector(img, [("convolve", orc.jitter_kernel(0.9, 5, 3)),
                             ("pixel_response", resp), ("downsample", 2)])
    err = _err(out.data, ref)
    print(f"  err {err:.2e}")
    assert err < PSF_TOL
    assert np.isclose(float(out.pixel_scale), 2 * cfg.p_out)


def test_dither_single_launch_matches_per_dither_models(dl):
    """Dither: every dither in one batched launch equals the Telescope modelled
    with the source moved by that dither (instruments.py:231-256)."""
    print("[dither]", flush=True)
    from dlux_amd import layers, runtime
    cfg, optics, star = _optics_and_oracle(dl)
    p = cfg.p_out
    dithers = np.array([[0.0, 0.0], [2.5 * p, -1.0 * p], [-3.0 * p, 4.0 * p]])
    src = dl.BinarySource([1e-6, 1.1e-6], separation=5 * p, contrast=2.0)
    det = dl.LayeredDetector([layers.Downsample(4), layers.AddConstant(0.25)])
    dith = dl.Dither(dithers, optics, src, det)
    with runtime.precision("bf16x3"):
        got = dith.model(return_psf=True)
        assert tuple(got.data.shape) == (3, cfg.n_out // 4, cfg.n_out // 4)
        for k, d in enumerate(dithers):
            one = dl.Telescope(optics, src.add("position", d), det).model()
            assert _err(got.data[k], one.cpu().numpy()) < 1e-6
    # dither 1 against the oracle
    f = 2 * np.array([2.0, 1.0]) / 3.0
    half = np.array([np.sin(np.pi / 2), np.cos(np.pi / 2)]) * 5 * p / 2
    c = dithers[1]
    img = star([1e-6, 1.1e-6], np.full(2, 0.5) * f[0], c + half) + \
        star([1e-6, 1.1e-6], np.full(2, 0.5) * f[1], c - half)
    ref = orc.detector(img, [("downsample", 4), ("add", 0.25)])
    assert _err(got.data[1], ref) < PSF_TOL


def test_resolved_sources_vs_oracle(dl):
    """ResolvedSource = convolve(PSF, distribution, "same") (sources.py:490-521);
    PointResolvedSource = point PSF + convolved resolved PSF with the flux split
    by contrast (sources.py:711-748); in a Scene beside a point source."""
    print("[resolved sources]", flush=True)
    cfg, optics, star = _optics_and_oracle(dl)
    p = cfg.p_out
    rng = np.random.default_rng(17)
    dist = rng.uniform(0, 1, (9, 9))
    wls = np.array([0.95e-6, 1.05e-6])
    pos = np.array([1.2 * p, -0.7 * p])
    res = dl.ResolvedSource(wls, position=pos, flux=2.0, distribution=dist)
    got = res.model(optics, return_psf=True)
    ref = orc.convolve_same(star(wls, np.full(2, 0.5) * 2.0, pos), dist / dist.sum())
    assert _err(got.data, ref) < PSF_TOL
    assert np.isclose(float(got.pixel_scale), cfg.p_out)

    prs = dl.PointResolvedSource(wls, position=pos, flux=3.0, distribution=dist, contrast=4.0,
                                 weights=np.array([[1.0, 3.0], [2.0, 2.0]]))
    # the reference's wavefront carries the default 1/L weights too (sources.py:728-741)
    f = 2 * 3.0 * np.array([4.0, 1.0]) / 5.0 / len(wls)
    ref2 = (star(wls, np.array([0.25, 0.75]) * f[0], pos)
            + orc.convolve_same(star(wls, np.array([0.5, 0.5]) * f[1], pos), dist / dist.sum()))
    assert _err(prs.model(optics), ref2) < PSF_TOL

    scene = dl.Scene([("disk", prs), ("star", dl.PointSource(wls, position=-pos, flux=1.5))])
    ref3 = ref2 + star(wls, np.full(2, 0.5) * 1.5, -pos)
    assert _err(scene.model(optics), ref3) < PSF_TOL


@pytest.mark.parametrize("where", ["pre", "post"])
def test_detector_parameter_gradients_vs_fd(dl, where):
    """dlx_detector_vjp_params through torch autograd: the cotangents of two
    pixel-response maps (one entering a convolution -- the load-side list --
    or applied after it, and one after a downsample), a saturation level and a
    constant, against central differences of the oracle chain (directional
    for the maps); the image cotangent against the oracle VJP.  Bound 2e-5
    relative (float32 sums over the image and batch)."""
    print(f"[detector] parameter gradients, map {where} the convolution", flush=True)
    rng = np.random.default_rng(71 if where == "pre" else 72)
    n, batch, k = 48, 3, 4
    x = rng.uniform(0.0, 1.0, (batch, n, n))
    kern = rng.uniform(0.0, 1.0, (5, 5))
    kern /= kern.sum()
    m1 = rng.uniform(0.8, 1.2, (n, n))
    m2 = rng.uniform(0.8, 1.2, (n // k, n // k))
    c = 0.37

    def chain_of(m1_, m2_, t_, c_):
        head = ([("pixel_response", m1_), ("convolve", kern)] if where == "pre"
                else [("convolve", kern), ("pixel_response", m1_)])
        return head + [("downsample", k), ("pixel_response", m2_), ("saturate", t_),
                       ("add", c_)]
    pre_sat = _oracle(x, chain_of(m1, m2, 1e30, 0.0)[:-2])
    t = float(np.float32(_threshold_in_gap(pre_sat, 0.7)))
    w = rng.normal(size=(batch, n // k, n // k))

    def loss(m1_, m2_, t_, c_):
        return float(np.sum(w * _oracle(x, chain_of(m1_, m2_, t_, c_))))

    dev = "cuda"
    f32 = lambda a: torch.tensor(np.asarray(a, np.float32), device=dev, requires_grad=True)
    xt, m1t, m2t, tt, ct = f32(x), f32(m1), f32(m2), f32(t), f32(c)
    chain_t = chain_of(m1t, m2t, tt, ct)
    chain_t[1 if where == "pre" else 0] = ("convolve", torch.as_tensor(
        kern.astype(np.float32), device=dev))
    out = dl.runtime.detector(xt, chain_t)
    assert _err(out, _oracle(x, chain_of(m1, m2, t, c))) < CHAIN_TOL
    (out * torch.as_tensor(w.astype(np.float32), device=dev)).sum().backward()
    h = 1e-4
    V1, V2 = rng.normal(size=m1.shape), rng.normal(size=m2.shape)
    fd = {
        "m1": (loss(m1 + h * V1, m2, t, c) - loss(m1 - h * V1, m2, t, c)) / (2 * h),
        "m2": (loss(m1, m2 + h * V2, t, c) - loss(m1, m2 - h * V2, t, c)) / (2 * h),
        "t": (loss(m1, m2, t + h * t, c) - loss(m1, m2, t - h * t, c)) / (2 * h * t),
        "c": (loss(m1, m2, t, c + h) - loss(m1, m2, t, c - h)) / (2 * h),
    }
    got = {
        "m1": float(np.sum(m1t.grad.double().cpu().numpy() * V1)),
        "m2": float(np.sum(m2t.grad.double().cpu().numpy() * V2)),
        "t": float(tt.grad), "c": float(ct.grad),
    }
    for key in fd:
        err = abs(got[key] - fd[key]) / max(abs(fd[key]), 1e-12)
        print(f"  d/d{key}: got {got[key]:.6e} fd {fd[key]:.6e} err {err:.2e}")
        assert err < 2e-5, key
    assert abs(got["c"] - np.sum(w)) < 2e-5 * abs(np.sum(w)) + 1e-4
    vref = np.stack([orc.detector_vjp(xi, chain_of(m1, m2, t, c), wi) for xi, wi in zip(x, w)])
    assert _err(xt.grad, vref) < CHAIN_TOL


def test_detector_kernel_and_jitter_gradients_vs_fd(dl):
    """Kernel cotangents of dlx_detector_vjp_params: a full convolution
    kernel (directional central difference of the oracle chain), separable
    factors, and ApplyJitter's sigma through its Gaussian factor.  Bound 2e-5
    relative."""
    print("[detector] kernel and jitter gradients", flush=True)
    from dlux_amd import layers
    rng = np.random.default_rng(81)
    n, batch = 64, 2
    x = rng.uniform(0.0, 1.0, (batch, n, n))
    kern = rng.uniform(0.0, 1.0, (7, 7))
    kern /= kern.sum()
    w = rng.normal(size=(batch, n // 2, n // 2))
    resp = rng.uniform(0.9, 1.1, (n, n))

    def loss(chain):
        return float(np.sum(w * _oracle(x, chain)))

    def chain_of(k_):
        return [("pixel_response", resp), ("convolve", k_), ("downsample", 2), ("add", 0.5)]
    dev = "cuda"
    xt = torch.as_tensor(x.astype(np.float32), device=dev)
    wt = torch.as_tensor(w.astype(np.float32), device=dev)
    kt = torch.tensor(kern.astype(np.float32), device=dev, requires_grad=True)
    rt = torch.as_tensor(resp.astype(np.float32), device=dev)
    out = dl.runtime.detector(xt, [("pixel_response", rt), ("convolve", kt), ("downsample", 2),
                                   ("add", 0.5)])
    (out * wt).sum().backward()
    h, V = 1e-4, rng.normal(size=kern.shape)
    fd = (loss(chain_of(kern + h * V)) - loss(chain_of(kern - h * V))) / (2 * h)
    got = float(np.sum(kt.grad.double().cpu().numpy() * V))
    print(f"  d/dkernel: got {got:.6e} fd {fd:.6e}")
    assert abs(got - fd) < 2e-5 * abs(fd)
    # separable factors
    wy, wx = rng.uniform(0, 1, 5), rng.uniform(0, 1, 5)
    wyt = torch.tensor(wy.astype(np.float32), device=dev, requires_grad=True)
    wxt = torch.tensor(wx.astype(np.float32), device=dev, requires_grad=True)
    out = dl.runtime.detector(xt, [("convolve_separable", (wyt, wxt)), ("downsample", 2)])
    (out * wt).sum().backward()
    Vy, Vx = rng.normal(size=5), rng.normal(size=5)

    def lsep(a, b):
        return loss([("convolve_separable", (a, b)), ("downsample", 2)])
    fdy = (lsep(wy + h * Vy, wx) - lsep(wy - h * Vy, wx)) / (2 * h)
    fdx = (lsep(wy, wx + h * Vx) - lsep(wy, wx - h * Vx)) / (2 * h)
    gy = float(np.sum(wyt.grad.double().cpu().numpy() * Vy))
    gx = float(np.sum(wxt.grad.double().cpu().numpy() * Vx))
    print(f"  d/dwy: got {gy:.6e} fd {fdy:.6e}; d/dwx: got {gx:.6e} fd {fdx:.6e}")
    assert abs(gy - fdy) < 2e-5 * abs(fdy) and abs(gx - fdx) < 2e-5 * abs(fdx)
    # ApplyJitter's sigma through a LayeredDetector
    sig = torch.tensor(0.8, dtype=torch.float64, requires_grad=True)
    det = dl.LayeredDetector([layers.ApplyJitter(sig), layers.Downsample(2)])
    img = det(dl.PSF(xt[0], 1.0))
    (img * wt[0]).sum().backward()

    def ljit(s_):
        return float(np.sum(w[0] * orc.detector(x[0], [("convolve", orc.jitter_kernel(s_)),
                                                       ("downsample", 2)])))
    fds = (ljit(0.8 + 1e-5) - ljit(0.8 - 1e-5)) / 2e-5
    print(f"  d/dsigma: got {float(sig.grad):.6e} fd {fds:.6e}")
    assert abs(float(sig.grad) - fds) < 1e-4 * abs(fds)
