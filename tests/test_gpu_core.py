"""GPU parity of the C-ABI kernels against the FP64 oracle.

Error metric: peak-normalised max |x - x_ref| / max |x_ref| (fields and PSFs).
Tolerances: FP32 form 2e-6 on fields / 1e-5 on PSFs; split-BF16x3 form 3e-5 on
fields / 1e-5 on PSFs (the PSF bound is the one the hot path is held to).
"""

import numpy as np
import pytest

from oracle import dlux_oracle as orc

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _peak_err(x, ref):
    x = np.asarray(x)
    ref = np.asarray(ref)
    return float(np.max(np.abs(x - ref)) / np.max(np.abs(ref)))


@pytest.fixture(scope="module")
def rt():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from dlux_amd import runtime
    return runtime


FIELD_TOL = {"fp32": 2e-6, "bf16x3": 3e-5}


def _rand_field(rng, n):
    return (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) / n


@pytest.mark.parametrize("prec", ["fp32", "bf16x3"])
@pytest.mark.parametrize("n_in,n_out", [(32, 16), (16, 8), (33, 17), (256, 128), (96, 200)])
def test_mft_random_field(rt, prec, n_in, n_out):
    print(f"[mft] prec={prec} n_in={n_in} n_out={n_out}", flush=True)
    rng = np.random.default_rng(n_in * 1000 + n_out)
    E = _rand_field(rng, n_in)
    wl, dx, p = 1e-6, 1.0 / n_in, 0.25e-6
    out = rt.mft(E, wl, dx, n_out, p, precision=prec).cpu().numpy()
    ref = orc.mft(E, wl, dx, n_out, p)
    assert out.shape == (n_out, n_out)
    assert _peak_err(out, ref) < FIELD_TOL[prec]


@pytest.mark.parametrize("prec", ["fp32", "bf16x3"])
@pytest.mark.parametrize("focal,inverse,shift", [
    (2.0, True, (1.0, -2.0)), (2.0, False, (1.0, -2.0)),
    (None, True, (0.5, 0.5)), (None, False, (3.0, -1.5)),
])
def test_mft_reference_fixture_modes(rt, prec, focal, inverse, shift):
    """The reference's MFT test fixture (32 -> 16, wl 1, dx 0.1, p 0.05), every mode."""
    print(f"[mft-modes] prec={prec} focal={focal} inverse={inverse} shift={shift}", flush=True)
    rng = np.random.default_rng(7)
    E = _rand_field(rng, 32)
    out = rt.mft(E, 1.0, 0.1, 16, 0.05, focal_length=focal, shift_x=shift[0],
                 shift_y=shift[1], inverse=inverse, precision=prec).cpu().numpy()
    ref = orc.mft(E, 1.0, 0.1, 16, 0.05, focal_length=focal, shift=shift, inverse=inverse)
    assert np.all(np.isfinite(out))
    assert _peak_err(out, ref) < FIELD_TOL[prec]


def test_mft_batched_per_instance_params(rt):
    print("[mft-batched]", flush=True)
    rng = np.random.default_rng(3)
    B, n_in, n_out = 5, 64, 48
    E = np.stack([_rand_field(rng, n_in) for _ in range(B)])
    wls = np.linspace(0.8e-6, 1.2e-6, B)
    dxs = np.linspace(0.9, 1.1, B) / n_in
    ps = np.linspace(2e-7, 3e-7, B)
    fs = np.linspace(0.5, 1.5, B)
    out = rt.mft(E, wls, dxs, n_out, ps, field_scale=fs).cpu().numpy()
    for b in range(B):
        ref = orc.mft(E[b], wls[b], dxs[b], n_out, ps[b]) * fs[b]
        assert _peak_err(out[b], ref) < FIELD_TOL["fp32"]


def test_transfer_matrix(rt):
    print("[transfer-matrix]", flush=True)
    wls = np.array([0.8e-6, 1.0e-6, 1.2e-6])
    W = rt.transfer_matrix_T(wls, 64, 1 / 64, 32, 2.5e-7, shifts=0.75).cpu().numpy()
    for b, wl in enumerate(wls):
        ref = orc.transfer_matrix(wl, 64, 1 / 64, 32, 2.5e-7, shift=0.75)
        assert np.max(np.abs(W[b].T - ref)) < 1e-6


def test_pupil_field(rt):
    print("[pupil-field]", flush=True)
    rng = np.random.default_rng(11)
    n, D = 64, 1.0
    amp = (rng.random((n, n)) > 0.3).astype(np.float32)
    opd = (rng.standard_normal((n, n)) * 50e-9).astype(np.float32)
    phase = (rng.standard_normal((n, n)) * 0.3).astype(np.float32)
    wls = np.array([0.9e-6, 1.1e-6])
    tilt = (1e-6, -2e-6)
    out = rt.pupil_field(n, wls, D / n, tilt=tilt, amp=amp, opd=opd, phase=phase,
                         scale=1.0 / n ** 2).cpu().numpy()
    for b, wl in enumerate(wls):
        E = orc.wavefront_init(n)
        E = orc.add_opd(E, wl, orc.tilt_opd(n, D / n, tilt))
        E = E * amp
        E = orc.add_opd(E, wl, opd.astype(np.float64))
        E = E * np.exp(1j * phase.astype(np.float64))
        assert _peak_err(out[b], E) < 1e-6


def test_eval_basis_power_intensity(rt):
    print("[basis/power/intensity]", flush=True)
    rng = np.random.default_rng(5)
    basis = rng.standard_normal((10, 37, 37)).astype(np.float32)
    c = rng.standard_normal(10).astype(np.float32)
    out = rt.eval_basis(basis, c).cpu().numpy()
    ref = np.tensordot(basis.astype(np.float64), c.astype(np.float64), axes=(0, 0))
    assert np.max(np.abs(out - ref)) < 1e-5 * np.max(np.abs(ref))
    F = (rng.standard_normal((3, 40, 40)) + 1j * rng.standard_normal((3, 40, 40))).astype(
        np.complex64)
    Ft = torch.as_tensor(F).cuda()
    pw = rt.power(Ft).cpu().numpy()
    assert np.allclose(pw, np.sum(np.abs(F.astype(np.complex128)) ** 2, axis=(1, 2)), rtol=1e-5)
    I = rt.intensity(Ft).cpu().numpy()
    assert np.allclose(I, np.abs(F) ** 2, rtol=1e-6)
    Is = rt.intensity(Ft, sum_batch=True).cpu().numpy()
    assert np.allclose(Is, np.sum(np.abs(F) ** 2, 0), rtol=1e-5)


# (modes, side): 36 modes at 512^2 (the Zernike OPD of the polychromatic PSF),
# pixel counts that are and are not multiples of 4 (16-byte loads or the
# scalar tail), a partial last workgroup (100^2), and more modes than the
# loop's unroll factor
@pytest.mark.parametrize("nb,side", [(36, 512), (37, 100), (9, 100), (8, 64), (7, 64), (40, 37)])
def test_eval_basis_vs_fp64(rt, nb, side):
    print(f"[basis] {nb} modes {side}^2", flush=True)
    rng = np.random.default_rng(nb * 1000 + side)
    basis = rng.standard_normal((nb, side, side)).astype(np.float32)
    c = rng.standard_normal(nb).astype(np.float32)
    base = rng.standard_normal((side, side)).astype(np.float32)
    ref = np.tensordot(basis.astype(np.float64), c.astype(np.float64), axes=(0, 0))
    out = rt.eval_basis(basis, c).cpu().numpy()
    assert np.max(np.abs(out - ref)) < 1e-5 * np.max(np.abs(ref))
    out_b = rt.eval_basis(basis, c, base=base).cpu().numpy()
    assert np.max(np.abs(out_b - (ref + base))) < 1e-5 * np.max(np.abs(ref + base))


@pytest.mark.parametrize("prec", ["fp32", "bf16x3"])
def test_psf_poly_small(rt, prec):
    """Fused chain vs oracle.poly_psf: aperture + spiders + Zernikes, 8 wavelengths."""
    print(f"[psf-poly] prec={prec}", flush=True)
    n, D, n_out = 128, 1.0, 64
    mask, basis = orc.circular_aperture(n, D, oversample=3, spider_width=0.02,
                                        spider_angles=[0, 120, 240],
                                        zernike_nolls=range(2, 12))
    rng = np.random.default_rng(1)
    coeffs = rng.normal(0, 30e-9, basis.shape[0])
    wls = np.linspace(0.8e-6, 1.2e-6, 8)
    w = np.ones(8) / 8
    p_out = 1e-6 / (4 * D)
    ref = orc.poly_psf(mask, basis, coeffs, wls, w, D, n_out, p_out)
    opd = np.tensordot(basis, coeffs, axes=(0, 0))
    norm = 1.0 / np.sqrt(np.sum((mask / n ** 2) ** 2))
    psf = rt.psf_poly(n, n_out, wls, w, D / n, p_out, amp=mask, opd=opd,
                      scale=norm / n ** 2, precision=prec).cpu().numpy()
    err = _peak_err(psf, ref)
    print(f"  peak-normalised error {err:.3e}")
    assert err < 1e-5


@pytest.mark.parametrize("prec", ["fp32", "bf16x3"])
def test_mft_batch_above_launch_chunk(rt, prec):
    """More instances than one launch carries (64): per-instance parameters stay aligned."""
    print(f"[mft-batch70] prec={prec}", flush=True)
    rng = np.random.default_rng(70)
    B, n_in, n_out = 70, 16, 12
    E = rng.standard_normal((B, n_in, n_in)) + 1j * rng.standard_normal((B, n_in, n_in))
    wls = np.linspace(0.5e-6, 1.5e-6, B)
    out = rt.mft(E, wls, 1.0 / n_in, n_out, 3e-7, shift_x=0.25, shift_y=-0.5,
                 precision=prec).cpu().numpy()
    for b in (0, 33, 63, 64, 69):
        ref = orc.mft(E[b], wls[b], 1.0 / n_in, n_out, 3e-7, shift=(0.25, -0.5))
        assert _peak_err(out[b], ref) < FIELD_TOL[prec]


def test_matrix_cache_hits_and_bypass(rt):
    print("[cache]", flush=True)
    n, n_out = 64, 32
    mask = orc.circular_aperture(n, 1.0, oversample=1)
    wls = np.linspace(0.9e-6, 1.1e-6, 5)
    w = np.ones(5) / 5
    rt.cache_clear()
    s0 = rt.cache_stats()
    a = rt.psf_poly(n, n_out, wls, w, 1.0 / n, 2.5e-7, amp=mask, scale=1.0 / n ** 2,
                    precision="bf16x3").cpu().numpy()
    b = rt.psf_poly(n, n_out, wls, w, 1.0 / n, 2.5e-7, amp=mask, scale=1.0 / n ** 2,
                    precision="bf16x3").cpu().numpy()
    s1 = rt.cache_stats()
    assert s1["misses"] == s0["misses"] + 1 and s1["hits"] == s0["hits"] + 1
    assert s1["entries"] == 1 and s1["bytes"] > 0
    assert np.array_equal(a, b)
    rt.cache_set_budget(0)
    try:
        c = rt.psf_poly(n, n_out, wls, w, 1.0 / n, 2.5e-7, amp=mask, scale=1.0 / n ** 2,
                        precision="bf16x3").cpu().numpy()
    finally:
        rt.cache_set_budget(4 << 30)
    assert np.array_equal(a, c)
    rt.cache_clear()
    assert rt.cache_stats()["entries"] == 0


def test_matrix_cache_eviction_reuses_blocks(rt):
    """A budget that holds about two entries: later wavelength grids evict earlier
    ones, their blocks are reused, and every result equals the uncached one."""
    print("[cache-evict]", flush=True)
    n, n_out = 64, 32
    mask = orc.circular_aperture(n, 1.0, oversample=1)
    grids = [np.linspace(0.8e-6 + 0.01e-6 * k, 1.2e-6, 6) for k in range(5)]
    w = np.ones(6) / 6
    kw = dict(amp=mask, scale=1.0 / n ** 2, precision="bf16x3")
    rt.cache_set_budget(0)
    try:
        refs = [rt.psf_poly(n, n_out, g, w, 1.0 / n, 2.5e-7, **kw).cpu().numpy() for g in grids]
    finally:
        rt.cache_set_budget(16 << 30)
    rt.cache_clear()
    per_entry = 6 * n_out * n * 8
    rt.cache_set_budget(int(2.5 * per_entry) + 4096)
    try:
        for rep in range(2):
            for g, ref in zip(grids, refs):
                out = rt.psf_poly(n, n_out, g, w, 1.0 / n, 2.5e-7, **kw).cpu().numpy()
                assert np.array_equal(out, ref)
        assert rt.cache_stats()["entries"] <= 2
    finally:
        rt.cache_set_budget(16 << 30)
        rt.cache_clear()


def test_psf_poly_vjp_under_cache_pressure(rt):
    """The gradient holds the forward and the adjoint matrices at once.  With a
    budget between one and two entries, and the forward entry already resident
    from a forward call, the adjoint acquire must not evict (and recycle) the
    forward entry the same call still reads: the gradient equals the uncached one."""
    print("[cache-vjp]", flush=True)
    n, n_out, nwl = 96, 48, 6
    mask = orc.circular_aperture(n, 1.0, oversample=1)
    rng = np.random.default_rng(5)
    opd = torch.as_tensor(rng.normal(0, 30e-9, (n, n)), dtype=torch.float32, device="cuda")
    wls = np.linspace(0.9e-6, 1.1e-6, nwl)
    w = np.ones(nwl) / nwl
    G = torch.as_tensor(rng.normal(size=(n_out, n_out)), dtype=torch.float32, device="cuda")
    kw = dict(amp=mask, opd=opd, scale=1.0 / n ** 2, precision="bf16x3")
    rt.cache_clear()
    rt.cache_set_budget(0)
    try:
        ref_psf = rt.psf_poly(n, n_out, wls, w, 1.0 / n, 2.5e-7, **kw).cpu().numpy()
        ref = rt.psf_poly_vjp(n, n_out, wls, w, 1.0 / n, 2.5e-7, G, **kw)[0].cpu().numpy()
    finally:
        rt.cache_set_budget(16 << 30)
    per_entry = nwl * n_out * n * 8
    rt.cache_set_budget(int(1.5 * per_entry))
    try:
        for _ in range(2):
            psf = rt.psf_poly(n, n_out, wls, w, 1.0 / n, 2.5e-7, **kw).cpu().numpy()
            g = rt.psf_poly_vjp(n, n_out, wls, w, 1.0 / n, 2.5e-7, G, **kw)[0].cpu().numpy()
            assert np.array_equal(psf, ref_psf)
            assert np.array_equal(g, ref)
        assert rt.cache_stats()["entries"] == 1
    finally:
        rt.cache_set_budget(16 << 30)
        rt.cache_clear()


def test_cache_entry_shared_by_two_streams_survives_eviction(rt):
    """An entry read by queued work on one stream and then by another stream,
    then evicted from the second stream: the block is recycled only after both
    streams' reads, so every result equals the uncached one."""
    print("[cache-streams]", flush=True)
    n, n_out, nwl = 256, 128, 8
    mask = orc.circular_aperture(n, 1.0, oversample=1)
    grids = [np.linspace(0.8e-6 + 0.01e-6 * k, 1.2e-6, nwl) for k in range(3)]
    w = np.ones(nwl) / nwl
    kw = dict(amp=mask, scale=1.0 / n ** 2, precision="bf16x3")
    rt.cache_clear()
    rt.cache_set_budget(0)
    try:
        refs = [rt.psf_poly(n, n_out, g, w, 1.0 / n, 2.5e-7, **kw).cpu().numpy() for g in grids]
    finally:
        rt.cache_set_budget(16 << 30)
    per_entry = nwl * n_out * n * 8
    rt.cache_set_budget(int(1.5 * per_entry))
    sa, sb = torch.cuda.Stream(), torch.cuda.Stream()
    try:
        torch.cuda.synchronize()
        with torch.cuda.stream(sa):
            outs_a = [rt.psf_poly(n, n_out, grids[0], w, 1.0 / n, 2.5e-7, **kw)
                      for _ in range(24)]
        with torch.cuda.stream(sb):
            outs_b = [rt.psf_poly(n, n_out, g, w, 1.0 / n, 2.5e-7, **kw)
                      for g in (grids[0], grids[1], grids[2], grids[0])]
        torch.cuda.synchronize()
        for o in outs_a:
            assert np.array_equal(o.cpu().numpy(), refs[0])
        for o, k in zip(outs_b, (0, 1, 2, 0)):
            assert np.array_equal(o.cpu().numpy(), refs[k])
    finally:
        torch.cuda.synchronize()
        rt.cache_set_budget(16 << 30)
        rt.cache_clear()


def test_cache_bypass_skips_the_cache_and_keeps_results(rt):
    """Calls inside runtime.cache_bypass() neither hit nor fill the matrix cache
    (dlx_cache_bypass is per thread) and give the cached calls' results."""
    print("[cache-bypass]", flush=True)
    n, n_out = 64, 32
    mask = orc.circular_aperture(n, 1.0, oversample=1)
    wls = np.linspace(0.9e-6, 1.1e-6, 4)
    w = np.ones(4) / 4
    kw = dict(amp=mask, scale=1.0 / n ** 2, precision="bf16x3")
    rt.cache_clear()
    a = rt.psf_poly(n, n_out, wls, w, 1.0 / n, 2.5e-7, **kw).cpu().numpy()
    s0 = rt.cache_stats()
    with rt.cache_bypass():
        b = rt.psf_poly(n, n_out, wls, w, 1.0 / n, 2.5e-7, **kw).cpu().numpy()
        c = rt.psf_poly(n, n_out, wls * 1.01, w, 1.0 / n, 2.5e-7, **kw).cpu().numpy()
    s1 = rt.cache_stats()
    assert (s1["hits"], s1["misses"], s1["entries"]) == (s0["hits"], s0["misses"], s0["entries"])
    assert np.array_equal(a, b)
    d = rt.psf_poly(n, n_out, wls * 1.01, w, 1.0 / n, 2.5e-7, **kw).cpu().numpy()
    assert np.array_equal(c, d)
    assert rt.cache_stats()["misses"] == s0["misses"] + 1
    rt.cache_clear()


def test_mfma_rate_probe(rt):
    """The matrix-core probe returns positive rates no larger than the nominal
    2.5 PFLOP/s (plus 10 %), full-entropy operands no faster than low-entropy ones
    (plus 10 %)."""
    print("[mfma-rate]", flush=True)
    hi = rt.mfma_rate(False)
    lo = rt.mfma_rate(True)
    print(f"  low-entropy {hi:.0f} TF/s, full-entropy {lo:.0f} TF/s")
    assert 0 < hi < 2750 and 0 < lo < 2750
    assert lo < 1.1 * hi
