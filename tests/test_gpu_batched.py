"""GPU parity of the uneven batched MFT (dlx_mft_batched_psf) against the FP64
oracle: mixed pupil and output sizes (odd and non-multiple-of-64 ones
included), per-instance wavelengths, shifts and weights, several outputs.

Error metric: peak-normalised max |I - I_ref| / max I_ref <= 1e-5 per output
image (the PSF bound of the hot path; the contractions are split-BF16x3).
"""

import numpy as np
import pytest

from oracle import dlux_oracle as orc

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

TOL = 1e-5


@pytest.fixture(scope="module")
def rt():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from dlux_amd import runtime
    return runtime


def _peak_err(x, ref):
    return float(np.max(np.abs(np.asarray(x) - ref)) / np.max(np.abs(ref)))


def _pupil(n, seed):
    """Unit-power circular pupil with a random low-order aberration (complex128)."""
    mask, basis = orc.circular_aperture(n, 1.0, oversample=1, zernike_nolls=range(2, 9))
    c = np.random.default_rng(seed).normal(0, 40e-9, basis.shape[0])
    # odd grids put a pixel at rho = 0, where some reference Zernike terms are NaN
    opd = np.tensordot(np.nan_to_num(basis), c, axes=(0, 0))
    f = mask * np.exp(2j * np.pi * opd / 1e-6)
    return f / np.sqrt(np.sum(np.abs(f) ** 2))


def _oracle_outputs(inst, n_outputs):
    outs = [None] * n_outputs
    for d in inst:
        n_in = d["field_np"].shape[0]
        E = orc.mft(d["field_np"], d["wavelength"], d["pixel_scale_in"], d["n_out"],
                    d["pixel_scale_out"], shift=d["shift"])
        I = d["weight"] * np.abs(E) ** 2
        o = d["out_index"]
        outs[o] = I if outs[o] is None else outs[o] + I
        del n_in
    return outs


def _make(rng, sizes, n_inst, n_outputs, fields_per_size=1):
    fields = {}
    for k, n in enumerate(sizes):
        for q in range(fields_per_size):
            f = _pupil(n, 100 * k + q)
            fields[(n, q)] = (f, torch.as_tensor(f.astype(np.complex64)).cuda())
    out_nout = {}
    inst = []
    for i in range(n_inst):
        n = sizes[rng.integers(len(sizes))]
        f_np, f_t = fields[(n, int(rng.integers(fields_per_size)))]
        o = int(rng.integers(n_outputs))
        n_out = out_nout.setdefault(o, int(rng.choice([8, 17, 40, 64, 96])))
        inst.append(dict(field=f_t, field_np=f_np, wavelength=float(rng.uniform(0.8e-6, 1.2e-6)),
                         pixel_scale_in=1.0 / n, n_out=n_out,
                         pixel_scale_out=float(rng.uniform(2e-7, 3e-7)),
                         shift=(float(rng.uniform(-3, 3)), float(rng.uniform(-3, 3))),
                         weight=float(rng.uniform(0.1, 2.0)), out_index=o))
    return inst


def test_batched_uneven_small_vs_oracle(rt):
    print("[batched] uneven small mix, 24 instances, 5 outputs", flush=True)
    rng = np.random.default_rng(11)
    inst = _make(rng, [16, 33, 64, 130], 24, 5)
    outs = rt.mft_batched_psf(inst, n_outputs=5)
    ref = _oracle_outputs(inst, 5)
    for o in range(5):
        if ref[o] is None:
            assert outs[o] is None
            continue
        err = _peak_err(outs[o].cpu().numpy(), ref[o])
        print(f"  output {o}: {tuple(outs[o].shape)} err {err:.3e}")
        assert err < TOL


def test_batched_many_problems_vs_oracle(rt):
    """Enough tiles for every XCD range and the persistent loop to run several tiles."""
    print("[batched] 160 instances, sizes 64..256", flush=True)
    rng = np.random.default_rng(12)
    inst = _make(rng, [64, 128, 200, 256], 160, 12, fields_per_size=2)
    outs = rt.mft_batched_psf(inst, n_outputs=12)
    ref = _oracle_outputs(inst, 12)
    worst = 0.0
    for o in range(12):
        if ref[o] is not None:
            worst = max(worst, _peak_err(outs[o].cpu().numpy(), ref[o]))
    print(f"  worst err {worst:.3e}")
    assert worst < TOL


def test_batched_matches_fused_tilted_psf(rt):
    """A tilt is an MFT shift of tilt / p_out pixels in intensity: the batched
    path with shifts equals the fused polychromatic PSF with the tilt applied
    in the pupil (point-source model of sources.py:392-411)."""
    print("[batched] shift == tilt", flush=True)
    n, n_out, p = 96, 48, 2.5e-7
    mask = orc.circular_aperture(n, 1.0, oversample=1)
    wls = np.linspace(0.9e-6, 1.1e-6, 4)
    w = np.array([0.1, 0.2, 0.3, 0.4])
    tilt = (3.3 * p, -1.7 * p)
    norm = 1.0 / np.sqrt(np.sum((mask / n ** 2) ** 2))
    fused = rt.psf_poly(n, n_out, wls, w, 1.0 / n, p, amp=mask, scale=norm / n ** 2,
                        tilt=tilt, precision="bf16x3").cpu().numpy()
    field = torch.as_tensor((mask * norm / n ** 2).astype(np.complex64)).cuda()
    inst = [dict(field=field, wavelength=wl, pixel_scale_in=1.0 / n, n_out=n_out,
                 pixel_scale_out=p, shift=(tilt[0] / p, tilt[1] / p), weight=wt)
            for wl, wt in zip(wls, w)]
    out = rt.mft_batched_psf(inst)[0].cpu().numpy()
    err = _peak_err(out, fused)
    print(f"  err {err:.3e}")
    assert err < TOL


def test_batched_leaves_empty_outputs_and_rejects_bad_input(rt):
    print("[batched] edge cases", flush=True)
    n = 32
    f = torch.as_tensor(_pupil(n, 1).astype(np.complex64)).cuda()
    sentinel = torch.full((8, 8), 7.0, device="cuda")
    o0 = torch.empty((16, 16), device="cuda")
    inst = [dict(field=f, wavelength=1e-6, pixel_scale_in=1 / n, n_out=16,
                 pixel_scale_out=2.5e-7, out_index=0, weight=0.0)]
    outs = rt.mft_batched_psf(inst, n_outputs=2, outputs=[o0, sentinel])
    assert torch.all(outs[0] == 0)
    assert torch.all(sentinel == 7.0)
    bad = dict(inst[0], n_out=12)
    with pytest.raises(RuntimeError, match="different n_out"):
        rt.mft_batched_psf([inst[0], bad])


def test_batched_matrices_stay_resident(rt):
    """A batch seen twice keeps its transfer matrices in the device cache: the
    third call hits.  Every call (first sighting generating into the
    workspace, second filling the cache, third reading it, and with the cache
    disabled) gives the same images."""
    print("[batched] matrix cache", flush=True)
    rng = np.random.default_rng(13)
    inst = _make(rng, [32, 48], 10, 3)
    rt.cache_clear()
    a = [o.clone() for o in rt.mft_batched_psf(inst, n_outputs=3)]
    b = [o.clone() for o in rt.mft_batched_psf(inst, n_outputs=3)]
    s0 = rt.cache_stats()
    c = [o.clone() for o in rt.mft_batched_psf(inst, n_outputs=3)]
    s1 = rt.cache_stats()
    assert s1["hits"] == s0["hits"] + 1 and s1["misses"] == s0["misses"]
    rt.cache_set_budget(0)
    try:
        d = rt.mft_batched_psf(inst, n_outputs=3)
    finally:
        rt.cache_set_budget(16 << 30)
    for x, y, z, u in zip(a, b, c, d):
        assert torch.equal(x, y) and torch.equal(x, z) and torch.equal(x, u)


def test_batched_from_arrays_equals_dict_path(rt):
    """The vectorised descriptor construction gives the same images as the
    per-instance dictionaries (stacked fields addressed by slice)."""
    print("[batched] from_arrays", flush=True)
    rng = np.random.default_rng(14)
    n, nl = 40, 5
    stack = torch.stack([torch.as_tensor(_pupil(n, 200 + l).astype(np.complex64))
                         for l in range(nl)]).cuda()
    wls = np.linspace(0.9e-6, 1.1e-6, nl)
    sx, sy = rng.uniform(-2, 2, 3), rng.uniform(-2, 2, 3)
    w = rng.uniform(0.1, 1.0, (3, nl))
    inst = [dict(field=stack[l], wavelength=wls[l], pixel_scale_in=1.0 / n, n_out=24,
                 pixel_scale_out=2.5e-7, shift=(sx[s], sy[s]), weight=w[s, l], out_index=s)
            for s in range(3) for l in range(nl)]
    a = rt.mft_batched_psf(inst, n_outputs=3)
    b = rt.MftBatch.from_arrays([stack], np.zeros(3 * nl, np.int64), np.tile(wls, 3), 1.0 / n,
                                24, 2.5e-7, w.reshape(-1), np.repeat(np.arange(3), nl),
                                shift_x=np.repeat(sx, nl), shift_y=np.repeat(sy, nl),
                                slice_index=np.tile(np.arange(nl), 3), n_outputs=3)()
    for x, y in zip(a, b):
        assert torch.equal(x, y)
    with pytest.raises(ValueError, match="outside"):
        rt.MftBatch.from_arrays([stack], [0], 1e-6, 1.0 / n, 24, 2.5e-7, 1.0, 0,
                                slice_index=nl)


def test_batched_graph_replay_equals_eager(rt):
    """The batched launches run as a captured HIP graph from the second call
    with the same workspace and batch shape; every graph launch equals the
    eager launches bit for bit (same kernels, same inputs), call after call,
    with new field contents (matrices resident) or new shifts (matrices
    generated inside the graph, cache off), and eager work and
    synchronisations between the calls; on the default stream and on a side
    stream."""
    import contextlib
    print("[batched] graph replay", flush=True)
    rng = np.random.default_rng(21)
    inst = _make(rng, [32, 64], 12, 3)
    uniq = list({d["field"].data_ptr(): d["field"] for d in inst}.values())

    def set_fields(k):
        for q, f in enumerate(uniq):
            g = _pupil(f.shape[0], 500 + 10 * k + q).astype(np.complex64)
            f.copy_(torch.as_tensor(g, device=f.device))

    def set_shifts(b, k):
        r = np.random.default_rng(700 + k)
        for i in range(b.n):
            b.arr[i].shift_x = float(r.uniform(-3, 3))
            b.arr[i].shift_y = float(r.uniform(-3, 3)) + 0.5   # never equal to shift_x

    def run(b):
        out = [None if o is None else o.clone() for o in b()]
        x = torch.randn(256, 256, device="cuda")
        (x @ x).sum().item()   # eager work and a synchronisation between calls
        return out

    def same(a, b):
        return all((x is None and y is None) or torch.equal(x, y) for x, y in zip(a, b))

    prev = rt.batched_graphs(None)
    try:
        for side in (False, True):
            ctx = torch.cuda.stream(torch.cuda.Stream()) if side else contextlib.nullcontext()
            with ctx:
                b = rt.MftBatch(inst, n_outputs=3)
                # new field contents, matrices resident after the first calls
                rt.batched_graphs(False)
                ref = []
                for k in range(3):
                    set_fields(k)
                    ref.append(run(b))
                rt.batched_graphs(True)
                s0 = rt.batched_graph_stats()
                for call in range(6):
                    set_fields(call % 3)
                    assert same(run(b), ref[call % 3]), f"side={side} call {call}"
                s1 = rt.batched_graph_stats()
                print(f"  side={side} fields: {s0} -> {s1}")
                assert s1["captures"] == s0["captures"] + 1
                assert s1["launches"] == s0["launches"] + 5
                # new shifts, cache off: matrices generated in every call
                rt.cache_set_budget(0)
                try:
                    rt.batched_graphs(False)
                    ref = []
                    for k in range(3):
                        set_shifts(b, k)
                        ref.append(run(b))
                    rt.batched_graphs(True)
                    s0 = rt.batched_graph_stats()
                    for call in range(6):
                        set_shifts(b, call % 3)
                        assert same(run(b), ref[call % 3]), f"side={side} shifts call {call}"
                    s1 = rt.batched_graph_stats()
                finally:
                    rt.cache_set_budget(16 << 30)
                print(f"  side={side} shifts: {s0} -> {s1}")
                assert s1["captures"] == s0["captures"] + 1
                assert s1["launches"] == s0["launches"] + 5
            torch.cuda.synchronize()
        # the replayed results are the oracle's too
        set_fields(0)
        b = rt.MftBatch(inst, n_outputs=3)
        for _ in range(3):
            outs = b()
        for d in inst:
            d["field_np"] = d["field"].cpu().numpy().astype(np.complex128)
        refo = _oracle_outputs(inst, 3)
        for o in range(3):
            if refo[o] is not None:
                assert _peak_err(outs[o].cpu().numpy(), refo[o]) < TOL
    finally:
        rt.batched_graphs(prev)
