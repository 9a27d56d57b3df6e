"""Continuous-batching request loop for PSF, gradient and angular-spectrum
requests on one GPU.

Requests arrive over time (a Poisson stream in the benchmark).  The loop
runs on one host thread:

  * arrivals join a FIFO queue;
  * a request is admitted when an in-flight slot is free and its HBM
    footprint fits the budget next to everything still on the device (the
    head of the queue waits; nothing overtakes it);
  * all PSF requests admitted in one pass are fused into ONE uneven batched
    MFT call (``dlx_mft_batched_psf``: grouped persistent launches over every
    (request, wavelength) instance, whatever their pupil, output and
    wavelength counts); gradient requests run forward + hand-written VJP +
    basis projection; ASM requests run the padded angular-spectrum
    propagator;
  * work is issued on a small pool of streams (default priority), each group
    on the stream with the least work queued (estimated device time), so a
    short request never waits behind a long one while another stream is free;
    each launch group records an event; the loop polls events, finalises the
    requests of completed groups (latency = completion - arrival) and frees
    their memory, so later arrivals are admitted while earlier work still
    runs and their kernels overlap on the device;
  * every stream has one workspace arena, grown to the largest group it ran
    and reused (work on a stream is ordered), and the loop's thread skips the
    library's transfer-matrix cache (the mix's wavelength grids rarely
    repeat), so serving allocates nothing per request beyond its results
    and never waits on the host for an eviction.

Pupil masks, Zernike bases and their normalisation are built once per pupil
size and stay resident.

The PSF of a request is sum_l w_l |MFT_l(P_l)|^2 with flat weights (the
OpticalSystem.propagate chain of optical_systems.py:147-223); a gradient
request returns the Gaussian loss 0.5 sum (F psf - d)^2 / max(d, 1) and its
gradient with respect to the Zernike coefficients.
"""

from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np
import torch

from . import runtime
from .workloads import D, P_OUT


@dataclass
class Request:
    rid: int
    kind: str                       # "psf" | "grad" | "asm"
    arrival: float = 0.0            # seconds after the start of the run
    n_in: int = 0
    n_out: int = 0
    wavelengths: np.ndarray | None = None
    coeffs: np.ndarray | None = None
    data_seed: int = 0              # asm: seed of the input field's phase
    data: np.ndarray | None = None  # grad: the observed image (n_out x n_out counts)
    flux: float = 1e6
    distance: float = 0.0           # asm: metres
    # filled in by the server
    t_admit: float | None = None
    t_done: float | None = None
    result: object = None
    error: str | None = None

    @property
    def latency(self) -> float | None:
        return None if self.t_done is None else self.t_done - self.arrival

    def copy_inputs(self) -> "Request":
        return Request(self.rid, self.kind, 0.0, self.n_in, self.n_out, self.wavelengths,
                       self.coeffs, self.data_seed, self.data, self.flux, self.distance)


@dataclass(frozen=True)
class MixConfig:
    """The serving mix: 75% PSF, 20% gradient, 5% ASM requests."""
    n_in_choices: tuple = (256, 512, 1024)
    n_out_choices: tuple = (64, 128, 256)
    max_wl: int = 64
    wl_lo: float = 0.8e-6
    wl_hi: float = 1.2e-6
    js: tuple = tuple(range(2, 38))
    coeff_std: float = 30e-9
    p_psf: float = 0.75
    p_grad: float = 0.20
    asm_n: int = 1024
    asm_m: int = 2048
    asm_wavelength: float = 1e-6
    asm_pixel: float = 2e-6
    seed: int = 6


MIX = MixConfig()


def make_requests(rate: float, duration: float, cfg: MixConfig = MIX, seed: int | None = None,
                  n_max: int | None = None) -> list:
    """Poisson arrivals at `rate` requests/s over `duration` seconds (or n_max requests)."""
    rng = np.random.default_rng(cfg.seed if seed is None else seed)
    reqs, t = [], 0.0
    while True:
        t += rng.exponential(1.0 / rate) if rate > 0 else 0.0
        if (duration and t > duration) or (n_max is not None and len(reqs) >= n_max):
            break
        u = rng.random()
        kind = "psf" if u < cfg.p_psf else ("grad" if u < cfg.p_psf + cfg.p_grad else "asm")
        rid = len(reqs)
        if kind == "asm":
            reqs.append(Request(rid, "asm", t, n_in=cfg.asm_n, n_out=cfg.asm_m,
                                wavelengths=np.array([cfg.asm_wavelength]),
                                data_seed=int(rng.integers(1 << 30)),
                                distance=float(rng.uniform(1e-3, 100e-3))))
            continue
        n_wl = int(rng.integers(1, cfg.max_wl + 1))
        n_out = int(rng.choice(cfg.n_out_choices))
        r = Request(rid, kind, t, n_in=int(rng.choice(cfg.n_in_choices)), n_out=n_out,
                    wavelengths=(np.linspace(cfg.wl_lo, cfg.wl_hi, n_wl)
                                 if n_wl > 1 else np.array([1e-6])),
                    coeffs=rng.normal(0.0, cfg.coeff_std, len(cfg.js)))
        if kind == "grad":
            # a synthetic observation: Poisson counts around a smooth blob
            y, x = np.mgrid[:n_out, :n_out] - (n_out - 1) / 2
            blob = np.exp(-(x ** 2 + y ** 2) / (2 * (n_out / 16) ** 2))
            r.data = rng.poisson(r.flux * blob / blob.sum()).astype(np.float32)
        reqs.append(r)
    return reqs


class _Group:
    def __init__(self, reqs, event, keep, nbytes, finish, stream=0, cost=0.0):
        self.reqs, self.event, self.keep, self.nbytes, self.finish = (reqs, event, keep,
                                                                      nbytes, finish)
        self.stream, self.cost = stream, cost


class Server:
    """Continuous-batching loop (see the module docstring)."""

    def __init__(self, hbm_budget: int = 48 << 30, max_inflight: int = 256,
                 n_streams: int = 4, precision: str = "bf16x3", cfg: MixConfig = MIX):
        self.dev = runtime.device()
        self.budget = int(hbm_budget)
        self.max_inflight = int(max_inflight)
        self.streams = [torch.cuda.Stream(device=self.dev) for _ in range(n_streams)]
        self.precision = precision
        self.cfg = cfg
        self._geo = {}
        # estimated device time (s) of the work queued on each stream and not yet done
        self._queued = [0.0] * n_streams
        self._arena = [None] * n_streams   # per-stream workspace, reused
        self._cur = 0                      # stream index of the group being launched

    # ---- resident pupil data --------------------------------------------
    def geometry(self, n: int):
        """(mask, basis, pupil scale) for pupil size n, built once and kept on the device."""
        if n not in self._geo:
            from .utils import apertures, coordinates, zernikes
            mask = torch.as_tensor(apertures.circular_aperture(n, D, oversample=1),
                                   device=self.dev).float().contiguous()
            basis = torch.as_tensor(
                zernikes.zernike_basis(self.cfg.js, coordinates.pixel_coords(n, diameter=D), D),
                device=self.dev).float().contiguous()
            # unit power after the reference's 1/N^2 initial phasor and normalise()
            scale = float(1.0 / torch.sqrt(torch.sum(mask.double() ** 2)))
            self._geo[n] = (mask, basis, scale)
        return self._geo[n]

    # ---- footprints -------------------------------------------------------
    def footprint(self, r: Request) -> int:
        if r.kind == "asm":
            lib = runtime._lib.load()
            return int(lib.dlx_asm_workspace_size(r.n_in, r.n_out, 1)) + 8 * (
                r.n_in ** 2 + r.n_out ** 2)
        nwl = len(r.wavelengths)
        if r.kind == "psf":
            # fields + split4 pupils + two matrices + U^T + intensity planes + output
            per = 16 * r.n_in ** 2 + 3 * 8 * r.n_out * r.n_in + 4 * r.n_out ** 2
            return nwl * per + 4 * r.n_out ** 2 + (1 << 16)
        g = runtime.geom(r.n_in, r.n_out)
        lib = runtime._lib.load()
        prec = runtime._prec_code(self.precision)
        return int(lib.dlx_psf_poly_vjp_workspace_size(g, nwl, prec)) + 16 * r.n_in ** 2 + \
            12 * r.n_out ** 2

    # ---- launches ---------------------------------------------------------
    def _upload(self, x, keep):
        """Asynchronous host-to-device copy on the current stream through pinned
        memory (a pageable copy would make the host wait for the stream's
        earlier work); the pinned staging tensor is kept until the group ends."""
        h = torch.from_numpy(np.ascontiguousarray(x, np.float32)).pin_memory()
        keep.append(h)
        return h.to(self.dev, non_blocking=True)

    # Device-time estimates used to place work: the contractions at ~300 TFLOP/s
    # (8 (N_out N_in^2 + N_out^2 N_in) FLOPs per wavelength, four contractions'
    # worth for a gradient) and the angular spectrum's three FFT launches (each
    # reading and writing the padded array) at ~2.3 TB/s.
    def cost(self, r: Request) -> float:
        if r.kind == "asm":
            return 3 * 16.0 * r.n_out ** 2 / 2.3e12
        f = 8.0 * (r.n_out * r.n_in ** 2 + r.n_out ** 2 * r.n_in) * len(r.wavelengths)
        return (4.0 if r.kind == "grad" else 1.0) * f / 300e12

    def _ws(self, nbytes: int) -> torch.Tensor:
        """The current stream's workspace arena, grown (x1.5) when too small."""
        a = self._arena[self._cur]
        if a is None or a.numel() < nbytes:
            size = max(int(nbytes), int(1.5 * (0 if a is None else a.numel())))
            a = self._arena[self._cur] = runtime.workspace(size, self.dev)
        return a

    def _stream(self):
        """The stream with the least queued work: a group never waits behind a
        long one on its stream while another stream is idle."""
        i = min(range(len(self.streams)), key=lambda k: self._queued[k])
        return i, self.streams[i]

    def _launch_psf(self, reqs):
        keep, fields = [], []
        cols = {k: [] for k in ("fi", "si", "wl", "dx", "nout", "w", "oi")}
        for j, r in enumerate(reqs):
            mask, basis, scale = self.geometry(r.n_in)
            c = self._upload(r.coeffs, keep)
            opd = runtime.eval_basis(basis, c)
            f = runtime.pupil_field(r.n_in, r.wavelengths, D / r.n_in, amp=mask, opd=opd,
                                    scale=scale)
            keep += [c, opd]
            nwl = len(r.wavelengths)
            cols["fi"].append(np.full(nwl, len(fields)))
            cols["si"].append(np.arange(nwl))
            cols["wl"].append(r.wavelengths)
            cols["dx"].append(np.full(nwl, D / r.n_in))
            cols["nout"].append(np.full(nwl, r.n_out))
            cols["w"].append(np.full(nwl, 1.0 / nwl))
            cols["oi"].append(np.full(nwl, j))
            fields.append(f)
        cat = {k: np.concatenate(v) for k, v in cols.items()}
        batch = runtime.MftBatch.from_arrays(
            fields, cat["fi"], cat["wl"], cat["dx"], cat["nout"], P_OUT, cat["w"], cat["oi"],
            slice_index=cat["si"], n_outputs=len(reqs), ws=self._arena[self._cur])
        if batch.ws is not self._arena[self._cur]:   # grew: keep the new one
            self._arena[self._cur] = batch.ws
        outs = batch()
        keep.append(batch)

        def finish():
            for r, o in zip(reqs, outs):
                r.result = o
        return keep, finish

    def _launch_grad(self, r: Request):
        mask, basis, scale = self.geometry(r.n_in)
        keep = []
        c = self._upload(r.coeffs, keep)
        opd = runtime.eval_basis(basis, c)
        w = np.full(len(r.wavelengths), 1.0 / len(r.wavelengths))
        kw = dict(amp=mask, opd=opd, scale=scale, precision=self.precision)
        lib = runtime._lib.load()
        g = runtime.geom(r.n_in, r.n_out)
        prec = runtime._prec_code(self.precision)
        ws = self._ws(max(int(lib.dlx_psf_poly_workspace_size(g, len(w), prec)),
                          int(lib.dlx_psf_poly_vjp_workspace_size(g, len(w), prec))))
        psf = runtime.psf_poly(r.n_in, r.n_out, r.wavelengths, w, D / r.n_in, P_OUT, ws=ws,
                               **kw)
        d = self._upload(r.data, keep)
        var = torch.clamp(d, min=1.0)
        resid = r.flux * psf - d
        loss = 0.5 * torch.sum(resid.double() ** 2 / var)
        G = (r.flux * resid / var).contiguous()
        g_opd, _ = runtime.psf_poly_vjp(r.n_in, r.n_out, r.wavelengths, w, D / r.n_in, P_OUT,
                                        G, ws=ws, **kw)
        g_c = runtime.project_basis(basis, g_opd)
        keep += [c, opd, psf, d, var, resid, G, g_opd]

        def finish():
            r.result = (loss, g_c)
        return keep, finish

    def _asm_amp(self, n: int):
        key = ("asm", n)
        if key not in self._geo:
            x = (torch.arange(n, dtype=torch.float64) - (n - 1) / 2) * self.cfg.asm_pixel
            rr = x[None, :] ** 2 + x[:, None] ** 2
            amp = torch.exp(-rr / (2 * (n * self.cfg.asm_pixel / 8) ** 2))
            self._geo[key] = amp.float().to(self.dev)
        return self._geo[key]

    def _launch_asm(self, r: Request):
        # the input field: a Gaussian beam with a seeded random phase screen, made on the device
        amp = self._asm_amp(r.n_in)
        g = torch.Generator(device=self.dev).manual_seed(r.data_seed)
        ph = torch.rand((r.n_in, r.n_in), generator=g, device=self.dev) * 0.1
        f = torch.polar(amp, ph).contiguous()
        out = runtime.asm_prop(f, r.wavelengths, self.cfg.asm_pixel, r.distance, r.n_out)
        keep = [f, ph]

        def finish():
            r.result = out
        return keep, finish

    # ---- the loop ---------------------------------------------------------
    def run(self, requests, realtime: bool = True, idle_sleep: float = 20e-6) -> dict:
        """Serve `requests` (sorted by arrival).  realtime=False admits every
        request as soon as slots and memory allow, ignoring arrival times."""
        reqs = sorted(requests, key=lambda q: q.arrival)
        pending, inflight = [], []
        used, n_inflight, nxt, done = 0, 0, 0, 0
        t0 = time.perf_counter()
        clock = (lambda: time.perf_counter() - t0) if realtime else (lambda: float("inf"))
        with runtime.precision(self.precision), runtime.cache_bypass():
            while done < len(reqs):
                now = clock()
                while nxt < len(reqs) and reqs[nxt].arrival <= now:
                    pending.append(reqs[nxt])
                    nxt += 1
                # completions
                still = []
                for grp in inflight:
                    if grp.event.query():
                        grp.finish()
                        t = time.perf_counter() - t0
                        for r in grp.reqs:
                            r.t_done = t
                        used -= grp.nbytes
                        self._queued[grp.stream] -= grp.cost
                        n_inflight -= len(grp.reqs)
                        done += len(grp.reqs)
                    else:
                        still.append(grp)
                inflight = still
                # admission (FIFO)
                admitted = []
                while pending and n_inflight + len(admitted) < self.max_inflight:
                    r = pending[0]
                    fp = self.footprint(r)
                    if fp > self.budget:
                        pending.pop(0)
                        r.error = f"request needs {fp} B, more than the HBM budget {self.budget}"
                        r.t_done = time.perf_counter() - t0
                        done += 1
                        continue
                    if used + fp > self.budget:
                        break
                    pending.pop(0)
                    r.t_admit = time.perf_counter() - t0
                    used += fp
                    admitted.append((r, fp))
                if admitted:
                    n_inflight += len(admitted)
                    psf = [(r, fp) for r, fp in admitted if r.kind == "psf"]
                    groups = []
                    if psf:
                        groups.append(([r for r, _ in psf], sum(fp for _, fp in psf),
                                       lambda rs=[r for r, _ in psf]: self._launch_psf(rs)))
                    for r, fp in admitted:
                        if r.kind == "grad":
                            groups.append(([r], fp, lambda r=r: self._launch_grad(r)))
                        elif r.kind == "asm":
                            groups.append(([r], fp, lambda r=r: self._launch_asm(r)))
                    for rs, nbytes, launch in groups:
                        c = sum(self.cost(r) for r in rs)
                        si, s = self._stream()
                        self._queued[si] += c
                        self._cur = si
                        with torch.cuda.stream(s):
                            keep, finish = launch()
                            ev = torch.cuda.Event()
                            ev.record(s)
                        inflight.append(_Group(rs, ev, keep, nbytes, finish, si, c))
                elif not inflight and nxt < len(reqs) and realtime:
                    wait = reqs[nxt].arrival - clock()
                    if wait > 0:
                        time.sleep(min(wait, 1e-3))
                elif inflight:
                    time.sleep(idle_sleep)
            wall = time.perf_counter() - t0
        lat = np.array([r.latency for r in reqs if r.error is None and r.latency is not None])
        by_kind = {}
        for k in ("psf", "grad", "asm"):
            lk = np.array([r.latency for r in reqs if r.kind == k and r.error is None])
            if lk.size:
                by_kind[k] = {"n": int(lk.size), "p50_ms": float(np.percentile(lk, 50) * 1e3),
                              "p99_ms": float(np.percentile(lk, 99) * 1e3)}
        return {"requests": len(reqs), "rejected": sum(r.error is not None for r in reqs),
                "wall_s": wall, "throughput_rps": len(reqs) / wall if wall > 0 else None,
                "p50_ms": float(np.percentile(lat, 50) * 1e3) if lat.size else None,
                "p99_ms": float(np.percentile(lat, 99) * 1e3) if lat.size else None,
                "by_kind": by_kind}
