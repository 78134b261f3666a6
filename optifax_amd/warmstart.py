"""Surrogate-warm-started OPO steady state.

The cavity is singly resonant (physics.OutputCoupler): only the signal band
is fed back, the pump is injected fresh and the idler is dumped every trip.
The state that carries from one trip to the next is therefore the crystal
input inside the resonated band, and one round trip is

    X_{n+1} = D(tau) . F(X_n + P) + P,    F = band feedback of one crystal pass

with P the pump, D(tau) = exp(-i tau dw) the cavity delay and F the
crystal pass seen through the resonator band, F(X) = cav * pass(X).  The
surrogate learns F on 1024 spectral bins of that band, conditioned on pump
energy and poling period, and replaces it for the first trips of a solve:

    1. surrogate trips: iterate the bin-level map from a weak smooth seed
       through the build-up, while the signal still grows by 2 % per trip
       and stays within the input norms seen in training (at most
       `max_surrogate_trips`); one batched inference per trip;
    2. handover: the bins become a full-grid field (band-limited
       interpolation) plus the exact pump;
    3. gate: the first plain trip's band output is compared with the
       surrogate's prediction for the same input; above `gate_tol` the
       surrogate state is discarded and the instance restarts plainly from its
       noise seed (fallback);
    4. plain FP32 round trips until the trip-to-trip change of the
       output-coupled spectrum is below `conv_tol`, within the trip budget.

The output-coupled spectrum is sqrt(T/R) * F in the band (the coupler's
masks share the band shape), so the convergence test needs no extra pass.

Features (SpectralBins over the band, D = 2 grid points per bin), in the
frame where the input's energy-weighted mean phase is zero (the map commutes
with a global signal phase; outputs are rotated back), with s = ||X||:
  input  [Re X / s, Im X / s (2 x 1024), log10(s / a_ref) / 4,
          pump energy / 4 nJ, (period - 30 um) / 1 um]            (2051)
  output [Re, Im of (F - Lin X) / s]                              (2048)
where Lin = cav exp(-i beta_op NSteps) alpha^NSteps is the crystal's linear
propagation through the band.  The network learns the nonlinear part of the
pass relative to linear propagation, scale-free; the input norm carries the
signal power that sets saturation.  A first representation by amplitude and
unwrapped phase (the phase of a pulse 0.3 ps off centre ramps over ~12 pi
across the signal band) trained to 27 % held-out relative error and made
the warm start fall back on every instance; this one is smooth in the field.
"""
from __future__ import annotations

import math
import time

import numpy as np

from . import physics as ph
from .surrogate import SpectralBins, SurrogateMLP
from .solver import OPOBatch

NBINS = 1024
N_IN = 2 * NBINS + 3
N_OUT = 2 * NBINS
PUMP_REF = 4e-9
PERIOD_REF = 30e-6


# ---------------------------------------------------------------------------
# band, features
# ---------------------------------------------------------------------------
def band_bins(simwin, f_lo=150e12, a_ref=1.0):
    """1024 bins, every second grid frequency from f_lo (150-278 THz on a 16 ps
    window): the resonated band 1.2-1.8 um (167-250 THz) with margin."""
    return SpectralBins(simwin, f_lo, nbins=NBINS, D=2, a_ref=a_ref)


def canonical_phase(Xb):
    """Energy-weighted mean phase of each row of complex bins."""
    Xb = np.atleast_2d(Xb)
    z = np.sum(Xb * np.abs(Xb), axis=-1)
    return np.where(np.abs(z) > 0, np.angle(z), 0.0)


def input_features(Xb, pump_energy, period, a_ref):
    """Rows of complex band bins -> network inputs [B, 2051], the rotation
    applied and the row norms (both needed to decode the output)."""
    Xb = np.atleast_2d(Xb)
    th = canonical_phase(Xb)
    Xr = Xb * np.exp(-1j * th)[:, None]
    s = np.maximum(np.linalg.norm(Xr, axis=1), 1e-300)
    B = Xb.shape[0]
    pe = np.broadcast_to(np.asarray(pump_energy, np.float64) / PUMP_REF, (B,))
    pr = np.broadcast_to((np.asarray(period, np.float64) - PERIOD_REF) / 1e-6, (B,))
    lg = np.log10(s / a_ref) / 4.0
    x = np.concatenate([Xr.real / s[:, None], Xr.imag / s[:, None], lg[:, None], pe[:, None],
                        pr[:, None]], axis=1)
    return x.astype(np.float32), th, s


def output_features(Fb, Xb, lin, th, s):
    """Targets [B, 2048]: the nonlinear part (F - Lin X), rotated and scaled
    like the input."""
    R = (np.atleast_2d(Fb) - lin * np.atleast_2d(Xb)) * np.exp(-1j * np.asarray(th))[:, None]
    R = R / np.asarray(s)[:, None]
    return np.concatenate([R.real, R.imag], axis=1).astype(np.float32)


def from_output_features(y, Xb, lin, th, s):
    y = np.asarray(y, np.float64)
    R = (y[:, :NBINS] + 1j * y[:, NBINS:]) * np.asarray(s)[:, None]
    return lin * np.atleast_2d(Xb) + R * np.exp(1j * np.asarray(th))[:, None]


def linear_band(sb, inst):
    """Lin = cav exp(-i beta_op NSteps) alpha^NSteps on the band bins."""
    si = inst["solver"]
    ns = float(si["nSteps"])
    beta = np.asarray(si["beta_op"], np.float64)
    alpha = np.asarray(si["alpha_w"], np.float64)
    full = np.asarray(inst["cav"]) * np.exp(-1j * beta * ns) * np.where(alpha > 0, alpha ** ns, 0.0)
    return sb.bins(full)


# ---------------------------------------------------------------------------
# instances
# ---------------------------------------------------------------------------
def opo_instance(cfg: ph.OPOConfig):
    """Device instance dict + the pieces the warm start needs."""
    b = cfg.build()
    ops = b["ops"]
    inst = dict(solver=b["solver"], field0=ops["field0"], pump=ops["pump"], cav=ops["cav"],
                ocm=ops["ocm"], delay=b["delay"])
    return inst, b


def smooth_seed(simwin, cfg: ph.OPOConfig, rel_amp=1e-3, fwhm=300e-15, t0=0.0):
    """A weak transform-limited Gaussian at the QPM signal wavelength, fftshifted
    time; rel_amp relative to the pump's peak field."""
    ls = ph.qpm_signal_wavelength(cfg.lambda_ref, cfg.period, cfg.temperature)
    fs = ph.C0 / ls - ph.C0 / cfg.lambda_ref
    t = simwin.times
    pump = ph.sech_pump(simwin, cfg.pump_fwhm, cfg.pump_energy, cfg.area)
    env = np.exp(-2 * math.log(2) * ((t - t0) / fwhm) ** 2)
    return rel_amp * np.abs(pump).max() * env * np.exp(2j * np.pi * fs * t)


class BandCavity:
    """Bin-level cavity of a batch of instances (same SimWindow)."""

    def __init__(self, sb: SpectralBins, insts, builds):
        self.sb = sb
        self.cav = np.stack([sb.bins(i["cav"]) for i in insts])
        self.ocm = np.stack([sb.bins(i["ocm"]) for i in insts])
        self.pump = np.stack([sb.bins(i["pump"]) for i in insts])
        dw = 2 * np.pi * (sb.freqs - ph.C0 / builds[0]["simwin"].reference_wave)
        self.dphase = np.exp(-1j * np.outer([i["delay"] for i in insts], dw))
        self.lin = np.stack([linear_band(sb, i) for i in insts])
        with np.errstate(divide="ignore", invalid="ignore"):
            r = np.abs(self.ocm) / np.abs(self.cav)
        self.oc_over_fb = np.where(np.abs(self.cav) > 0, r, 0.0)

    def next_input(self, Fb):
        return self.dphase * Fb + self.pump

    def band_out(self, Xin_next):
        """F from a next crystal input: exp(+i tau dw) (X_next - P)."""
        return (Xin_next - self.pump) / self.dphase

    def oc_power(self, Fb):
        return (self.oc_over_fb * np.abs(Fb)) ** 2


def conv_metric(S, Sp):
    num = np.sqrt(np.sum((S - Sp) ** 2, axis=-1))
    den = np.sqrt(np.sum(S ** 2, axis=-1))
    return np.where(den > 0, num / np.where(den > 0, den, 1.0), np.inf)


# ---------------------------------------------------------------------------
# training data from device trajectories
# ---------------------------------------------------------------------------
def generate_pairs(cfgs, trips, seeds_rel=(1e-6, 1e-2), rng_seed=0, log=print, ctx=None):
    """Run each configuration from a smooth seed of random strength and phase
    for `trips` plain round trips (one persistent launch per trip, all
    instances together) and return (inputs X [B*trips, 1024] complex bins,
    band outputs F, pump energies, periods) for every trip."""
    rng = np.random.default_rng(rng_seed)
    insts, builds = [], []
    for cfg in cfgs:
        inst, b = opo_instance(cfg)
        sw = b["simwin"]
        amp = math.exp(rng.uniform(math.log(seeds_rel[0]), math.log(seeds_rel[1])))
        seed = smooth_seed(sw, cfg, amp, t0=rng.uniform(-0.3e-12, 0.3e-12)) * \
            np.exp(1j * rng.uniform(0, 2 * np.pi))
        Xk = ph.SimWindow.fft_order(ph.SimWindow.t2k(seed)) + inst["pump"]
        inst["field0"] = np.fft.ifft(Xk)
        insts.append(inst)
        builds.append(b)
    sb = band_bins(builds[0]["simwin"])
    cav = BandCavity(sb, insts, builds)
    run = OPOBatch(insts, trip_cap=trips, ctx=ctx)
    B = len(insts)
    X = np.empty((trips, B, NBINS), np.complex128)
    F = np.empty((trips, B, NBINS), np.complex128)
    cur = np.stack([sb.bins(np.fft.fft(i["field0"])) for i in insts])
    t0 = time.time()
    for t in range(trips):
        run.launch(1)
        run.sync(allow_nonfinite=True)
        nxt = np.stack([sb.bins(np.fft.fft(run.read(i)["field"])) for i in range(B)])
        X[t] = cur
        F[t] = cav.band_out(nxt)
        cur = nxt
        if log and (t % 10 == 9 or t == trips - 1):
            log(f"  data: trip {t + 1}/{trips}, {time.time() - t0:.0f} s")
    run.close()
    pe = np.array([c.pump_energy for c in cfgs])
    pr = np.array([c.period for c in cfgs])
    lin = np.broadcast_to(cav.lin[None], (trips,) + cav.lin.shape).reshape(-1, NBINS)
    return (X.reshape(-1, NBINS), F.reshape(-1, NBINS), np.tile(pe, trips), np.tile(pr, trips),
            lin, sb)


# ---------------------------------------------------------------------------
# training
# ---------------------------------------------------------------------------
def make_dataset(X, F, pe, pr, lin, a_ref=None):
    """Network inputs / targets from complex band pairs; a_ref (the norm
    scale of the log-power input) defaults to the largest input norm."""
    if a_ref is None:
        a_ref = float(np.max(np.linalg.norm(X, axis=1)))
    x, th, s = input_features(X, pe, pr, a_ref)
    y = output_features(F, X, lin, th, s)
    return x, y, a_ref


class BandSurrogate:
    """Trained network + the band it was trained on."""

    def __init__(self, net: SurrogateMLP, a_ref: float, max_norm: float = np.inf):
        self.net = net
        self.a_ref = float(a_ref)
        self.max_norm = float(max_norm)   # largest band-signal norm seen in training

    def __call__(self, Xb, pump_energy, period, lin):
        x, th, s = input_features(Xb, pump_energy, period, self.a_ref)
        y = self.net.forward(x)
        return from_output_features(y, Xb, lin, th, s)


def train_surrogate(x, y, steps, batch=512, lr=1e-3, width=2048, seed=0, val_frac=0.1,
                    log=print):
    """Adam on the device kernels; returns (net, train losses, validation loss,
    device ms).  The last val_frac of the (shuffled) rows is held out."""
    rng = np.random.default_rng(seed)
    perm = rng.permutation(x.shape[0])
    nv = max(64, int(val_frac * x.shape[0]) // 64 * 64)
    vi, ti = perm[:nv], perm[nv:]
    net = SurrogateMLP(N_IN, width, N_OUT, max_batch=batch, seed=seed)
    net.upload_dataset(x[ti], y[ti])
    losses, ms = [], 0.0
    done = 0
    chunk = 500
    while done < steps:
        k = min(chunk, steps - done)
        # cosine decay from lr to lr/20 over the run
        frac = done / max(1, steps)
        lr_k = lr * (0.05 + 0.95 * 0.5 * (1 + math.cos(math.pi * frac)))
        l, t = net.train(k, batch=batch, lr=lr_k, seed=seed + done)
        losses.extend(l.tolist())
        ms += t
        done += k
        if log and (done % 2000 == 0 or done == steps):
            log(f"  train: step {done}/{steps} loss {np.mean(l[-20:]):.4f} lr {lr_k:.2e}")
    yv = net.forward(x[vi])
    num = np.linalg.norm(yv - y[vi], axis=1)
    den = np.linalg.norm(y[vi], axis=1)
    val = float(np.mean(num / np.where(den > 0, den, 1.0)))
    return net, np.array(losses), val, ms


# ---------------------------------------------------------------------------
# solves
# ---------------------------------------------------------------------------
def _run_to_convergence(run, n_inst, conv_tol, budget, chunk, log, tag, t_start, max_seconds):
    """Launch chunks of trips until every instance converged (or diverged),
    the trip budget or max_seconds is reached.  Returns the wall time (from
    t_start) at the end of the chunk in which each instance converged (inf if
    it did not)."""
    done = 0
    conv_at = np.full(n_inst, np.inf)
    while done < budget:
        k = min(chunk, budget - done)
        run.launch(k, conv_tol)
        run.sync(allow_nonfinite=True)
        done += k
        now = time.time() - t_start
        st = np.array([run.read(i)["status"] for i in range(n_inst)])
        newly = (st == 2) & ~np.isfinite(conv_at)
        conv_at[newly] = now
        if log:
            log(f"  {tag}: {done} trips launched, {int((st == 2).sum())}/{n_inst} converged, "
                f"{now:.1f} s")
        if np.all((st == 2) | (st == 3)) or (max_seconds and now > max_seconds):
            break
    return conv_at


def solve_plain(insts, conv_tol=1e-4, budget=1000, chunk=10, log=print, ctx=None,
                max_seconds=None, precision="single"):
    """Plain round trips (FP32 fields unless precision="double") from each
    instance's own initial field until converged (or the budget).  Returns
    dict(seconds, conv_seconds[] (time to steady state per instance, inf if
    not reached), trips[], status[], ...)."""
    t0 = time.time()
    run = OPOBatch(insts, trip_cap=budget, ctx=ctx, precision=precision)
    conv_at = _run_to_convergence(run, len(insts), conv_tol, budget, chunk, log, "plain", t0,
                                  max_seconds)
    rs = [run.read(i) for i in range(len(insts))]
    run.close()
    return dict(seconds=time.time() - t0, conv_seconds=conv_at,
                trips=np.array([r["trips_done"] for r in rs]),
                status=np.array([r["status"] for r in rs]),
                fields=[r["field"] for r in rs], oc=[r["oc_spec"] for r in rs])


def _growth_band(sur, cav, X, pe, pr, max_trips, growth_stop, max_norm):
    """Surrogate trips through the build-up: iterate the bin-level round trip
    while the signal norm still grows by at least `growth_stop` per trip and
    stays within the largest input norm the surrogate was trained on; each
    instance keeps its last state inside those bounds.  Returns (crystal
    inputs X, trips per instance)."""
    B = X.shape[0]
    trips = np.zeros(B, int)
    active = np.ones(B, bool)
    P = cav.pump
    for _ in range(max_trips):
        idx = np.where(active)[0]
        if idx.size == 0:
            break
        F = sur(X[idx], pe[idx], pr[idx], cav.lin[idx])
        Xn = cav.dphase[idx] * F + P[idx]
        n_old = np.linalg.norm(X[idx] - P[idx], axis=1)
        n_new = np.linalg.norm(Xn - P[idx], axis=1)
        ok = np.isfinite(n_new) & (n_new <= max_norm)
        take = idx[ok]
        X[take] = Xn[ok]
        trips[idx] += 1
        stop = ~ok | (n_new < growth_stop * n_old)
        active[idx[stop]] = False
    return X, trips


def solve_warm(cfgs, insts, builds, sur: BandSurrogate, conv_tol=1e-4, budget=1000,
               max_surrogate_trips=100, surrogate_tol=1e-5, gate_tol=0.3, seed_rel=1e-3,
               chunk=10, log=print, ctx=None, max_seconds=None, accel="growth",
               growth_stop=1.02):
    """Surrogate trips on the band bins, handover, gated plain trips.
    Returns dict(seconds, surrogate_seconds, surrogate_trips[], trips[],
    status[], gate_err[], fallback[])."""
    t0 = time.time()
    sw = builds[0]["simwin"]
    sb = band_bins(sw)
    cav = BandCavity(sb, insts, builds)
    B = len(insts)
    pe = np.array([c.pump_energy for c in cfgs])
    pr = np.array([c.period for c in cfgs])
    # smooth seeds in the band (+ pump), bin level
    seeds = [smooth_seed(b["simwin"], c, seed_rel) for c, b in zip(cfgs, builds)]
    X = np.stack([sb.bins(ph.SimWindow.fft_order(ph.SimWindow.t2k(s))) for s in seeds]) + cav.pump
    s_trips = np.zeros(B, int)
    if accel == "growth":
        X, s_trips = _growth_band(sur, cav, X, pe, pr, max_surrogate_trips, growth_stop,
                                  getattr(sur, "max_norm", np.inf))
    else:
        active = np.ones(B, bool)
        S_prev = np.zeros((B, NBINS))
        for t in range(max_surrogate_trips):
            idx = np.where(active)[0]
            if idx.size == 0:
                break
            F = sur(X[idx], pe[idx], pr[idx], cav.lin[idx])
            X[idx] = cav.dphase[idx] * F + cav.pump[idx]
            S = (cav.oc_over_fb[idx] * np.abs(F)) ** 2
            s_trips[idx] += 1
            if t > 0:
                done = conv_metric(S, S_prev[idx]) < surrogate_tol
                active[idx[done]] = False
            S_prev[idx] = S
    t_sur = time.time() - t0
    # handover: full-grid crystal inputs = pump + band-limited signal
    fields = []
    for i in range(B):
        Xk = insts[i]["pump"] + sb.to_grid(X[i] - cav.pump[i])
        fields.append(np.fft.ifft(Xk))
    warm = [dict(inst, field0=f) for inst, f in zip(insts, fields)]
    run = OPOBatch(warm, trip_cap=budget, ctx=ctx)
    # gate: one plain trip vs the surrogate's prediction for the same input
    F_pred = sur(X, pe, pr, cav.lin)
    run.launch(1)
    run.sync(allow_nonfinite=True)
    nxt = np.stack([sb.bins(np.fft.fft(run.read(i)["field"])) for i in range(B)])
    F_plain = cav.band_out(nxt)
    jerr = np.linalg.norm(F_pred - F_plain, axis=1) / np.maximum(np.linalg.norm(F_plain, axis=1),
                                                                 1e-300)
    fallback = jerr > gate_tol
    for i in np.where(fallback)[0]:
        run.write_field(i, insts[i]["field0"])      # restart plainly from the noise seed
    if log:
        log(f"  warm: {t_sur:.2f} s of surrogate trips (max {s_trips.max()}), gate error "
            f"median {np.median(jerr):.3f}, fallbacks {int(fallback.sum())}/{B}")
    conv_at = _run_to_convergence(run, B, conv_tol, budget - 1, chunk, log, "warm", t0,
                                  max_seconds)
    rs = [run.read(i) for i in range(B)]
    run.close()
    return dict(seconds=time.time() - t0, conv_seconds=conv_at, surrogate_seconds=t_sur,
                surrogate_trips=s_trips,
                trips=np.array([r["trips_done"] for r in rs]),
                status=np.array([r["status"] for r in rs]), gate_err=jerr, fallback=fallback,
                fields=[r["field"] for r in rs], oc=[r["oc_spec"] for r in rs])
