"""Learned closure for coarse lattices: data, training and evaluation (host side).

A coarse D3Q19 lattice (4x coarser than a reference run) gets a per-cell
relaxation-time correction from a small network of its own non-equilibrium
moments (the network and its inputs: ``aslhip_closure_set`` in
``include/aslhip.h``).  It is trained on box-filtered runs of a 4x finer
lattice produced by this same library:

* the fine run starts from a seeded random-phase solenoidal field with
  energy spectrum E(k) ~ k^4 exp(-2 (k/4)^2) scaled to u_rms;
* at each snapshot the fine velocity is box-filtered by 4; the subgrid
  stress tau_ij = <u_i u_j> - <u_i><u_j> is projected on the filtered strain
  rate S_ij (central differences) to an eddy viscosity
  nu_t = -<tau^d : S>_3 / (2 <S : S>_3) with <.>_3 a 3^3 box average (the
  least-squares estimate of the dynamic Smagorinsky procedure; negative
  values, backscatter, are kept down to -nu_coarse / 2), and the target
  relaxation-time correction is 3 nu_t;
* the coarse lattice, started from the filtered fields and run two steps,
  gives the network inputs of every cell.

The network trains on the GPU with the library's own forward / backward /
Adam kernels (``asl_amd.Mlp``).  Plain LBGK stays the fallback
(``set_closure(None)``) and the reference for its error.  Acoustic scaling: one
coarse step is four fine steps at the same lattice velocity, and the coarse
viscosity is a quarter of the fine one.
"""
import time

import numpy as np

from .lattice import Lattice, Mlp


def turbulence_field(n, seed, u_rms=0.03, kpeak=4.0):
    """Random-phase solenoidal velocity on a periodic n^3 grid (float32, shape (3, n, n, n)).

    White noise filtered in Fourier space by k exp(-(k/kpeak)^2) (so the shell
    spectrum is E(k) ~ k^4 exp(-2 (k/kpeak)^2)), projected onto divergence-free
    modes, mean removed, scaled to u_rms = sqrt(<|u|^2> / 3).
    """
    rng = np.random.default_rng(seed)
    k1 = np.fft.fftfreq(n, d=1.0 / n).astype(np.float32)
    kz = np.fft.rfftfreq(n, d=1.0 / n).astype(np.float32)
    KX, KY, KZ = np.meshgrid(k1, k1, kz, indexing="ij")
    k2 = KX * KX + KY * KY + KZ * KZ
    kk = np.sqrt(k2)
    amp = kk * np.exp(-(kk / np.float32(kpeak)) ** 2)
    amp[0, 0, 0] = 0.0
    uh = []
    for _ in range(3):
        w = rng.standard_normal((n, n, n), dtype=np.float32)
        uh.append(np.fft.rfftn(w) * amp)
    k2[0, 0, 0] = 1.0
    div = (KX * uh[0] + KY * uh[1] + KZ * uh[2]) / k2
    uh[0] -= KX * div
    uh[1] -= KY * div
    uh[2] -= KZ * div
    del div, KX, KY, KZ, k2, kk, amp
    u = np.stack([np.fft.irfftn(c, s=(n, n, n), axes=(0, 1, 2)).astype(np.float32) for c in uh])
    del uh
    u -= u.mean(axis=(1, 2, 3), keepdims=True)
    u *= np.float32(u_rms / np.sqrt((u * u).sum(0).mean() / 3.0))
    return u


def pressure_field(u):
    """Pressure of an incompressible periodic field: lap p = -d_i d_j (u_i u_j), spectrally
    (float32, mean zero); the lattice density is then rho = 1 + 3 p."""
    n = u.shape[-1]
    k1 = np.fft.fftfreq(n, d=1.0 / n).astype(np.float32) * np.float32(2 * np.pi / n)
    kz = np.fft.rfftfreq(n, d=1.0 / n).astype(np.float32) * np.float32(2 * np.pi / n)
    K = np.meshgrid(k1, k1, kz, indexing="ij")
    k2 = K[0] * K[0] + K[1] * K[1] + K[2] * K[2]
    k2[0, 0, 0] = 1.0
    ph = np.zeros(K[0].shape, dtype=np.complex64)
    for i in range(3):
        for j in range(i, 3):
            w = 1.0 if i == j else 2.0
            ph += np.float32(w) * K[i] * K[j] * np.fft.rfftn(u[i] * u[j], axes=(0, 1, 2))
    ph = -ph / k2
    ph[0, 0, 0] = 0.0
    return np.fft.irfftn(ph, s=(n, n, n), axes=(0, 1, 2)).astype(np.float32)


def populations(rho, u, tau, q=19):
    """ASL-normalised populations f_i / w_i of a periodic field: the equilibrium
    rho + 3 c.m + 4.5 (c.m)^2 / rho - 1.5 m.m / rho (m = rho u) plus the first
    Chapman-Enskog term -3 tau rho (c_a c_b - delta_ab / 3) d_a u_b (central
    differences), so a run starts with the stress of its velocity field."""
    from .lattice import VectorTemplate
    t = VectorTemplate(q)
    rho = np.asarray(rho, dtype=np.float64)
    u = np.asarray(u, dtype=np.float64)
    m = rho * u
    mm = (m * m).sum(0) / rho
    grad = [[(np.roll(u[b], -1, axis=a) - np.roll(u[b], 1, axis=a)) * 0.5 for b in range(3)] for a in range(3)]
    div = grad[0][0] + grad[1][1] + grad[2][2]
    f = np.empty((q,) + rho.shape, dtype=np.float32)
    for i in range(q):
        c = t.c[i]
        cm = c[0] * m[0] + c[1] * m[1] + c[2] * m[2]
        cgc = sum(c[a] * c[b] * grad[a][b] for a in range(3) for b in range(3) if c[a] * c[b])
        neq = -3.0 * tau * rho * (cgc - div / 3.0)
        f[i] = rho + 3.0 * cm + 4.5 * cm * cm / rho - 1.5 * mm + neq
    return f


def with_ghosts(a):
    """Periodic ghost layer around the last three axes (ASL's ghost-inclusive layout)."""
    return np.pad(a, [(0, 0)] * (a.ndim - 3) + [(1, 1)] * 3, mode="wrap")


def box_filter(a, r=4):
    """Mean over r^3 blocks of the last three axes."""
    s = a.shape[-3:]
    b = a.reshape(a.shape[:-3] + (s[0] // r, r, s[1] // r, r, s[2] // r, r))
    nd = a.ndim - 3
    return b.mean(axis=(nd + 1, nd + 3, nd + 5))


def box3(a):
    """Periodic 3^3 box average over the last three axes."""
    out = np.zeros_like(a)
    for dx in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dz in (-1, 0, 1):
                out += np.roll(a, (dx, dy, dz), axis=(-3, -2, -1))
    return out / 27.0


def strain_rate(u):
    """S_ij = (d_i u_j + d_j u_i) / 2 by periodic central differences, as (xx, yy, zz, xy, xz, yz)."""
    d = [[(np.roll(u[j], -1, axis=i) - np.roll(u[j], 1, axis=i)) * 0.5 for j in range(3)] for i in range(3)]
    return np.stack([d[0][0], d[1][1], d[2][2], 0.5 * (d[0][1] + d[1][0]), 0.5 * (d[0][2] + d[2][0]),
                     0.5 * (d[1][2] + d[2][1])])


def eddy_viscosity(u_fine, r=4, floor=-np.inf):
    """Least-squares eddy viscosity of the box-filtered field (coarse lattice units), at least floor."""
    ub = box_filter(u_fine, r)
    pairs = [(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)]
    tau = np.stack([box_filter(u_fine[i] * u_fine[j], r) - ub[i] * ub[j] for i, j in pairs])
    tr = (tau[0] + tau[1] + tau[2]) / 3.0
    tau[:3] -= tr
    S = strain_rate(ub)
    wts = np.array([1.0, 1.0, 1.0, 2.0, 2.0, 2.0]).reshape(6, 1, 1, 1)  # off-diagonal pairs count twice
    num = box3((wts * tau * S).sum(0))
    den = box3((wts * S * S).sum(0))
    return np.maximum(floor, -num / (2.0 * np.maximum(den, 1e-30))), ub


class Experiment:
    """Fine / coarse decaying-turbulence runs sharing one initial field."""

    def __init__(self, n_coarse=128, tau_fine=0.55, u_rms=0.03, seed=20261202, r=4, ce_init=False):
        """ce_init: False (both lattices start from equilibrium), True (both
        from equilibrium + the first Chapman-Enskog term) or "coarse" (only
        the coarse lattice; the fine run resolves its own start)."""
        self.nc, self.r = n_coarse, r
        self.nf = n_coarse * r
        self.nu_f = (tau_fine - 0.5) / 3.0
        self.nu_c = self.nu_f / r
        self.seed, self.u_rms = seed, u_rms
        self.ce_init = ce_init  # start from equilibrium + first Chapman-Enskog term (else equilibrium)
        self.u0 = turbulence_field(self.nf, seed, u_rms)
        # the pressure that goes with the field, so neither lattice starts with an acoustic transient
        self.rho0 = np.float32(1.0) + np.float32(3.0) * pressure_field(self.u0)

    def coarse(self, u=None, rho=None):
        L = Lattice((self.nc,) * 3, 19, dtype="f32", periodic=True, nu=self.nu_c)
        ub = box_filter(self.u0) if u is None else u
        rb = box_filter(self.rho0) if rho is None else rho
        if self.ce_init:
            L.write_f(with_ghosts(populations(rb, ub, 3.0 * self.nu_c + 0.5)))
        else:
            L.init_equilibrium(with_ghosts(rb.astype(np.float32)), with_ghosts(ub.astype(np.float32)))
        return L

    def fine(self):
        L = Lattice((self.nf,) * 3, 19, dtype="f32", periodic=True, nu=self.nu_f)
        if self.ce_init is True:
            L.write_f(with_ghosts(populations(self.rho0, self.u0, 3.0 * self.nu_f + 0.5)))
        else:
            L.init_equilibrium(with_ghosts(self.rho0), with_ghosts(self.u0))
        return L

    def training_data(self, coarse_steps_between, snapshots, log=print):
        """(features n x 7, targets n) from the fine run at `snapshots` coarse-step times."""
        F = self.fine()
        inner = (slice(None),) + (slice(1, -1),) * 3
        xs, ts, done = [], [], 0
        for snap in range(1, snapshots + 1):
            F.step((snap * coarse_steps_between - done) * self.r)
            done = snap * coarse_steps_between
            u = F.velocity()[inner]
            rho = box_filter(F.rho()[inner[1:]])
            nut, ub = eddy_viscosity(u, self.r, floor=-0.5 * self.nu_c)
            del u
            C = self.coarse(ub.astype(np.float32), rho.astype(np.float32))
            C.step(2)
            xs.append(C.closure_features().reshape(-1, 7))
            ts.append((3.0 * nut).astype(np.float32).reshape(-1))
            log("# closure data: snapshot %d (coarse step %d): mean target %.3e" % (snap, done, ts[-1].mean()))
        del F
        return np.concatenate(xs), np.concatenate(ts)

    def ke_history(self, lattice, steps_between, count, fine=False):
        """Mean kinetic energy per cell after each of `count` chunks (coarse-step units)
        and the wall time of the stepping alone."""
        mult = self.r if fine else 1
        ncell = float((self.nf if fine else self.nc) ** 3)
        out = [lattice.conservation()[4] / ncell]
        wall = 0.0
        for _ in range(count):
            lattice.synchronize()
            t0 = time.perf_counter()
            lattice.step(steps_between * mult, write_macro=False, graph=True)
            lattice.synchronize()
            wall += time.perf_counter() - t0
            out.append(lattice.conservation()[4] / ncell)
        return np.array(out), wall

    def filtered_ke_history(self, steps_between, count):
        """Kinetic energy of the box-filtered fine run, per coarse cell: mean |<rho u>|^2 / (2 <rho>)."""
        F = self.fine()
        inner = (slice(None),) + (slice(1, -1),) * 3
        out = []
        for c in range(count + 1):
            if c:
                F.step(steps_between * self.r, write_macro=True, graph=True)
            rho = F.rho()[inner[1:]].astype(np.float64)
            m = box_filter(F.velocity()[inner].astype(np.float64) * rho, self.r)
            rb = box_filter(rho, self.r)
            out.append(float(((m * m).sum(0) / (2.0 * rb)).mean()))
        return np.array(out)


def init_params(seed=0):
    """Random initial network parameters (1345 floats): uniform He-style
    weights (bound sqrt(6 / fan_in), sqrt(3 / fan_in) for the output unit),
    zero biases; layout W1 (32x7), b1, W2 (32x32), b2, w3, b3."""
    rng = np.random.default_rng(seed)
    W1 = rng.uniform(-1, 1, (32, 7)) * np.sqrt(6.0 / 7)
    W2 = rng.uniform(-1, 1, (32, 32)) * np.sqrt(6.0 / 32)
    w3 = rng.uniform(-1, 1, 32) * np.sqrt(3.0 / 32)
    return np.concatenate([W1.ravel(), np.zeros(32), W2.ravel(), np.zeros(32), w3, [0.0]]).astype(np.float32)


def normalisation(x, t):
    """Input shift / scale and output scale of a training set."""
    mu = x.mean(0).astype(np.float32)
    sd = x.std(0)
    sc = (1.0 / np.where(sd > 0, sd, 1.0)).astype(np.float32)
    out = float(t.std()) if t.std() > 0 else 1.0
    return mu, sc, out


def train(x, t, params0, steps=3000, batch=8192, lr=2e-3, seed=1):
    """Train on the GPU; returns (params, mu, sc, out_scale, losses)."""
    mu, sc, out = normalisation(x, t)
    xn = ((x - mu) * sc).astype(np.float32)
    m = Mlp(params0)
    m.set_data(xn, (t / out).astype(np.float32))
    losses = m.train(steps, batch, lr, seed)
    return m.params(), mu, sc, out, losses


# output scales tried on the training field (0: plain LBGK)
SCALE_GRID = (0.0, 0.1, 0.15, 0.2, 0.25, 0.3, 0.35, 0.5, 0.75, 1.0, 1.5, 2.0)


def _ke_decay_error(ke, ke_ref):
    """Largest relative deviation of the kinetic-energy decay ke / ke[0] from
    the reference's (inf if the run blew up)."""
    if not np.isfinite(ke).all():
        return float("inf")
    return float(np.abs((ke / ke[0]) / (ke_ref / ke_ref[0]) - 1.0).max())


def _ke_error(ke, ke_ref):
    if not np.isfinite(ke).all():
        return float("inf")
    return float(np.abs(ke / ke_ref - 1.0).max())


def _shifted(params, shift):
    """The network with its output bias b3 moved by shift (a uniform change of
    y, i.e. of the relaxation-time correction by shift x output scale)."""
    p = np.array(params, dtype=np.float32, copy=True)
    p[-1] += np.float32(shift)
    return p


def study(n_coarse, tau_fine, chunk=50, snapshots=16, eval_chunks=16, target="total", seed=20261202,
          train_steps=4000, batch=16384, grid=SCALE_GRID, shifts=(0.0,), ce_init=False, log=print):
    """Coarse n^3 with and without the learned closure against the 4x finer run.

    The network is trained, and its output scale chosen from `grid` (with
    its output bias moved by one of `shifts`), on one
    realisation of the initial field (seed + 1): against the fine run's
    kinetic-energy decay (target "total") or the box-filtered fine kinetic
    energy ("filtered"), over the evaluation horizon of eval_chunks chunks of
    `chunk` coarse steps.  The comparison runs on another realisation (seed).
    ce_init: the lattices' start (Experiment).  Returns a dict of the
    histories, errors and wall times.
    """
    t0 = time.perf_counter()
    tr = Experiment(n_coarse=n_coarse, tau_fine=tau_fine, seed=seed + 1, ce_init=ce_init)
    log("# closure study: coarse %d^3, fine %d^3, tau_fine %.4f; training field ready (%.1f s)"
        % (n_coarse, tr.nf, tau_fine, time.perf_counter() - t0))
    x, t = tr.training_data(chunk, snapshots, log=log)
    tt = time.perf_counter()
    params, mu, sc, out, losses = train(x, t, init_params(1), steps=train_steps, batch=batch, lr=2e-3)
    train_s = time.perf_counter() - tt
    log("# trained on %d samples in %.1f s: loss %.4f -> %.4f" % (t.size, train_s, losses[:20].mean(),
                                                                 losses[-20:].mean()))
    del x, t
    if target == "total":
        F = tr.fine()
        ke_tr, _ = tr.ke_history(F, chunk, eval_chunks, fine=True)
        del F
    else:
        ke_tr = tr.filtered_ke_history(chunk, eval_chunks)
    calib = {}
    for a in grid:
        for sh in (shifts if a > 0 else (0.0,)):
            C = tr.coarse()
            if a > 0:
                C.set_closure(_shifted(params, sh), mu, sc, out * a)
            ke_a, _ = tr.ke_history(C, chunk, eval_chunks)
            calib[(a, sh)] = _ke_decay_error(ke_a, ke_tr) if target == "total" else _ke_error(ke_a, ke_tr)
            del C
    alpha, shift = min(calib, key=calib.get)
    log("# output scale, bias shift on the training field (%s kinetic energy): %s -> %s, %s"
        % (target, calib, alpha, shift))
    del tr
    ex = Experiment(n_coarse=n_coarse, tau_fine=tau_fine, seed=seed, ce_init=ce_init)
    F = ex.fine()
    ke_f, wall_f = ex.ke_history(F, chunk, eval_chunks, fine=True)
    del F
    log("# fine run done (%.1f s)" % wall_f)
    ke_ff = ex.filtered_ke_history(chunk, eval_chunks)
    log("# filtered fine kinetic energy done")
    P = ex.coarse()
    ke_p, wall_p = ex.ke_history(P, chunk, eval_chunks)
    del P
    C = ex.coarse()
    C.set_closure(_shifted(params, shift), mu, sc, out * max(alpha, 1e-30))
    ke_c, wall_c = ex.ke_history(C, chunk, eval_chunks)
    del C
    return {"study": "learned closure, D3Q19 decaying turbulence", "coarse": n_coarse, "fine": ex.nf,
            "tau_fine": tau_fine, "scale_chosen_against": target, "tau_coarse": 0.5 + 3 * ex.nu_c,
            "start": {False: "equilibrium", True: "equilibrium + Chapman-Enskog term",
                      "coarse": "coarse: equilibrium + Chapman-Enskog term; fine: equilibrium"}[ce_init],
            "coarse_steps_per_chunk": chunk, "train_snapshots": snapshots, "eval_chunks": eval_chunks,
            "output_scale_factor": alpha, "output_bias_shift": shift,
            "calibration_errors": {"%s/%s" % k: v for k, v in calib.items()},
            "ke_fine": ke_f.tolist(), "ke_fine_filtered": ke_ff.tolist(), "ke_coarse_plain": ke_p.tolist(),
            "ke_coarse_closure": ke_c.tolist(),
            "max_rel_ke_decay_error_plain": _ke_decay_error(ke_p, ke_f),
            "max_rel_ke_decay_error_closure": _ke_decay_error(ke_c, ke_f),
            "max_rel_ke_error_vs_filtered_plain": _ke_error(ke_p, ke_ff),
            "max_rel_ke_error_vs_filtered_closure": _ke_error(ke_c, ke_ff),
            "wall_s_fine": wall_f, "wall_s_coarse_plain": wall_p, "wall_s_coarse_closure": wall_c,
            "speedup_closure_vs_fine": wall_f / wall_c, "train_s": train_s,
            "train_loss_first": float(losses[:20].mean()), "train_loss_last": float(losses[-20:].mean())}


__all__ = ["Experiment", "turbulence_field", "box_filter", "eddy_viscosity", "strain_rate", "train", "init_params",
           "normalisation", "with_ghosts", "study", "SCALE_GRID"]
