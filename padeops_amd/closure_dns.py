"""Learned SGS closure on filtered-DNS data that this library's solver
generates, and its judgement against the dynamic Smagorinsky model.

1. DNS: forced isotropic turbulence on the N^3 grid of side 2 pi with the
   igrid time advancement on the device (FP64, SSP-RK45, HIT shell forcing,
   viscous term).  After a spin-up, snapshots of u, v (cells) and w (edges).
2. LES filter: the sharp spectral cutoff of the m^3 LES grid, the box
   |k_d| <= kc = floor(m/3) that its 2/3 dealiasing keeps, applied to u_i and
   to the products u_i u_j (formed on the DNS cells, w moved there
   spectrally).  The filtered fields are band-limited below m/2, so their
   samples on the LES grid (cells for u, v, tau; edges for w) hold them
   exactly: tau_ij = bar(u_i u_j) - bar(u_i) bar(u_j), deviatoric.
   (Spectral plumbing with torch.fft; the network, features, stress and
   solver are this library's kernels.)
3. Features: the filtered field set into an m^3 igrid object (its operators
   give duidxjC, as in the LES), then pdo_closure_features: X (10 invariant
   features), a = A / g (A9); targets T = tau / (Delta^2 g^2), Delta = pi / kc.
4. Training: Adam on batches of 65536 points, the network's BF16 MFMA
   forward / backward kernels, loss and dL/dG from pdo_closure_loss.
5. A priori: held-out snapshots, the network's stress against the exact one
   and against the dynamic Smagorinsky stress (the standard dynamic
   procedure, pdo_dynproc, on the LES field): correlation, relative error,
   mean dissipation.
6. A posteriori: m^3 LES (FP32 + refinement) from a filtered snapshot, with
   no model, dynamic Smagorinsky, and the learned closure with the dynamic
   fallback where its dissipation is negative; time-averaged energy spectra
   against the filtered DNS spectrum.
"""
import math

import numpy as np

import padeops_amd as P

L = 2 * math.pi
IDX6 = ((0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2))


def _torch():
    import torch
    return torch


def kc_of(m):
    """largest wavenumber the m-grid's 2/3 dealiasing keeps"""
    return int(math.floor((2.0 / 3.0) * (m / 2) - 1e-9))


def delta_of(m):
    return math.pi / kc_of(m)


# ---------------------------------------------------------------------------
# 1. DNS
# ---------------------------------------------------------------------------
class DNS:
    """Forced isotropic DNS on the device (FP64 igrid)."""

    def __init__(self, n, Re, forcing, seed=1, u0=0.5, k0=3.0, CFL=1.0, device="cuda"):
        torch = _torch()
        from padeops_amd.synthetic import turbulence
        self.n = n
        self.g = P.igrid()
        assert self.g.init(n, n, n, L, L, L, Re=Re, scheme=2, CFL=CFL, forcing=forcing, prec="f64") == 0
        u, v, w = (u0 * c for c in turbulence(n, k0, seed, torch.float64, torch.device(device)))
        w3 = w.view(n, n, n)
        wE = torch.cat([w3, w3[:1]], 0).reshape(-1)
        assert self.g.set_velocity(u, v, wE) == 0

    def step(self, k):
        assert self.g.step(k) == 0

    def fields(self):
        """physical u, v (n, n, n) and w on the n+1 edges, torch (z, y, x) views"""
        n = self.n
        return (self.g.field("u").view(n, n, n), self.g.field("v").view(n, n, n),
                self.g.field("w").view(n + 1, n, n))

    def time(self):
        return self.g.time()


# ---------------------------------------------------------------------------
# 2. LES filter
# ---------------------------------------------------------------------------
def _keep(n, kc, device):
    torch = _torch()
    return torch.cat([torch.arange(0, kc + 1, device=device), torch.arange(n - kc, n, device=device)])


def _lowpass_sample(F, N, m, kc, zshift):
    """the field of spectrum F (torch fftn of an N^3 (z, y, x) field), cut to
    |k_d| <= kc and sampled on the m^3 grid whose z origin lies zshift DNS
    points above the DNS grid's"""
    torch = _torch()
    dev = F.device
    i = _keep(N, kc, dev)
    sub = F.index_select(0, i).index_select(1, i).index_select(2, i)
    kz = torch.fft.fftfreq(N, d=1.0 / N, device=dev, dtype=torch.float64).index_select(0, i)
    sub = sub * torch.exp(1j * (2 * math.pi / N) * kz * zshift).view(-1, 1, 1)
    j = _keep(m, kc, dev)
    G = torch.zeros((m, m, m), dtype=F.dtype, device=dev)
    G[j.view(-1, 1, 1), j.view(1, -1, 1), j.view(1, 1, -1)] = sub
    return torch.fft.ifftn(G).real * (m ** 3 / N ** 3)


def filter_snapshot(u, v, wE, grids):
    """Filtered fields and exact SGS stress on each m^3 LES grid of `grids`
    (an int or a sequence) from DNS u, v (N^3 cells) and w (N+1 edges): per
    grid (ub, vb on cells, wb on the m+1 edges, tau (6, m, m, m) deviatoric
    11 12 13 22 23 33 on cells), torch (z, y, x).  The DNS transforms are
    taken once for all grids."""
    torch = _torch()
    single = isinstance(grids, int)
    grids = (grids,) if single else tuple(grids)
    N = u.shape[0]
    for m in grids:
        assert (N // m) * m == N and kc_of(m) < m // 2
    kz = torch.fft.fftfreq(N, d=1.0 / N, device=u.device, dtype=torch.float64).view(-1, 1, 1)
    Fw = torch.fft.fftn(wE[:N])
    # w on the DNS cells (half a point up), for the products
    wc = torch.fft.ifftn(Fw * torch.exp(1j * (2 * math.pi / N) * kz * 0.5)).real
    out = {m: {} for m in grids}

    def cut(F, name, shift_cells):
        # shift_cells(r): the LES grid's z origin above the DNS grid's, in
        # DNS points, for r = N / m
        for m in grids:
            out[m][name] = _lowpass_sample(F, N, m, kc_of(m), shift_cells(N // m))

    cell = lambda r: 0.5 * (r - 1)  # noqa: E731  LES cell centres above DNS cell centres
    cut(Fw, "wbc", lambda r: 0.5 * r)
    cut(Fw, "wbe", lambda r: 0.0)
    del Fw
    vel = (u, v, wc)
    for q, c in enumerate((u, v)):
        cut(torch.fft.fftn(c), "u%d" % q, cell)
    for q, (i, j) in enumerate(IDX6):
        cut(torch.fft.fftn(vel[i] * vel[j]), "p%d" % q, cell)
    res = []
    for m in grids:
        o = out[m]
        velb = (o["u0"], o["u1"], o["wbc"])
        tau = torch.stack([o["p%d" % q] - velb[i] * velb[j] for q, (i, j) in enumerate(IDX6)])
        trace = (tau[0] + tau[3] + tau[5]) / 3.0
        for q in (0, 3, 5):
            tau[q] -= trace
        wb = torch.cat([o["wbe"], o["wbe"][:1]], 0)
        res.append((o["u0"], o["u1"], wb, tau))
    return res[0] if single else res


def spectrum(u, v, w):
    """shell-summed energy spectrum E(k), k = 0 .. m/2, of cell fields (m^3)"""
    torch = _torch()
    m = u.shape[0]
    k = torch.fft.fftfreq(m, d=1.0 / m, device=u.device, dtype=torch.float64)
    K = torch.sqrt(k.view(-1, 1, 1) ** 2 + k.view(1, -1, 1) ** 2 + k.view(1, 1, -1) ** 2)
    sh = torch.round(K).long().clamp(max=m // 2)
    E = torch.zeros(m // 2 + 1, dtype=torch.float64, device=u.device)
    for c in (u, v, w):
        Fc = torch.fft.fftn(c.double()) / m ** 3
        E.index_add_(0, sh.reshape(-1), 0.5 * (Fc.abs() ** 2).reshape(-1))
    return E.cpu().numpy()


# ---------------------------------------------------------------------------
# 3. features of the filtered field on its LES grid
# ---------------------------------------------------------------------------
class LesFeatures:
    """An m^3 FP64 igrid used as the operator set of the LES grid: the filtered
    velocity goes in, duidxjC and the closure's features come out."""

    def __init__(self, m, closure, d0, deltaRat=2.0):
        self.m = m
        self.g = P.igrid()
        assert self.g.init(m, m, m, L, L, L, Re=1.0, scheme=2, CFL=0.5, prec="f64") == 0
        self.cl = closure
        self.delta = delta_of(m)
        self.dratio = self.delta / d0
        # the dynamic procedure of the comparison: its test filter acts in x and y
        # only, through a spectral object of this grid without the periodic-z
        # masks (pdo_dynproc refuses those, as InitTestFilter does)
        dx = L / m
        self.sp = P.spectral()
        self.sp.init(m, m, m, dx, dx, dx, fixOddball=False, dealiasF=2.0 / 3.0, init_periodicInZ=False)
        self.dyn = P.DynamicProcedure()
        assert self.dyn.init(self.sp, deltaRat) == 0

    def set(self, ub, vb, wb):
        assert self.g.set_velocity(ub.reshape(-1), vb.reshape(-1), wb.reshape(-1)) == 0
        return self.g.field("duidxjC")

    def dataset(self, ub, vb, wb, tau):
        """(X (n, 10), A9 (n, 9), T (n, 6), g) of one filtered snapshot"""
        torch = _torch()
        n = self.m ** 3
        A = self.set(ub, vb, wb)
        X = torch.empty((n, 10), dtype=torch.float32, device=A.device)
        A9 = torch.empty((n, 9), dtype=torch.float32, device=A.device)
        assert self.cl.features(A, n, self.dratio, X, A9) == 0
        g = self.cl.scale()
        T = (tau.reshape(6, n).t() / (self.delta ** 2 * g * g)).float().contiguous()
        return X, A9, T, g


# ---------------------------------------------------------------------------
# 4. training
# ---------------------------------------------------------------------------
# offsets of the output layer in the network's parameter vector (W4 16 x 64, b4 16)
_W4, _B4 = 9408, 10432


def train(net, closure, X, A9, T, steps, batch=65536, lr=1e-3, seed=0, log=None):
    """Adam on random batches of targets scaled to unit rms (sigma: the
    weighted rms of T); the output layer is multiplied by sigma at the end, so
    the trained network gives the unscaled coefficients.  Returns the history
    of the scaled loss (every 50 steps: the fraction of the targets' mean
    square left unexplained) and sigma."""
    torch = _torch()
    gen = torch.Generator(device=X.device)
    gen.manual_seed(seed)
    N = X.shape[0]
    w = torch.tensor([1, 2, 2, 1, 2, 1], dtype=torch.float64, device=X.device)
    sigma = float(torch.sqrt((T.double() ** 2 * w).sum() / (9.0 * N)))
    G = torch.empty((batch, 6), dtype=torch.float32, device=X.device)
    dG = torch.empty_like(G)
    hist = []
    for it in range(steps):
        idx = torch.randint(0, N, (batch,), generator=gen, device=X.device)
        xb, ab = X[idx].contiguous(), A9[idx].contiguous()
        tb = (T[idx] / sigma).contiguous()
        assert net.forward_train(xb, G) == 0
        sync = it % 50 == 0 or it == steps - 1
        lv = closure.loss(G, ab, tb, dG, sync=sync)
        assert net.backward(dG) == 0
        # cosine decay of the step size
        lr_t = lr * 0.5 * (1 + math.cos(math.pi * it / steps))
        assert net.adam_step(lr=lr_t) == 0
        if sync:
            hist.append((it, lv))
            if log:
                log("train step %d scaled loss %.5f" % (it, lv))
    p = net.get_params()
    p[_W4:_B4 + 16] *= sigma
    net.set_params(p)
    return hist, sigma


# ---------------------------------------------------------------------------
# 5. a priori judgement
# ---------------------------------------------------------------------------
def stress_scores(pred, tau):
    """(mean correlation over the six components, relative L2 error of the
    deviatoric tensor with off-diagonals counted twice); (6, n) tensors"""
    torch = _torch()
    p, t = pred.double(), tau.double()
    cors = []
    for q in range(6):
        a, b = p[q] - p[q].mean(), t[q] - t[q].mean()
        cors.append(float((a * b).sum() / torch.sqrt((a * a).sum() * (b * b).sum())))
    w = torch.tensor([1, 2, 2, 1, 2, 1], dtype=p.dtype, device=p.device).view(6, 1)
    err = float(torch.sqrt((w * (p - t) ** 2).sum() / (w * t * t).sum()))
    return float(np.mean(cors)), err


def dissipation(tau, A):
    """mean of -tau_ij S_ij over the field (tau (6, n), A duidxj planes)"""
    n = tau.shape[1]
    g = A.reshape(9, n).double()
    S = [0.5 * (g[3 * i + j] + g[3 * j + i]) for i, j in IDX6]
    w = (1, 2, 2, 1, 2, 1)
    return float(-sum(w[q] * (tau[q].double() * S[q]).mean() for q in range(6)))


def dynamic_smagorinsky_stress(les, deltaRat=2.0):
    """tau = -2 lambda(z) |S| S_ij of the standard dynamic procedure
    (pdo_dynproc: test filter ratio deltaRat, planar coefficient) on the field
    currently set in the LesFeatures grid; (6, n), and lambda (m,)"""
    torch = _torch()
    m = les.m
    n = m ** 3
    g = les.g
    A = g.field("duidxjC")
    S = torch.empty(6 * n, dtype=torch.float64, device=A.device)
    D = torch.empty(n, dtype=torch.float64, device=A.device)
    assert P.sgs_kernel(A, S, D, n, prec="f64") == 0
    lam = torch.empty(m, dtype=torch.float64, device=A.device)
    assert les.dyn.compute(g.field("u"), g.field("v"), g.field("wC"), g.field("uhat"), g.field("vhat"),
                           g.field("whatC"), S, D, lam, prec="f64") == 0
    nu = D.view(m, m * m) * lam.view(m, 1)
    S6 = S.view(6, n)
    tau = (-2.0 * nu.reshape(1, n)) * S6
    trace = (tau[0] + tau[3] + tau[5]) / 3.0
    tau = tau.clone()
    for q in (0, 3, 5):
        tau[q] -= trace
    return tau, lam


AMD_C = 1.67   # the AMD model constant of the reference's HIT input (Csgs, SGSModelID = 2)


def amd_constants(m, C=AMD_C):
    """camd_x, camd_y, camd_z of init_amd (AMD.F90:1-12) on the m^3 grid of
    side 2 pi: C dx sqrt(1/12) in every direction (the z direction takes the
    same factor here, periodic z having no wall-normal Poincare constant)"""
    dx = L / m
    c = C * dx * math.sqrt(1.0 / 12.0)
    return (c, c, c)


def amd_stress(les, C=AMD_C):
    """tau = -2 nu_AMD S_ij (AMD kernel with the reference's constants) on the
    field set in the LesFeatures grid; (6, n) deviatoric"""
    torch = _torch()
    n = les.m ** 3
    A = les.g.field("duidxjC")
    S = torch.empty(6 * n, dtype=torch.float64, device=A.device)
    nu = torch.empty(n, dtype=torch.float64, device=A.device)
    assert P.sgs_kernel(A, S, nu, n, model="amd", camd=amd_constants(les.m, C), prec="f64") == 0
    tau = (-2.0 * nu.view(1, n)) * S.view(6, n)
    trace = (tau[0] + tau[3] + tau[5]) / 3.0
    tau = tau.clone()
    for q in (0, 3, 5):
        tau[q] -= trace
    return tau


def apriori(les, net, closure, snaps, deltaRat=2.0):
    """scores of the network and the dynamic Smagorinsky stress on held-out
    filtered snapshots [(ub, vb, wb, tau)]"""
    torch = _torch()
    n = les.m ** 3
    out = {"mlp_corr": [], "mlp_err": [], "dyn_corr": [], "dyn_err": [], "amd_corr": [], "amd_err": [],
           "diss_dns": [], "diss_mlp": [], "diss_dyn": [], "diss_amd": [], "mlp_backscatter_fraction": []}
    for ub, vb, wb, tau in snaps:
        A = les.set(ub, vb, wb)
        t6 = tau.reshape(6, n)
        pred = torch.empty(6 * n, dtype=torch.float64, device=A.device)
        assert closure.stress(A, n, les.delta, les.dratio, pred, which=2, prec="f64") == 0
        pred = pred.view(6, n)
        tdyn, _ = dynamic_smagorinsky_stress(les, deltaRat)
        tamd = amd_stress(les)
        c, e = stress_scores(pred, t6)
        cd, ed = stress_scores(tdyn, t6)
        ca, ea = stress_scores(tamd, t6)
        out["amd_corr"].append(ca)
        out["amd_err"].append(ea)
        out["diss_amd"].append(dissipation(tamd, A))
        out["mlp_corr"].append(c)
        out["mlp_err"].append(e)
        out["dyn_corr"].append(cd)
        out["dyn_err"].append(ed)
        out["diss_dns"].append(dissipation(t6, A))
        out["diss_mlp"].append(dissipation(pred, A))
        out["diss_dyn"].append(dissipation(tdyn, A))
        mark = torch.zeros(6 * n, dtype=torch.float64, device=A.device)
        assert closure.stress(A, n, les.delta, les.dratio, mark, which=2, fallback=True, prec="f64") == 0
        out["mlp_backscatter_fraction"].append(closure.fallback_count() / n)
    return {k: float(np.mean(v)) for k, v in out.items()}


# ---------------------------------------------------------------------------
# 6. a posteriori LES
# ---------------------------------------------------------------------------
def les_run(m, Re, forcing, ub, vb, wb, kind, closure=None, d0=None, T=4.0, sample_every=10, CFL=1.0,
            prec="f32_refine", deltaRat=2.0, log=None):
    """m^3 LES from (ub, vb, wb); kind None (no model), "amd" (the AMD model
    with the reference's constant), "dynamic", "learned+dynamic" or "learned".  Returns the time-averaged spectrum over
    the second half of [0, T], the final time and steps, and whether the run
    stayed finite."""
    torch = _torch()
    g = P.igrid()
    if kind == "amd":
        sgs = dict(model="amd", camd=amd_constants(m), cglobal=1.0)
    else:
        sgs = None if kind is None else dict(model="smagorinsky", cglobal=0.0)
    assert g.init(m, m, m, L, L, L, Re=Re, scheme=2, CFL=CFL, sgs=sgs, forcing=forcing, prec=prec) == 0
    dt = torch.float64 if prec == "f64" else torch.float32
    if kind not in (None, "amd"):
        delta = delta_of(m)
        g.set_closure(kind, closure, delta=delta, dratio=delta / (d0 or delta), deltaRat=deltaRat)
    assert g.set_velocity(ub.reshape(-1).to(dt), vb.reshape(-1).to(dt), wb.reshape(-1).to(dt)) == 0
    specs = []
    t = 0.0
    steps = 0
    ok = True
    while t < T:
        assert g.step(sample_every) == 0
        t, steps, _ = g.time()
        u = g.field("u").view(m, m, m)
        if not bool(torch.isfinite(u).all()):
            ok = False
            break
        if t >= T / 2:
            specs.append(spectrum(u, g.field("v").view(m, m, m), g.field("wC").view(m, m, m)))
        if log and steps % (20 * sample_every) == 0:
            log("les %s m=%d t=%.3f steps=%d" % (kind, m, t, steps))
    E = np.mean(specs, axis=0) if specs else None
    return {"E": E, "t": t, "steps": steps, "finite": ok}


def spectrum_error(E, Eref, kc):
    """mean |log(E / Eref)| over shells 1 .. kc, and over kc/2 .. kc"""
    k = np.arange(1, kc + 1)
    r = np.abs(np.log(np.maximum(E[k], 1e-300) / Eref[k]))
    return float(r.mean()), float(r[len(k) // 2:].mean())


# ---------------------------------------------------------------------------
# the whole experiment
# ---------------------------------------------------------------------------
def init_params(seed=0):
    """He initialisation of the 10-64-64-64-6 network, output layer scaled by
    0.1 (small initial stresses); FP32 [(W (out, in), b)]"""
    rng = np.random.default_rng(seed)
    sizes = (10, 64, 64, 64, 6)
    out = []
    for i in range(4):
        W = rng.standard_normal((sizes[i + 1], sizes[i])) * math.sqrt(2.0 / sizes[i])
        if i == 3:
            W *= 0.1
        out.append((W.astype(np.float32), np.zeros(sizes[i + 1], np.float32)))
    return out


def run(N=256, les_grids=(64, 128), Re=600.0, nsnap=200, every=10, spinup_time=6.0, ntest=40, ntest_keep=8,
        train_steps=3000, sub_points=262144, les_grid=128, les_T=4.0, forcing=None, seed=1, log=print):
    """DNS -> filtered data -> training -> a priori and a posteriori
    judgement; returns a dict of results (JSON-ready).  The last ntest
    snapshots are held out (ntest_keep of them, evenly spaced, are scored a
    priori); grids finer than the coarsest contribute sub_points random
    points per snapshot to the training set."""
    import time
    torch = _torch()
    forcing = forcing or dict(kmin=2.0, kmax=3.0, nwaves=20, eps=0.1, seed=seed)
    t0 = time.time()
    clock = lambda: "%.0f s" % (time.time() - t0)  # noqa: E731
    res = {"N": N, "les_grids": list(les_grids), "Re": Re, "forcing": forcing, "nsnap": nsnap, "every": every,
           "ntest": ntest}
    dns = DNS(N, Re, forcing, seed=seed)
    while dns.time()[0] < spinup_time:
        dns.step(20)
        t, st, dt = dns.time()
        log("dns spin-up t=%.3f steps=%d dt=%.4g (%s)" % (t, st, dt, clock()))
    res["spinup"] = {"t": dns.time()[0], "steps": dns.time()[1]}
    m0 = min(les_grids)
    d0 = delta_of(m0)
    net = P.SGSNet(65536)
    net.set_params(P.SGSNet.pack(init_params(seed)))
    mmax = max(max(les_grids), les_grid)
    closure = P.Closure(net, mmax * mmax * (mmax + 1))
    les = {m: LesFeatures(m, closure, d0) for m in les_grids}
    Xs, As, Ts = [], [], []
    test = {m: [] for m in les_grids}
    Eacc = {m: 0.0 for m in les_grids}
    keep_at = set(int(round(nsnap - ntest + i * ntest / ntest_keep)) for i in range(ntest_keep))
    gen = torch.Generator(device="cuda")
    gen.manual_seed(seed)
    urms = []
    for snap in range(nsnap):
        dns.step(every)
        u, v, w = dns.fields()
        urms.append(float(torch.sqrt((u * u + v * v + w[:N] * w[:N]).mean() / 3)))
        held = snap >= nsnap - ntest
        for m, (ub, vb, wb, tau) in zip(les_grids, filter_snapshot(u, v, w, les_grids)):
            Eacc[m] = Eacc[m] + spectrum(ub, vb, wb[:m])
            if held:
                if snap in keep_at:
                    test[m].append((ub, vb, wb, tau))
                continue
            X, A9, T, _ = les[m].dataset(ub, vb, wb, tau)
            if m != m0 and X.shape[0] > sub_points:
                idx = torch.randint(0, X.shape[0], (sub_points,), generator=gen, device=X.device)
                X, A9, T = X[idx], A9[idx], T[idx]
            Xs.append(X)
            As.append(A9)
            Ts.append(T)
        if snap % 10 == 9:
            t, st, dt = dns.time()
            log("snapshot %d/%d t=%.3f steps=%d u_rms=%.4f (%s)" % (snap + 1, nsnap, t, st, urms[-1], clock()))
    res["dns_u_rms_mean"] = float(np.mean(urms))
    res["dns_end"] = {"t": dns.time()[0], "steps": dns.time()[1]}
    X, A9, T = torch.cat(Xs), torch.cat(As), torch.cat(Ts)
    del Xs, As, Ts
    res["train_points"] = int(X.shape[0])
    log("training on %d points (%s)" % (X.shape[0], clock()))
    hist, sigma = train(net, closure, X, A9, T, train_steps, log=log)
    res["train_loss_scaled"] = hist
    res["target_rms"] = sigma
    del X, A9, T
    res["apriori"] = {}
    for m in les_grids:
        res["apriori"][str(m)] = apriori(les[m], net, closure, test[m])
        log("a priori m=%d %s (%s)" % (m, res["apriori"][str(m)], clock()))
    # a posteriori: the LES from the first held-out snapshot, filtered DNS
    # spectrum (all snapshots) as the reference
    m = les_grid
    Eref = Eacc[m] / nsnap
    ub, vb, wb, _ = test[m][0]
    post = {"Eref": Eref.tolist(), "kc": kc_of(m)}
    for kind in (None, "amd", "dynamic", "learned+dynamic", "learned"):
        r = les_run(m, Re, forcing, ub, vb, wb, kind, closure=closure, d0=d0, T=les_T, log=log)
        name = kind or "no_model"
        entry = {"t": r["t"], "steps": r["steps"], "finite": r["finite"]}
        if r["E"] is not None:
            entry["E"] = r["E"].tolist()
            entry["spectrum_err_all"], entry["spectrum_err_high"] = spectrum_error(r["E"], Eref, kc_of(m))
        post[name] = entry
        log("a posteriori %s %s (%s)" % (name, {k: v for k, v in entry.items() if k != "E"}, clock()))
    res["aposteriori"] = post
    res["params"] = net.get_params().tolist()
    res["time_s"] = time.time() - t0
    return res


def main(argv=None):
    import argparse
    import json
    import os
    ap = argparse.ArgumentParser(description="learned SGS closure on filtered DNS (C4 workload)")
    ap.add_argument("--N", type=int, default=256)
    ap.add_argument("--les-grids", type=int, nargs="+", default=[64, 128])
    ap.add_argument("--les-grid", type=int, default=128)
    ap.add_argument("--Re", type=float, default=600.0)
    ap.add_argument("--nsnap", type=int, default=200)
    ap.add_argument("--every", type=int, default=10)
    ap.add_argument("--spinup", type=float, default=6.0)
    ap.add_argument("--ntest", type=int, default=40)
    ap.add_argument("--train-steps", type=int, default=3000)
    ap.add_argument("--les-T", type=float, default=4.0)
    ap.add_argument("--out", default="closure_c4.json")
    ap.add_argument("--params-out", default=None, help="trained parameters (.npy, FP32)")
    a = ap.parse_args(argv)

    def log(msg):
        print(msg, flush=True)

    res = run(N=a.N, les_grids=tuple(a.les_grids), Re=a.Re, nsnap=a.nsnap, every=a.every, spinup_time=a.spinup,
              ntest=a.ntest, train_steps=a.train_steps, les_grid=a.les_grid, les_T=a.les_T, log=log)
    params = np.asarray(res.pop("params"), np.float32)
    if a.params_out:
        np.save(a.params_out, params)
    os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
    with open(a.out, "w") as fh:
        json.dump(res, fh, indent=1)
    log("wrote %s" % a.out)


if __name__ == "__main__":
    main()
