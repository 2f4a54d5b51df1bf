"""Warm start of loop 3 by the U-Net (no counterpart in the reference).

Training data come from the plain FP64 solver: random confined-square
channel cases (Ma ~ U[0.1, 2.5], Kn ~ log-U[1e-3, 5e-2], square side ~
U[0.15 H, 0.35 H], centre x ~ U[2 H, 6 H]) are stepped from a uniform start,
and after every time step the pair (state at the start of the step, change of
p and T over the step) is recorded on the device.  The network is trained on
the GPU by its own kernels (sts_unet_train_step), then used by
sts_time_step_ws, which keeps the prediction only when it lowers the first
sweep's residual.
"""
import importlib
import math
import os
import shutil
import tempfile

import numpy as np

PKG = importlib.import_module(__package__)
CASES = importlib.import_module(__package__ + ".cases")
SQ = math.sqrt(15 * math.pi / 128)  # u_in = Re Kn sqrt(15 pi / 128) (S:2022)


def conv_shapes(cin, cout, base):
    """(ci, co, k) of the network's convolutions in parameter order
    (include/sts_unet.h)."""
    B = base
    return [(cin, B, 3), (B, B, 3), (B, 2 * B, 3), (2 * B, 2 * B, 3), (2 * B, 4 * B, 3), (4 * B, 4 * B, 3),
            (4 * B, 8 * B, 3), (8 * B, 8 * B, 3), (12 * B, 4 * B, 3), (4 * B, 4 * B, 3), (6 * B, 2 * B, 3),
            (2 * B, 2 * B, 3), (3 * B, B, 3), (B, B, 3), (B, cout, 1)]


def he_init(cin, cout, base, seed):
    """Flat FP32 parameters: He-normal weights (the linear head: 1/fan-in
    variance) and zero biases."""
    rng = np.random.default_rng(seed)
    parts = []
    shapes = conv_shapes(cin, cout, base)
    for q, (ci, co, k) in enumerate(shapes):
        var = (2.0 if q < len(shapes) - 1 else 1.0) / (k * k * ci)
        parts += [rng.standard_normal(k * k * ci * co) * math.sqrt(var), np.zeros(co)]
    return np.concatenate(parts).astype(np.float32)


def case_params(rng):
    """One seeded draw of the training distribution."""
    Ma = rng.uniform(0.1, 2.5)
    Kn = 10 ** rng.uniform(math.log10(1e-3), math.log10(5e-2))
    return dict(Ma=Ma, Kn=Kn, Re=Ma / (Kn * SQ), side=rng.uniform(0.15, 0.35), xc=rng.uniform(2.0, 6.0))


def write_case(dst, Nx, Ny, prm, steps, dt_over_tin=None):
    """Case directory of a confined square; dt keeps the Courant number of the
    240 x 48 grid at 0.01 H/u_in."""
    dt = dt_over_tin if dt_over_tin is not None else 0.01 * 240 / Nx
    return CASES.confined_square(dst, Nx, Ny, prm["Re"], prm["Kn"], steps, dt_over_tin=dt, side=prm["side"],
                                 xc=prm["xc"])


class Run:
    """A case set up on the GPU the way the program sets up a fresh start
    (uniform field, body values, boundary conditions, fluxes), ready for
    time steps."""

    def __init__(self, case_dir, precision=64):
        self.case = PKG.Case(case_dir)
        c = self.case
        self.ctx = PKG.Context(c.problem(), precision)
        ctx = self.ctx
        Na = c.Nx * c.Ny
        for name, key in (("u", "u_gas_nd"), ("v", "v_gas_nd"), ("p", "p_gas_nd"), ("Temper", "T_gas_nd"),
                          ("rho", "rho_gas_nd")):
            ctx.set(name, np.full(Na, c.value(key)))
        ctx.apply_bodies()
        ctx.save_previous()
        ctx.rho_in_middle_points()
        ctx.apply_bodies()
        ctx.boundary_conditions(PKG.BC_VELOCITIES)
        ctx.save_previous()
        ctx.rho_in_middle_points()
        ctx.properties()
        ctx.rho_in_middle_points()
        ctx.save_previous()
        ctx.rho_in_middle_points()
        ctx.boundary_conditions(PKG.BC_SLIP_U | PKG.BC_SLIP_V | PKG.BC_VELOCITIES)
        ctx.boundary_conditions(PKG.BC_JUMP | PKG.BC_TEMPERATURE | PKG.BC_PRESSURE)
        ctx.fluxes_diffusion()
        self.loop = (int(c.value("N_I")), int(c.value("N_I_p_c")), c.value("MaxError_Velocities"),
                     c.value("MaxError_p_c"), c.value("MaxError_T"))
        self.norm = PKG.UNetNorm(c.value("u_gas_nd"), c.value("p_gas_nd"), c.value("ht") * c.value("u_gas_nd"),
                                 c.value("ht") * c.value("u_gas_nd"))

    def step(self):
        return self.ctx.time_step(*self.loop)[0]

    def step_ws(self, net, mixed=None):
        used, n32, n64, _, _ = self.ctx.time_step_ws(net, self.norm, *self.loop, mixed=mixed)
        return used, n32, n64

    def close(self):
        self.ctx.close()
        self.case.close()


def generate(net, n_cases, steps, seed, log=print):
    """Records up to n_cases * steps samples into the network's training set
    (allocated here); cases whose solve fails are skipped.  Returns the
    number of samples and the case parameters used."""
    rng = np.random.default_rng(seed)
    net.dataset(n_cases * steps)
    work = tempfile.mkdtemp(prefix="sts_ws_")
    slot, used = 0, []
    try:
        for k in range(n_cases):
            prm = case_params(rng)
            d = write_case(os.path.join(work, f"c{k}"), net.W, net.H, prm, steps)
            run = Run(d)
            start = slot
            try:
                for _ in range(steps):
                    run.step()
                    net.record(slot, run.ctx, run.norm)
                    slot += 1
            except PKG.StsError as e:
                log(f"case {k} {prm}: {e}; dropped")
                slot = start
            finally:
                run.close()
            if slot > start:
                used.append(prm)
            log(f"case {k}: Ma {prm['Ma']:.2f} Kn {prm['Kn']:.4f} side {prm['side']:.2f} xc {prm['xc']:.2f}: "
                f"{slot} samples")
    finally:
        shutil.rmtree(work, ignore_errors=True)
    return slot, used


def train(net, n_samples, epochs, batch, lr, seed, holdout=0.1, log=print):
    """Adam on shuffled mini-batches; returns (train losses per epoch,
    held-out loss per epoch)."""
    rng = np.random.default_rng(seed)
    idx = rng.permutation(n_samples)
    n_hold = max(1, int(holdout * n_samples))
    hold, tr = idx[:n_hold], idx[n_hold:]
    hist, hist_hold = [], []
    for ep in range(epochs):
        rng.shuffle(tr)
        losses = [net.train_step(tr[b:b + batch], lr) for b in range(0, len(tr) - batch + 1, batch)]
        hl = [net.eval(hold[b:b + batch]) for b in range(0, len(hold), batch)]
        hist.append(float(np.mean(losses)))
        hist_hold.append(float(np.mean(hl)))
        log(f"epoch {ep}: train loss {hist[-1]:.4e}, held-out {hist_hold[-1]:.4e}")
    return hist, hist_hold


# ---------------------------------------------------------------------------
# The corrector (sts_set_corrector): the same network with cin = 9, cout = 4,
# refining the start predicted by the polynomial extrapolation in time.
# ---------------------------------------------------------------------------
CORR_CIN, CORR_COUT = 9, 4
CORR_FRACTION = 0.05  # the correction's scale as a fraction of the step's change scale
ACCEL = dict(extrapolate=3, guard=64, anderson=1)  # the predicted, accelerated FP64 step the corrector serves


def corr_norm(case):
    """Normalisation of a case: the step's change scale ht u_in (in the
    normalised units of u/u_in, p/p_ref, T) and the correction's scale."""
    u_in, ht = case.value("u_gas_nd"), case.value("ht")
    return PKG.CorrNorm(u_in, case.value("p_gas_nd"), ht * u_in, CORR_FRACTION * ht * u_in)


def accelerated_step(run):
    """One step of the predicted, accelerated FP64 method; returns (FP64
    outer iterations, whether the step ran from a predicted start)."""
    before = run.ctx.predict_stats()[0]
    _, n64, _, _ = run.ctx.time_step_mixed(*run.loop, 0, (1e-7,) * 3, **ACCEL)
    return n64, run.ctx.predict_stats()[0] > before


def generate_corrections(net, n_cases, samples_per_case, seed, max_steps=None, log=print):
    """Records n_cases * samples_per_case corrector samples into the network's
    training set (allocated here): seeded cases of the training distribution
    are stepped with the predicted, accelerated FP64 method, and every step
    that ran from a predicted start gives one sample (x^n, the predicted
    start, the converged x^(n+1)).  A case whose solve fails is dropped and
    the next seeded case takes its place, so the set is always full.
    Returns the parameters of the cases used."""
    rng = np.random.default_rng(seed)
    max_steps = max_steps or samples_per_case + 20
    net.dataset(n_cases * samples_per_case)
    work = tempfile.mkdtemp(prefix="sts_corr_")
    slot, used, k = 0, [], 0
    try:
        while len(used) < n_cases:
            prm = case_params(rng)
            d = write_case(os.path.join(work, f"c{k}"), net.W, net.H, prm, max_steps)
            run = Run(d)
            run.ctx.keep_predicted_start(True)
            nrm = corr_norm(run.case)
            start, ok = slot, False
            try:
                for _ in range(max_steps):
                    _, predicted = accelerated_step(run)
                    if predicted:
                        net.record_correction(slot, run.ctx, nrm)
                        slot += 1
                        if slot - start == samples_per_case:
                            ok = True
                            break
            except PKG.StsError as e:
                log(f"case {k} {prm}: {e}; dropped")
            finally:
                run.close()
            if not ok:
                slot = start
            else:
                used.append(prm)
                log(f"case {k}: Ma {prm['Ma']:.2f} Kn {prm['Kn']:.4f} side {prm['side']:.2f} xc {prm['xc']:.2f}: "
                    f"{slot} samples")
            k += 1
    finally:
        shutil.rmtree(work, ignore_errors=True)
    return used
