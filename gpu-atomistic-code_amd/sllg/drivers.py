"""Host-side drivers around the time step: the M(T) driver of the reference.

`mvt` restates sim::MvT (src/mvt.cpp:21-392): for each temperature, run MinRunTime without
convergence checks, then up to MaxRunTime more, pushing |m_s| of every magnetic species
into a running mean/variance (util::RunningStat, inc/util.h:26-80) on every step t with
t % update == 0; a temperature has converged when for every species the running mean moved
by less than Mean_tol since the previous sample and the running variance is below
Variance_tol.  Each temperature adds one line "T<TAB>|m_0|<TAB>|m_1|..." to the output file
(the last sample's |m_s|, as the reference writes modm), after the header
"#Temperature<TAB>Mean".  The temperature loop is the reference's: T = lower, lower + step,
... while T < upper for step > 0, and T = upper, upper + step, ... while T > lower for
step < 0.

Differences, stated: the step number handed to the time step is the global step count, so
every temperature draws fresh counter noise (the reference's local t restarts at 0 per
temperature; its cuRAND stream simply continues).  The per-sample log files
(util::output_mag, VTU, real-space correlation) are not written here.
"""
import math

import numpy as np


class RunningStat:
    """Welford running mean and (n - 1)-normalised variance, as util::RunningStat."""

    def __init__(self):
        self.n = 0
        self.m = 0.0
        self.s = 0.0

    def clear(self):
        self.n = 0

    def push(self, x):
        self.n += 1
        if self.n == 1:
            self.m = x
            self.s = 0.0
        else:
            m_new = self.m + (x - self.m) / self.n
            self.s = self.s + (x - self.m) * (x - m_new)
            self.m = m_new

    def mean(self):
        return self.m if self.n > 0 else 0.0

    def variance(self):
        return self.s / (self.n - 1) if self.n > 1 else 0.0


def _format_g(x):
    """std::ostream's default formatting of a double (%g, 6 significant digits)."""
    return "%g" % x


def temperatures(lower, upper, step):
    """The reference's temperature loop (floating accumulation of the step)."""
    out = []
    if step > 0:
        T = lower
        while T < upper:
            out.append(T)
            T += step
    elif step < 0:
        T = upper
        while T > lower:
            out.append(T)
            T += step
    return out


def mvt(ctx, lower, upper, step, dt, update, min_run_time=5e-12, max_run_time=25e-12, mean_tol=1e-5,
        var_tol=1e-6, nspecies=None, path="MvT.dat", t_global=0):
    """Run the M(T) driver on a handle (`sllg.Context` or anything with set_temperature,
    step(t0, n) and reduce(energy=False) -> (m[nsub, 3], E)).  Returns a list of dicts
    (T, modm, converged, steps) and writes `path` (None: no file)."""
    eminrts = int(min_run_time / dt)
    mrts = int(max_run_time / dt)
    rows = []
    globt = int(t_global)
    lines = ["#Temperature\tMean"]

    def sample():
        m = np.asarray(ctx.reduce(energy=False)[0])
        ns = nspecies if nspecies is not None else m.shape[0]
        return [math.sqrt(float(m[i, 0]) ** 2 + float(m[i, 1]) ** 2 + float(m[i, 2]) ** 2) for i in range(ns)]

    def run(t_from, t_to, on_sample):
        """integrate(t) for t in [t_from, t_to), calling on_sample() after every t with
        t % update == 0; stops early when on_sample returns True.  Returns the steps run."""
        nonlocal globt
        t = t_from
        while t < t_to:
            ts = t + (-t) % update  # next sampling step at or after t
            last = min(ts, t_to - 1)
            n = last - t + 1
            ctx.step(globt, n)
            globt += n
            t = last + 1
            if last == ts and on_sample():
                return t - t_from
        return t - t_from

    for T in temperatures(lower, upper, step):
        ctx.set_temperature(T)
        if eminrts > 0:  # MinRunTime: no convergence checks (the reference only logs <M> here)
            ctx.step(globt, eminrts)
            globt += eminrts
        stats = None
        oldmean = None
        state = {"modm": None, "conv": False}

        def check():
            nonlocal stats, oldmean
            modm = sample()
            if stats is None:
                stats = [RunningStat() for _ in modm]
                oldmean = [0.0] * len(modm)
            counter = 0
            for i, x in enumerate(modm):
                stats[i].push(x)
                if abs(stats[i].mean() - oldmean[i]) < mean_tol and stats[i].variance() < var_tol:
                    counter += 1
                oldmean[i] = stats[i].mean()
            state["modm"] = modm
            state["conv"] = counter == len(modm)
            return state["conv"]

        steps = eminrts + run(eminrts, eminrts + mrts, check)
        # no sample at all (MaxRunTime shorter than one update interval): the reference
        # writes its zero-initialised modm
        modm = state["modm"] if state["modm"] is not None else [0.0] * len(sample())
        rows.append(dict(T=T, modm=modm, converged=state["conv"], steps=steps))
        lines.append(_format_g(T) + "".join("\t" + _format_g(x) for x in modm))
    if path is not None:
        with open(path, "w") as f:
            f.write("\n".join(lines) + "\n")
    return rows
