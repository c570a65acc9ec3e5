"""Worker functions run inside a fresh process holding the compiled reference
(oracle/_ref), plus the matching oracle-side replay.  Used by
tests/test_oracle_vs_reference.py and tests/golden/make_golden.py."""
import os

import numpy as np

import hc_testlib as T

DUMP = ("u", "up", "um", "v", "vp", "vm", "w", "wp", "wm", "p", "fx", "fy", "fz",
        "t", "s", "c", "q_t", "q_s", "q_c", "kxt", "kyt", "kzt", "rhop", "rrho",
        "ss", "ss_v", "ri", "epsh", "epsv", "chi_v", "chi_h", "axu", "ayu", "axv", "ayv", "azu", "azv",
        "uc", "vc", "wc", "ud", "vd", "wd", "cp", "cc", "worku", "workv", "workw",
        "alphaee", "alphau", "alphav", "alphae", "isource_map", "alphaw", "t_bkg",
        "s_bkg", "c_bkg", "rho_bkg", "azzmin", "pgradx", "pgrady")

PROFILE = os.path.join(T.REF_DIR, "site_stratification_linear.inc")


def ref_init_smooth(ref, seed, u0=0.03, amp=0.01, tnoise=1.0e-4):
    """Reference set-up followed by a smooth crossflow with a few low modes
    and seeded uniform T noise: a state the leapfrog scheme carries stably
    over many steps (used by the whole-step comparisons)."""
    ref.init_profiles(PROFILE)
    ref.call("density")
    ref.call("initialize")
    ref.call("sponge_maker")
    rng = np.random.default_rng(seed)
    nx, ny, nz = ref.nx, ref.ny, ref.nz
    ph = rng.random(6) * 2 * np.pi
    for name, comp in (("u", 0), ("um", 0), ("v", 1), ("vm", 1), ("w", 2), ("wm", 2)):
        a = ref.get(name)
        b = T.bounds(name, nx, ny, nz)
        I, J, K = np.meshgrid(*[np.arange(l, h + 1) for l, h in b], indexing="ij")
        x, y, z = 2 * np.pi * I / nx, 2 * np.pi * J / ny, np.pi * K / nz
        if comp == 0:
            f = u0 + amp * np.sin(x + ph[0]) * np.cos(y + ph[1]) * np.sin(z)
        elif comp == 1:
            f = amp * np.cos(x + ph[2]) * np.sin(y + ph[3]) * np.sin(z)
        else:
            f = 0.2 * amp * np.sin(x + ph[4]) * np.sin(y + ph[5]) * np.sin(z)
        core = (a != -1.0e34)
        a[core] = f[core]
        ref.set(name, a)
    t = ref.get("t")
    t += tnoise * (2.0 * rng.random(t.shape) - 1.0)
    ref.set("t", t)
    ref.call("density")


def ref_init(ref, seed, amp=0.02):
    """Reference set-up (INITIALIZE_PROFILES, DENSITY, INITIALIZE,
    SPONGE_MAKER: hotcross80-revised2014.f:210-255), then a seeded random
    perturbation of the prognostic fields so that every term is exercised."""
    ref.init_profiles(PROFILE)
    ref.call("density")
    ref.call("initialize")
    ref.call("sponge_maker")
    rng = np.random.default_rng(seed)
    nx, ny, nz = ref.nx, ref.ny, ref.nz
    for name in ("u", "um", "v", "vm", "w", "wm"):
        a = ref.get(name)
        core = (a != -1.0e34)
        a[core] = a[core] + amp * (rng.random(core.sum()) - 0.5)
        ref.set(name, a)
    for name, scale in (("t", 2e-3), ("s", 1e-4)):
        a = ref.get(name)
        a += scale * (rng.random(a.shape) - 0.5)
        ref.set(name, a)
    c = ref.get("c")
    c[...] = 0.5 * rng.random(c.shape)
    ref.set("c", c)
    ref.call("density")
    # smooth positive eddy viscosities so that FORCES sees realistic inputs
    for name in ("ss", "ss_v"):
        a = ref.get(name)
        a[...] = 1e-4 * (1.0 + rng.random(a.shape))
        ref.set(name, a)


def dump(ref):
    return {n: ref.get(n) for n in DUMP}


def worker_sequence(ref, seed, steps, calls, amp=0.02):
    """Initial state, then for each step the given calls; returns the dump
    before the first call, the dump after, and the recorded source noise."""
    ref_init(ref, seed, amp=amp)
    before = dump(ref)
    noises = []
    for m in steps:
        ref.call("begin_step", m)
        for c in calls:
            if c == "source_terms":
                ref.call("source_terms")
                noises.append(ref.noise())
            elif c.startswith("tracer"):
                ref.call("tracer", int(c[-1]))
            else:
                ref.call(c)
    return before, dump(ref), noises


def worker_steps(ref, seed, nsteps, dump_every=0):
    ref_init_smooth(ref, seed)
    before = dump(ref)
    noises, mids = [], []
    for m in range(nsteps):
        ref.call("step", m)
        noises.append(ref.noise())
        if dump_every and (m + 1) % dump_every == 0:
            mids.append(dump(ref))
    return before, dump(ref), noises, mids


def oracle_from(prm, state):
    o = T.Oracle(prm)
    for n, a in state.items():
        o.set(n, a)
    return o


def oracle_sequence(o, steps, calls, noises):
    prm = o.prm
    q = 0
    for m in steps:
        o.begin_step(m)
        for c in calls:
            if c == "source_terms":
                o.source_terms(noises[q]); q += 1
            elif c == "tracer1":
                tb = o.get("t_bkg")
                o.tracer("t", tb.min(), 1.0e10, prm.dt)
                o.vent("t", prm.tsource + tb[prm.nz])
            elif c == "tracer2":
                sb = o.get("s_bkg")
                o.tracer("s", 0.0, sb.max(), prm.dt)
                o.vent("s", prm.ssource + sb[prm.nz])
            elif c == "tracer3":
                o.set("c_bkg", np.zeros(prm.nz + 2))
                o.tracer("c", 0.0, 1.0e10, prm.dt)
            else:
                o.call(c)


def plant_cfl_extremes(state, prm, seed):
    """Velocities for CFL_VALUES checks: seeded values of magnitude below
    0.2 dx/dt inside the scanned ranges (u: i=0..nx, j=1..ny, k=1..nz; v:
    i=1..nx, j=0..ny, k=1..nz; w: i=1..nx, j=1..ny, k=0..nz,
    hotcross80-revised2014.f:395-402), the range's extremes planted on its
    first and last cells, and larger values (0.6 dx/dt, still under the
    bailout) on the cells just outside every range face, which a scan over
    the wrong range would report."""
    rng = np.random.default_rng(seed)
    nx, ny, nz = prm.nx, prm.ny, prm.nz
    out = dict(state)
    for name, d, lo in (("u", prm.dx, (0, 1, 1)), ("v", prm.dy, (1, 0, 1)), ("w", prm.dz, (1, 1, 0))):
        b = T.bounds(name, nx, ny, nz)
        a = np.array(state[name], dtype=np.float64, order="F", copy=True)
        c = d / prm.dt
        a[...] = 0.2 * c * (2.0 * rng.random(a.shape) - 1.0)
        hi = (nx, ny, nz)
        ix = lambda i, j, k: (i - b[0][0], j - b[1][0], k - b[2][0])
        a[ix(*lo)] = 0.45 * c
        a[ix(*hi)] = -0.4 * c
        for ax in range(3):
            for edge, val in ((lo[ax] - 1, 0.6 * c), (hi[ax] + 1, -0.6 * c)):
                if b[ax][0] <= edge <= b[ax][1]:
                    sl = [slice(None)] * 3
                    sl[ax] = edge - b[ax][0]
                    a[tuple(sl)] = val
        out[name] = a
    return out


def worker_cfl(ref, seed):
    """The reference's CFL_VALUES on its initial state with planted
    velocities: returns (state, six Courant numbers as it printed them)."""
    ref.init_profiles(PROFILE)
    ref.call("density")
    ref.call("initialize")
    ref.call("sponge_maker")
    st = {n: ref.get(n) for n in ("u", "v", "w")}
    st = plant_cfl_extremes(st, ref.prm, seed)
    for n in ("u", "v", "w"):
        ref.set(n, st[n])
    ref.call("cfl_values")
    with open("fort.81") as fh:
        line = fh.read().splitlines()[-1]
    vals = [float(line[15 * q:15 * (q + 1)]) for q in range(6)]
    return st, np.array(vals)


def worker_profiles(ref, path):
    """INITIALIZE_PROFILES of the reference on the table at path: its
    t_bkg and s_bkg."""
    ref.init_profiles(path)
    return ref.get("t_bkg"), ref.get("s_bkg")
