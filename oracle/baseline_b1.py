"""Reference-path CPU baseline: a pure Python/NumPy FP64 restatement of the
reference's per-event work, run under multiprocessing like doScatter.py.

TEST INFRASTRUCTURE ONLY (the checker and the CPU baseline): used by tests/ and
bench.py's CPU-baseline leg, never by the product.

It repeats the reference's per-step object work so its cost is representative:
a cross-section dict per step (scatter_discrete.__init__, MC/tools/scattering.py:440-527),
a descending sort + np.cumsum + bisect for the channel (pickFromSigmas, :39-59),
bisect on numpy.ma float32 CDF rows (pickTable, :63-94), a quaternion product
and a rotation-matrix rotation of (d, y_local) (rotation.py:41-53), and
per-step record appends (fileTools.thingsToSave, :193-242).  Electrons are run
4 per job in a multiprocessing.Pool (doScatter.py:46,87-94,216).  The random
numbers are the Philox stream of the transport contract (4 per step), generated
vectorised per trajectory block, so trajectories agree with the C oracle.
"""
from __future__ import annotations

import bisect
import math
import operator
import os
import time
from multiprocessing import Pool

import numpy as np
import numpy.ma as ma

PI_EFOUR = 6.51408491409531e-14
BOHR_R = 5.29177210903e-11
M0, M1, W0, W1 = 0xD2511F53, 0xCD9E8D57, 0x9E3779B9, 0xBB67AE85


def philox_block(electron, config, steps, seed):
    """Philox4x32-10 for counters (electron low 32 bits, config | electron
    bits 32..47 << 16, step, 0), step in `steps`; returns (len(steps), 4)
    uniforms (2*(x>>9)+1)*2^-24."""
    n = len(steps)
    mask = np.uint64(0xFFFFFFFF)
    c0 = np.full(n, electron & 0xFFFFFFFF, dtype=np.uint64)
    c1 = np.full(n, config | (electron >> 32) << 16, dtype=np.uint64)
    c2 = np.asarray(steps, dtype=np.uint64)
    c3 = np.zeros(n, dtype=np.uint64)
    k0, k1 = seed & 0xFFFFFFFF, (seed >> 32) & 0xFFFFFFFF
    for r in range(10):
        if r:
            k0 = (k0 + W0) & 0xFFFFFFFF
            k1 = (k1 + W1) & 0xFFFFFFFF
        p0 = np.uint64(M0) * c0
        p1 = np.uint64(M1) * c2
        n0 = (p1 >> np.uint64(32)) ^ c1 ^ np.uint64(k0)
        n2 = (p0 >> np.uint64(32)) ^ c3 ^ np.uint64(k1)
        c1 = p1 & mask
        c3 = p0 & mask
        c0, c2 = n0, n2
    x = np.stack([c0, c1, c2, c3], axis=1)
    return ((x >> np.uint64(9)) * np.uint64(2) + np.uint64(1)).astype(np.float64) * 2.0 ** -24


class Rng:
    """Per-electron Philox stream consumed in the reference's draw order."""

    def __init__(self, electron, config, seed, mode_ds=True, block=512):
        self.e, self.c, self.seed, self.block = electron, config, seed, block
        self.buf, self.base, self.step, self.slot = None, 0, -1, 0
        self.ds = mode_ds

    def new_step(self):
        self.step += 1
        if self.buf is None or self.step >= self.base + self.block:
            self.base = self.step
            self.buf = philox_block(self.e, self.c, range(self.base, self.base + self.block), self.seed)
        self.row = self.buf[self.step - self.base]

    def u(self, slot):
        return float(self.row[slot])


# -- crossSections.py / stoppingPowers.py (FP64) ------------------------------
def alpha(E, Z):
    return 3.4 * (Z ** (0.666667)) / E


def ruther_sigma(E, Z):
    a = alpha(E, Z)
    Ek = E * 1e-3
    return 5.21e-21 * (Z * Z / (Ek * Ek)) * ((Ek + 511.) / (Ek + 1024.)) ** 2 * (4. * math.pi) / (a * (1. + a))


def moller_sigma(E, Emin, nfree):
    if E <= Emin:
        return 0.
    eps = Emin * 1. / E
    return nfree * PI_EFOUR * (1. / (E * E)) * (1. / eps - 1. / (1. - eps) + np.log(eps / (1. - eps)))


def gryz_sigma(E, Esi, nsi):
    if E <= Esi:
        return 0.
    U = E * 1. / Esi
    return nsi * PI_EFOUR * ((U - 1.) / (U + 1.)) ** 1.5 * (1. + (2. * (1. - (1. / (2. * U)))) / 3. *
                                                            np.log(2.7 + ((U - 1.) ** 0.5))) / (E * Esi)


def quinn_sigma(E, Epl, Ef, n):
    if E <= Epl:
        return 0.
    E_Ef = E * 1. / Ef
    Epl_Ef = Epl * 1. / Ef
    tot = Epl * np.log(((1. + Epl_Ef) ** 0.5 - 1.) / (E_Ef ** 0.5 - (E_Ef - Epl_Ef) ** 0.5)) / (2. * BOHR_R * E)
    return (tot / n) * 1e4


def mfp_from_sigma(sigma, n):
    return 1. / (n * sigma) * 1e14


def pick_from_sigmas(sigma_dict, u):
    sorted_sigmas = dict(sorted(sigma_dict.items(), key=operator.itemgetter(1), reverse=True))
    probs = np.cumsum(list(sorted_sigmas.values())) / sum(sorted_sigmas.values())
    return list(sorted_sigmas.keys())[bisect.bisect_left(probs, u)]


class MTable:
    """A probability table as read back by the reference (masked zeros)."""

    def __init__(self, Es, Ws, cdf):
        self.Es = np.asarray(Es, dtype=np.float32)
        self.Ws = np.asarray(Ws, dtype=np.float32)
        cdf = np.asarray(cdf, dtype=np.float32)
        self.table = ma.masked_array(cdf, cdf == 0)


def pick_table(table, energy, u):
    i = bisect.bisect_left(table.Es, energy)
    j = bisect.bisect_left(table.table[i], u)
    return table.Ws[j]


# -- quaternion rotation (numpy-quaternion semantics) --------------------------
def qmul(a, b):
    a1, b1, c1, d1 = a
    a2, b2, c2, d2 = b
    return (a1 * a2 - b1 * b2 - c1 * c2 - d1 * d2, a1 * b2 + b1 * a2 + c1 * d2 - d1 * c2,
            a1 * c2 - b1 * d2 + c1 * a2 + d1 * b2, a1 * d2 + b1 * c2 - c1 * b2 + d1 * a2)


def rotation_matrix(q):
    w, x, y, z = q
    n = w * w + x * x + y * y + z * z
    return np.array([[1.0 - 2 * (y * y + z * z) / n, 2 * (x * y - z * w) / n, 2 * (x * z + y * w) / n],
                     [2 * (x * y + z * w) / n, 1.0 - 2 * (x * x + z * z) / n, 2 * (y * z - x * w) / n],
                     [2 * (x * z - y * w) / n, 2 * (y * z + x * w) / n, 1.0 - 2 * (x * x + y * y) / n]])


def newdir(s_hT, c_hT, s_hP, c_hP, y, d):
    q_polar = (c_hT, y[0] * s_hT, y[1] * s_hT, y[2] * s_hT)
    q_az = (c_hP, d[0] * s_hP, d[1] * s_hP, d[2] * s_hP)
    m = rotation_matrix(qmul(q_az, q_polar))
    return m @ d, m @ y


class Things:
    def __init__(self, args):
        self.dict = {a: [] for a in args}

    def add(self, k, v):
        if k in self.dict:
            self.dict[k].append(v)


class Electron:
    __slots__ = ["energy", "Emin", "xyz", "dir", "outcome", "y_local", "el", "sc"]

    def __init__(self, energy, Emin, position, direction, out):
        self.energy, self.Emin, self.xyz, self.dir = energy, Emin, position, direction
        self.outcome = None
        self.y_local = np.array([0., 1., 0.])
        self.el, self.sc = Things(out["el_output"]), Things(out["scat_output"])

    def save(self):
        self.el.add("outcome", self.outcome)
        self.el.add("final_E", self.energy)
        self.el.add("final_dir", self.dir)
        self.el.add("last_pos", self.xyz)

    def update_energy(self, W):
        e = self.energy - W
        if e <= self.Emin:
            self.outcome = "abs"
            self.save()
        self.energy = e
        self.sc.add("energy", self.energy)

    def update_xyz(self, l, zmax):
        p = self.xyz + l * self.dir
        if p[2] <= 0.:
            self.outcome = "bks"
            self.save()
        elif p[2] >= zmax:
            self.outcome = "trsm"
            self.save()
        else:
            self.xyz = p

    def update_direction(self, c2, hp):
        s_hT = (1. - c2) ** 0.5
        c_hT = c2 ** 0.5
        s_hP = math.sin(hp)
        c_hP = (1. - s_hP ** 2) ** 0.5
        nd, ny = newdir(s_hT, c_hT, s_hP, c_hP, self.y_local, self.dir)
        self.dir, self.y_local = nd / np.linalg.norm(nd), ny / np.linalg.norm(ny)


_W = {}


def _init_worker(params):
    _W.update(params)
    _W["tm"] = MTable(*params["tm"]) if params.get("tm") else None
    _W["tg"] = MTable(*params["tg"]) if params.get("tg") else None


def trajectory_ds(el, P, rng):
    """trajectory_DS (singleScatter.py:10-41) with the scatter_discrete steps."""
    nsc = 0
    while el.outcome not in ("abs", "bks", "trsm", "trdf", "long"):
        E = el.energy
        sig = {"Ruth": ruther_sigma(E, P["Z"]), "Moller": moller_sigma(E, P["Wc"], P["n_val"])}
        for name, Eb, ns in P["shells"]:
            sig["Gryz" + name] = gryz_sigma(E, Eb, ns)
        sig["Quinn"] = quinn_sigma(E, P["Epl"], P["Ef"], P["atnd"])
        mfp = mfp_from_sigma(sum(sig.values()), P["atnd"])
        el.sc.add("position", el.xyz)
        nsc += 1
        if P["maxScatt"] and nsc > P["maxScatt"]:
            el.outcome = "long"
            el.save()
            return nsc - 1
        rng.new_step()
        typ = pick_from_sigmas(sig, rng.u(0))
        el.sc.add("type", typ)
        l = -mfp * np.log(rng.u(1))
        el.sc.add("pathl", l)
        el.update_xyz(l, P["thickness"])
        if "Ruth" in typ:
            W = 0.
            a = alpha(E, P["Z"])
            u = rng.u(2)
            c2 = 1. - (a * u / (1. + a - u))
        elif typ == "Quinn":
            W = P["Epl"]
            c2 = 1.
        else:
            W = float(pick_table(_W["tm"] if typ == "Moller" else _W["tg"], E, rng.u(2)))
            c2 = float(np.where(W <= (E - P["Ef"]) * 0.5, 0.5 * (np.sqrt(1 - (W / E)) + 1), np.nan))
        el.sc.add("E", E)
        el.sc.add("E_loss", W)
        el.sc.add("pol_angle", c2)
        hp = math.pi * rng.u(3)
        el.sc.add("az_angle", hp)
        el.update_energy(W)
        el.update_direction(c2, hp)
    return nsc


def trajectory_cont(el, P, rng):
    """trajectory_cont_cl (singleScatter.py:205-239) with the Bethe law chosen by
    the configuration (stoppingPowers.py:112-215)."""
    nsc = 0
    Z, n, J = P["Z"], P["atnd"], P["J"]
    while el.outcome not in ("abs", "bks", "trsm"):
        E = el.energy
        mfp = mfp_from_sigma(ruther_sigma(E, Z), n)
        nsc += 1
        if P["maxScatt"] and nsc > P["maxScatt"]:
            el.outcome = "long"
            el.save()
            return nsc - 1
        rng.new_step()
        l = -mfp * np.log(rng.u(0))
        el.update_xyz(l, P["thickness"])
        if P["bethe"] == "classical":
            S = 2. * PI_EFOUR * n * (Z / E) * math.log(1.166 * E / J) * 1e-14
        elif P["bethe"] == "JL":
            S = 2. * PI_EFOUR * n * (Z / E) * math.log(1.166 * (E + P["k"] * J) / J) * 1e-14
        else:
            s = P["n_val"] * math.log(E / P["E_val"])
            for _, Eb, ns in P["shells"]:
                s = s + ns * math.log(E / Eb)
            S = (2. * PI_EFOUR * n / E) * 1e-14 * s
        W = l * S
        el.update_energy(W)
        a = alpha(E, Z)
        u = rng.u(1)
        c2 = 1. - (a * u / (1. + a - u))
        hp = math.pi * rng.u(2)
        el.update_direction(c2, hp)
    return nsc


def multi_traj(first, num):
    """multiTraj_DS / multiTraj_cont (multiScatter.py:66-116,171-229) for a job
    of `num` electrons starting at `first`; returns (events, outcomes)."""
    P = _W
    t = P["s_tilt"] * (math.pi / 180.)
    pos0 = np.array([0., 0., 0.])
    dir0 = np.array([-np.sin(t), 0., np.cos(t)])
    out = {"el_output": ["outcome", "final_E", "final_dir", "last_pos"], "scat_output": P["scat_output"]}
    events, results = 0, []
    for k in range(num):
        el = Electron(P["E0"], P["Emin"], pos0, dir0, out)
        rng = Rng(first + k, P["config"], P["seed"])
        if P["mode"] == "DS":
            events += trajectory_ds(el, P, rng)
        else:
            events += trajectory_cont(el, P, rng)
        results.append({"els": el.el.dict, "scats": el.sc.dict})
    return events, results


def params_from(cfg, moller=None, gryz=None, seed=1, config=0, scat_output=("type", "E", "energy")):
    """Worker parameters from an mcd_config (ctypes) and reference-layout tables."""
    m = cfg.mat
    return dict(Z=int(m.Z), n_val=float(m.n_val), E_val=float(m.E_val), k=float(m.bethe_k), atnd=float(m.atnd),
                Ef=float(m.fermi_e), Epl=float(m.plasmon_e),
                shells=[(str(i), float(m.shell_Eb[i]), float(m.shell_n[i])) for i in range(m.n_shells)],
                Wc=float(cfg.Wc), E0=float(cfg.E0), Emin=float(cfg.Emin), s_tilt=float(cfg.s_tilt),
                thickness=float(cfg.thickness), maxScatt=int(cfg.max_scatt), J=float(cfg.J),
                bethe=["classical", "JL", "explicit"][cfg.bethe_model], mode="DS" if cfg.mode == 0 else "cont",
                seed=int(seed), config=int(config), scat_output=list(scat_output),
                tm=(moller.Es, moller.Ws, moller.cdf) if moller is not None else None,
                tg=(gryz.Es, gryz.Ws, gryz.cdf) if gryz is not None else None)


def run_serial(params, first, n):
    _init_worker(params)
    return multi_traj(first, n)


def run_pool(params, n_electrons, workers=None, per_job=4):
    """Time n_electrons through a Pool of `workers` processes (tables loaded once
    per worker).  Returns (events, seconds, workers)."""
    workers = workers or os.cpu_count()
    jobs = [(i, min(per_job, n_electrons - i)) for i in range(0, n_electrons, per_job)]
    t = time.perf_counter()
    with Pool(workers, initializer=_init_worker, initargs=(params,)) as pool:
        res = pool.starmap(_job_events, jobs, chunksize=max(1, len(jobs) // (8 * workers)))
    dt = time.perf_counter() - t
    return int(sum(res)), dt, workers


def _job_events(first, num):
    ev, _ = multi_traj(first, num)
    return ev


def _job_events_shipped(params, first, num):
    """One job as the reference ships it: the job's arguments carry the
    parameters and every table (functools.partial over multiTraj_DS with
    `tables`, doScatter.py:229-233, pickled with each apply_async call,
    doScatter.py:88-90), so the worker loads them per job."""
    _init_worker(params)
    ev, _ = multi_traj(first, num)
    return ev


def run_pool_shipped(params, n_electrons, workers=None, per_job=4, n_gryz_tables=1):
    """Time n_electrons through a Pool of `workers` processes with one
    apply_async call per job of `per_job` electrons, each carrying the tables:
    the Moller table and `n_gryz_tables` Gryzinski tables (the reference's
    tables dict holds one per shell, genTables probTables.py:541-614; the
    extra ones are copies of the outermost shell's, of the same size).
    Returns (events, seconds, workers)."""
    workers = workers or os.cpu_count()
    shipped = dict(params)
    if params.get("tg") is not None:
        # distinct copies, so that pickle writes each one (it would share a
        # repeated reference to one array)
        shipped["tg_other_shells"] = [tuple(np.array(a, copy=True) for a in params["tg"])
                                      for _ in range(max(0, n_gryz_tables - 1))]
    t = time.perf_counter()
    with Pool(workers) as pool:
        res = [pool.apply_async(_job_events_shipped, (shipped, i, min(per_job, n_electrons - i)))
               for i in range(0, n_electrons, per_job)]
        ev = sum(r.get() for r in res)
    dt = time.perf_counter() - t
    return int(ev), dt, workers
