"""ctypes binding of the CPU oracle (oracle/liblto.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker.  The product package
(lettuce_amd) never imports this module.

Fields are numpy float64 arrays of shape (V, 4, 18) in the reference AoS layout
(LettuceGauge/lattice.hpp:366-382).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liblto.so")


class Dims(ctypes.Structure):
    _fields_ = [("nt", ctypes.c_int), ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int)]


def build():
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE, "liblto.so"], check=True)


def _load():
    if not os.path.exists(_LIB_PATH):
        build()
    lib = ctypes.CDLL(_LIB_PATH)
    D = ctypes.POINTER(ctypes.c_double)
    dims = Dims
    sig = {
        "lto_philox4x32_10": (None, [ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_uint32)]),
        "lto_u01": (ctypes.c_double, [ctypes.c_uint32, ctypes.c_uint32]),
        "lto_volume": (ctypes.c_size_t, [dims]),
        "lto_threads": (ctypes.c_int, []),
        "lto_set_identity": (None, [D, dims]),
        "lto_haar": (None, [D, dims, ctypes.c_uint64]),
        "lto_reunitarize": (None, [D, dims]),
        "lto_su3_exp": (None, [D, D]),
        "lto_gram_schmidt": (None, [D]),
        "lto_algebra": (None, [D, D]),
        "lto_staple": (None, [D, dims, ctypes.c_size_t, ctypes.c_int, D]),
        "lto_staple_field": (None, [D, dims, D]),
        "lto_or_sweep": (None, [D, dims, ctypes.c_int, ctypes.c_int]),
        "lto_hb_sweep": (None, [D, dims, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64)]),
        "lto_hb_link": (None, [D, dims, ctypes.c_double, ctypes.c_size_t, ctypes.c_int, ctypes.c_int, ctypes.c_uint64, ctypes.c_uint64, ctypes.POINTER(ctypes.c_int)]),
        "lto_plaquette_sum": (ctypes.c_double, [D, dims]),
        "lto_wilson_action": (ctypes.c_double, [D, dims, ctypes.c_double]),
        "lto_clover": (None, [D, dims, D]),
        "lto_topcharge_clover": (ctypes.c_double, [D, dims]),
        "lto_topcharge_density": (None, [D, dims, D]),
        "lto_topcharge_clover_tensor": (ctypes.c_double, [D, dims]),
        "lto_topcharge_timeslice": (None, [D, dims, D]),
        "lto_energy_clover_timeslice": (None, [D, dims, D]),
        "lto_energy_plaquette": (ctypes.c_double, [D, dims]),
        "lto_stout": (None, [D, D, dims, ctypes.c_double]),
        "lto_stout_tape": (None, [D, D, D, dims, ctypes.c_double]),
        "lto_flow_euler": (None, [D, D, dims, ctypes.c_double, ctypes.c_int]),
        "lto_flow_rk3": (None, [D, D, dims, ctypes.c_double, ctypes.c_int]),
        "lto_clover_derivative": (None, [D, D, dims, D]),
        "lto_stout_force_recursion": (None, [D, D, D, D, dims, ctypes.c_double, D]),
        "lto_cv_force": (ctypes.c_double, [D, dims, ctypes.c_double, ctypes.c_int, ctypes.c_double, D]),
        "lto_cv_charge": (ctypes.c_double, [D, dims, ctypes.c_double, ctypes.c_int]),
        "lto_random_momentum": (None, [D, dims, ctypes.c_uint64, ctypes.c_uint64]),
        "lto_gauge_force_update": (None, [D, D, dims, ctypes.c_double, ctypes.c_double]),
        "lto_field_update": (None, [D, D, dims, ctypes.c_double]),
        "lto_kinetic": (ctypes.c_double, [D, dims]),
        "lto_third_row_fma": (None, [D, ctypes.c_size_t]),
        "lto_rect_sums": (None, [D, dims, D, D]),
        "lto_topcharge_plaquette": (ctypes.c_double, [D, dims]),
        "lto_wilson_loop": (ctypes.c_double, [D, D, dims, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
        "lto_polyakov_loop": (None, [D, dims, D, D]),
        "lto_parity_update": (None, [D, D, dims]),
        "lto_staple_action_field": (None, [D, dims, ctypes.c_double, ctypes.c_double, D]),
        "lto_or_sweep_action": (None, [D, dims, ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int]),
        "lto_hb_sweep_action": (None, [D, dims, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                       ctypes.c_int, ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = _load()
    return _lib


def _p(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def dims(shape):
    nt, nx, ny, nz = shape
    return Dims(nt, nx, ny, nz)


def threads():
    """OpenMP threads of the oracle's parallel regions."""
    return lib().lto_threads()


def volume(shape):
    return int(np.prod(shape))


def new_field(shape):
    return np.zeros((volume(shape), 4, 18), dtype=np.float64)


def haar(shape, seed):
    U = new_field(shape)
    lib().lto_haar(_p(U), dims(shape), seed)
    return U


def identity(shape):
    U = new_field(shape)
    lib().lto_set_identity(_p(U), dims(shape))
    return U


def philox(ctr, key):
    c = (ctypes.c_uint32 * 4)(*ctr)
    k = (ctypes.c_uint32 * 2)(*key)
    o = (ctypes.c_uint32 * 4)()
    lib().lto_philox4x32_10(c, k, o)
    return list(o)


def su3_exp(Q):
    Q = np.ascontiguousarray(Q, dtype=np.float64).reshape(18)
    out = np.zeros(18)
    lib().lto_su3_exp(_p(Q), _p(out))
    return out


def staple_field(U, shape):
    ST = np.zeros_like(U)
    lib().lto_staple_field(_p(U), dims(shape), _p(ST))
    return ST


def or_sweep(U, shape, n=1, schedule=2):
    lib().lto_or_sweep(_p(U), dims(shape), n, schedule)


def hb_sweep(U, shape, beta, n=1, schedule=2, seed=0, counter=0):
    c = ctypes.c_uint64(counter)
    lib().lto_hb_sweep(_p(U), dims(shape), beta, n, schedule, seed, ctypes.byref(c))
    return c.value


def hb_link(U, shape, beta, site, mu, max_iter, seed, counter):
    att = (ctypes.c_int * 3)()
    lib().lto_hb_link(_p(U), dims(shape), beta, site, mu, max_iter, seed, counter, att)
    return list(att)


def plaquette_sum(U, shape):
    return lib().lto_plaquette_sum(_p(U), dims(shape))


def wilson_action(U, shape, beta):
    return lib().lto_wilson_action(_p(U), dims(shape), beta)


def clover(U, shape):
    C6 = np.zeros((volume(shape), 6, 18))
    lib().lto_clover(_p(U), dims(shape), _p(C6))
    return C6


def topcharge(U, shape):
    return lib().lto_topcharge_clover(_p(U), dims(shape))


def topcharge_density(U, shape):
    """Per-site summand of TopChargeClover (topological_charge.hpp:184-199), shape (Nt, Nx, Ny, Nz)."""
    q = np.zeros(tuple(shape))
    lib().lto_topcharge_density(_p(U), dims(shape), _p(q))
    return q


def topcharge_tensor(C6, shape):
    return lib().lto_topcharge_clover_tensor(_p(C6), dims(shape))


def topcharge_timeslice(U, shape):
    q = np.zeros(shape[0])
    lib().lto_topcharge_timeslice(_p(U), dims(shape), _p(q))
    return q


def energy_clover_timeslice(U, shape):
    e = np.zeros(shape[0])
    lib().lto_energy_clover_timeslice(_p(U), dims(shape), _p(e))
    return e


def energy_plaquette(U, shape):
    return lib().lto_energy_plaquette(_p(U), dims(shape))


def observables(U, shape, beta=6.0):
    """Same vector layout as the reference harness 'obs' command."""
    C6 = clover(U, shape)
    qt = topcharge_timeslice(U, shape)
    et = energy_clover_timeslice(U, shape)
    head = [plaquette_sum(U, shape), wilson_action(U, shape, beta), topcharge(U, shape),
            topcharge_tensor(C6, shape), float(qt.sum()), energy_plaquette(U, shape), float(et.sum())]
    return np.concatenate([np.array(head), qt, et])


def stout(U, shape, rho, n=1):
    a = U.copy()
    b = np.zeros_like(U)
    for _ in range(n):
        lib().lto_stout(_p(a), _p(b), dims(shape), rho)
        a, b = b, a
    return a


def stout_tape(U, shape, rho):
    out = np.zeros_like(U)
    tape = np.zeros_like(U)
    lib().lto_stout_tape(_p(U), _p(out), _p(tape), dims(shape), rho)
    return out, tape


def flow(U, shape, eps, n, integrator="rk3"):
    W = U.copy()
    Z = np.zeros_like(U)
    f = lib().lto_flow_rk3 if integrator == "rk3" else lib().lto_flow_euler
    f(_p(W), _p(Z), dims(shape), eps, n)
    return W


def clover_derivative(U, shape):
    C6 = clover(U, shape)
    F = np.zeros_like(U)
    lib().lto_clover_derivative(_p(U), _p(C6), dims(shape), _p(F))
    return F


def cv_force(U, shape, rho, n_levels, dvdq=1.0):
    F = np.zeros_like(U)
    Q = lib().lto_cv_force(_p(U), dims(shape), rho, n_levels, dvdq, _p(F))
    return Q, F


def cv_charge(U, shape, rho, n_levels):
    return lib().lto_cv_charge(_p(U), dims(shape), rho, n_levels)


def random_momentum(shape, seed, counter):
    P = new_field(shape)
    lib().lto_random_momentum(_p(P), dims(shape), seed, counter)
    return P


def gauge_force_update(U, P, shape, eps, beta):
    lib().lto_gauge_force_update(_p(U), _p(P), dims(shape), eps, beta)


def field_update(U, P, shape, eps):
    lib().lto_field_update(_p(U), _p(P), dims(shape), eps)


def kinetic(P, shape):
    return lib().lto_kinetic(_p(P), dims(shape))


def omf2(U, P, shape, beta, tau, n_step):
    """Integrators::HMC::OMF_2 (LettuceGauge/updates/hmc_gauge.hpp:155-190) on (U, P) in place."""
    eps = tau / n_step
    lam = 0.1931833275037836
    alpha, b, gamma = lam * eps, 0.5 * eps, (1.0 - 2.0 * lam) * eps
    gauge_force_update(U, P, shape, alpha, beta)
    field_update(U, P, shape, b)
    gauge_force_update(U, P, shape, gamma, beta)
    field_update(U, P, shape, b)
    for _ in range(n_step - 1):
        gauge_force_update(U, P, shape, 2.0 * alpha, beta)
        field_update(U, P, shape, b)
        gauge_force_update(U, P, shape, gamma, beta)
        field_update(U, P, shape, b)
    gauge_force_update(U, P, shape, alpha, beta)


def hamiltonian(U, P, shape, beta):
    return wilson_action(U, shape, beta) + kinetic(P, shape)


class Bias:
    """MetaBiasPotential (LettuceGauge/actions/bias_potential/metadynamics.hpp:29-507)."""

    def __init__(self, cv_min, cv_max, bins, weight, width, wt_param, threshold, well_tempered=True):
        if not cv_min < cv_max:
            raise ValueError("CV_min < CV_max required")      # :70-73
        if bins <= 0:
            raise ValueError("bin_number > 0 required")       # :74-77
        self.cv_min, self.cv_max, self.bins = cv_min, cv_max, int(bins)
        self.bin_width = (cv_max - cv_min) / bins
        self.inv = 1.0 / self.bin_width
        self.weight, self.width, self.wt, self.thr = weight, width, wt_param, threshold
        self.well_tempered = well_tempered
        self.grid = np.zeros(self.bins + 1)

    def deposit(self, cv):
        """UpdatePotential (:103-124)"""
        x = self.cv_min + np.arange(self.bins + 1) * self.bin_width
        d = cv - x
        g = -0.5 / (self.width * self.width) * d * d
        if self.well_tempered:
            g = g - self.grid / self.wt
        self.grid += self.weight * np.exp(g)

    def deposit_symmetric(self, cv):
        """UpdatePotentialSymmetric (:134-147)"""
        self.deposit(cv)
        self.deposit(-cv)

    def potential(self, cv):
        """ReturnPotential (:151-179)"""
        i = int(np.floor((cv - self.cv_min) * self.inv))
        if 0 <= i < self.bins:
            c = (cv - (self.cv_min + i * self.bin_width)) * self.inv
            return self.grid[i] * (1.0 - c) + c * self.grid[i + 1]
        if i < 0:
            return self.grid[0] + self.thr * (cv - self.cv_min) ** 2
        return self.grid[-1] + self.thr * (cv - self.cv_max) ** 2

    def derivative(self, cv):
        """ReturnDerivative (:181-187)"""
        h = 0.5 * self.bin_width
        return self.inv * (self.potential(cv + h) - self.potential(cv - h))


def metad_omf2(U, P, shape, beta, tau, n_step, rho, n_levels, bias):
    """HMCMetaDKernel with Integrators::HMC::OMF_2 on given momenta (hmc_metadynamics.hpp:374-459,
    hmc_gauge.hpp:155-190).  Returns (H_old, H_new, cv_old, cv_new, samples); U, P evolve in place."""
    eps = tau / n_step
    lam = 0.1931833275037836
    alpha, b, gamma = lam * eps, 0.5 * eps, (1.0 - 2.0 * lam) * eps
    samples = []

    def mom(e, sub):
        q = cv_charge(U, shape, rho, n_levels)
        if sub:
            samples.append(q)
        _, F = cv_force(U, shape, rho, n_levels, bias.derivative(q))
        gauge_force_update(U, P, shape, e, beta)
        P[...] += e * F

    cv_old = cv_charge(U, shape, rho, n_levels)
    H_old = hamiltonian(U, P, shape, beta) + bias.potential(cv_old)
    mom(alpha, False)
    field_update(U, P, shape, b)
    mom(gamma, False)
    field_update(U, P, shape, b)
    for _ in range(n_step - 1):
        mom(2.0 * alpha, True)
        field_update(U, P, shape, b)
        mom(gamma, False)
        field_update(U, P, shape, b)
    mom(alpha, True)
    cv_new = cv_charge(U, shape, rho, n_levels)
    H_new = hamiltonian(U, P, shape, beta) + bias.potential(cv_new)
    return H_old, H_new, cv_old, cv_new, samples


OMF2_LAMBDA = 0.1931833275037836
OMF4_ABGD = (0.08398315262876693, 0.2539785108410595, 0.6822365335719091, -0.03230286765269967)


def integrator_schedule(name, tau, n_step):
    """Kick / drift events [(kind, eps, sub)] of Integrators::HMC::* (hmc_gauge.hpp:21-295,
    hmc_metadynamics.hpp:28-238): kind 'P' UpdateMomenta, 'G' UpdateGaugeMomenta,
    'M' UpdateMetadynamicsMomenta, 'U' UpdateFields.  The reference records path samples
    only in UpdateMomenta; for the multiple-timescale schedules sub marks the merged
    and last bias kicks (the last one of each step for omf2_omf4_slow)."""
    if n_step <= 0:
        return []
    e = tau / n_step
    ev = []

    def add(kind, x, sub=0):
        ev.append((kind, x, sub))

    a4, b4, g4, d4 = (c * e for c in OMF4_ABGD)
    m4 = (0.5 - OMF4_ABGD[2] - OMF4_ABGD[0]) * e
    n4 = (1.0 - 2.0 * OMF4_ABGD[3] - 2.0 * OMF4_ABGD[1]) * e

    if name == "leapfrog_slow":
        for _ in range(n_step):
            add("P", 0.5 * e)
            add("U", e)
            add("P", 0.5 * e, 1)
    elif name == "leapfrog":
        add("P", 0.5 * e)
        for _ in range(n_step - 1):
            add("U", e)
            add("P", e, 1)
        add("U", e)
        add("P", 0.5 * e, 1)
    elif name in ("omf2", "omf2_slow"):
        al, be, ga = OMF2_LAMBDA * e, 0.5 * e, (1.0 - 2.0 * OMF2_LAMBDA) * e
        if name == "omf2_slow":
            for _ in range(n_step):
                ev.extend([("P", al, 0), ("U", be, 0), ("P", ga, 0), ("U", be, 0), ("P", al, 1)])
        else:
            ev.extend([("P", al, 0), ("U", be, 0), ("P", ga, 0), ("U", be, 0)])
            for _ in range(n_step - 1):
                ev.extend([("P", 2.0 * al, 1), ("U", be, 0), ("P", ga, 0), ("U", be, 0)])
            add("P", al, 1)
    elif name in ("omf4", "omf4_slow"):
        body = [("U", b4, 0), ("P", g4, 0), ("U", d4, 0), ("P", m4, 0), ("U", n4, 0), ("P", m4, 0), ("U", d4, 0),
                ("P", g4, 0), ("U", b4, 0)]
        if name == "omf4_slow":
            for _ in range(n_step):
                ev.extend([("P", a4, 0)] + body + [("P", a4, 1)])
        else:
            ev.extend([("P", a4, 0)] + body)
            for _ in range(n_step - 1):
                ev.extend([("P", 2.0 * a4, 1)] + body)
            add("P", a4, 1)
    elif name in ("leapfrog_omf4", "omf2_omf4_slow", "omf2_omf4"):
        h = 0.5 if name == "omf2_omf4" else 1.0
        ca, cb, cg, cd = (h * c * e for c in OMF4_ABGD)
        cm = h * (0.5 - OMF4_ABGD[2] - OMF4_ABGD[0]) * e
        cn = h * (1.0 - 2.0 * OMF4_ABGD[3] - 2.0 * OMF4_ABGD[1]) * e
        inner = [("G", ca, 0), ("U", cb, 0), ("G", cg, 0), ("U", cd, 0), ("G", cm, 0), ("U", cn, 0), ("G", cm, 0),
                 ("U", cd, 0), ("G", cg, 0), ("U", cb, 0), ("G", ca, 0)]
        if name == "leapfrog_omf4":
            add("M", 0.5 * e)
            for _ in range(n_step - 1):
                ev.extend(inner)
                add("M", e, 1)
            ev.extend(inner)
            add("M", 0.5 * e, 1)
        else:
            oa, ob = OMF2_LAMBDA * e, (1.0 - 2.0 * OMF2_LAMBDA) * e
            if name == "omf2_omf4_slow":
                for _ in range(n_step):
                    ev.extend([("M", oa, 0)] + inner + [("M", ob, 0)] + inner + [("M", oa, 1)])
            else:
                add("M", oa)
                for _ in range(n_step - 1):
                    ev.extend(inner + [("M", ob, 0)] + inner + [("M", 2.0 * oa, 1)])
                ev.extend(inner + [("M", ob, 0)] + inner + [("M", oa, 1)])
    else:
        raise ValueError(f"unknown integrator {name}")
    return ev


def metad_trajectory(U, P, shape, beta, tau, n_step, rho, n_levels, bias, integrator="omf2"):
    """HMCMetaDKernel's molecular dynamics on given momenta with any integrator
    (hmc_metadynamics.hpp:374-491): 'P' kicks add both forces, 'G' the gauge force,
    'M' the bias force evaluated at the current field.  A zero bias (bias=None) gives
    plain HMC.  Returns (H_old, H_new, cv_old, cv_new, samples); U, P evolve in place."""
    samples = []

    def charge():
        return cv_charge(U, shape, rho, n_levels) if bias is not None else 0.0

    def energy(q):
        return hamiltonian(U, P, shape, beta) + (bias.potential(q) if bias is not None else 0.0)

    cv_old = charge()
    H_old = energy(cv_old)
    for kind, e, sub in integrator_schedule(integrator, tau, n_step):
        if kind == "U":
            field_update(U, P, shape, e)
            continue
        F = None
        if kind in ("P", "M") and bias is not None:
            q = cv_charge(U, shape, rho, n_levels)
            if sub:
                samples.append(q)
            _, F = cv_force(U, shape, rho, n_levels, bias.derivative(q))
        if kind in ("P", "G"):
            gauge_force_update(U, P, shape, e, beta)
        if F is not None:
            P[...] += e * F
    cv_new = charge()
    return H_old, energy(cv_new), cv_old, cv_new, samples


def uniform(seed, site, stream, counter, k=0, sub=0, mu=0):
    """First uniform of the Philox draw with counter (site, stream<<24 | k<<4 | sub<<2 | mu, counter)
    and key seed: the stream-3 accept draw is uniform(seed, 0xFFFFFFFF, 3, counter)."""
    o = philox([site & 0xFFFFFFFF, (stream << 24) | (k << 4) | (sub << 2) | mu, counter & 0xFFFFFFFF,
                (counter >> 32) & 0xFFFFFFFF], [seed & 0xFFFFFFFF, (seed >> 32) & 0xFFFFFFFF])
    return lib().lto_u01(o[0], o[1])


def tempering_accept(bias, cv_unbiased, cv_biased, u):
    """MetadynamicsTemperingKernel::operator() (LettuceGauge/updates/tempering.hpp:57-72):
    CV_old is the biased stream's current CV, CV_new the unbiased stream's CV,
    delta_V = V(CV_new) - V(CV_old), swap if u <= exp(-delta_V)."""
    delta_v = bias.potential(cv_unbiased) - bias.potential(cv_biased)
    return u <= np.exp(-delta_v)


def ladder_exchange(labels, strengths, cvs, shapes, bias, pt_seed, pt_counter, mode):
    """Replica exchange over a ladder of bias strengths s_0 < ... < s_{n-1} (the
    library's generalisation of tempering.hpp:57-72 to n replicas): for adjacent
    labels k, k+1 held by replicas i, j of equal extents, swap labels if
    u <= exp(-(s_k - s_{k+1}) (V(Q_j) - V(Q_i))), u = uniform(pt_seed, k, 4, pt_counter).
    mode 1: k = 0, 2, ...; 2: k = 1, 3, ...; 3: alternate by pt_counter parity.
    Returns the new labels and the list of (k, accepted)."""
    labels = list(labels)
    n = len(labels)
    out = []
    if mode == 0:
        return labels, out
    first = (pt_counter & 1) if mode == 3 else mode - 1
    for k in range(first, n - 1, 2):
        i, j = labels.index(k), labels.index(k + 1)
        if tuple(shapes[i]) != tuple(shapes[j]):
            continue
        delta = (strengths[k] - strengths[k + 1]) * (bias.potential(cvs[j]) - bias.potential(cvs[i]))
        acc = uniform(pt_seed, k, 4, pt_counter) <= np.exp(-delta)
        out.append((k, bool(acc)))
        if acc:
            labels[i], labels[j] = k + 1, k
    return labels, out


# ---------------------------------------------------------------------------
# BMW configuration files (LettuceGauge/IO/config_io/bmw_format.hpp) and the
# metadynamics bias-potential text file (actions/bias_potential/metadynamics.hpp)

ADLER64_PRIME = 4294967291


def third_row_fma(U):
    """Copy of U with row 2 rebuilt from rows 0 and 1 as the device does."""
    V = np.ascontiguousarray(U, dtype=np.float64).copy()
    lib().lto_third_row_fma(_p(V), V.size // 18)
    return V


def bmw_words(U, shape, full):
    """Native 64-bit words of the link data in file order: sites t, z, y, x (slowest
    to fastest), directions x, y, z, t, rows 0-1 (or 0-2) row-major (re, im)
    (bmw_format.hpp:29-49, 220-260).  Shape (V, 4 * K)."""
    nt, nx, ny, nz = shape
    K = 18 if full else 12
    A = np.asarray(U, dtype=np.float64).reshape(nt, nx, ny, nz, 4, 18)[..., :K]
    A = A.transpose(0, 3, 2, 1, 4, 5)[:, :, :, :, [1, 2, 3, 0], :]
    return np.ascontiguousarray(A).reshape(-1, 4 * K).view(np.uint64)


def bmw_checksum(words):
    """Checksum_Adler64 (bmw_format.hpp:55-100) over blocks = rows of `words`:
    word i of block n has r = (N - n) * 2 * B - 2 i (B words per block, N blocks);
    A += h + l, B += (2 + r h + (r - 1) l) mod p with 64-bit wrapping products;
    final = ((B mod p) << 32) + (A + 1) mod p."""
    W = np.asarray(words, dtype=np.uint64)
    nblk, bsz = W.shape
    p = np.uint64(ADLER64_PRIME)
    n = np.arange(nblk, dtype=np.uint64)[:, None]
    i = np.arange(bsz, dtype=np.uint64)[None, :]
    r = (np.uint64(nblk) - n) * np.uint64(2 * bsz) - np.uint64(2) * i
    h = W >> np.uint64(32)
    l = W & np.uint64(0xFFFFFFFF)
    with np.errstate(over="ignore"):
        b = (np.uint64(2) + r * h + (r - np.uint64(1)) * l) % p
    A = int((h + l).sum(dtype=np.uint64)) % ADLER64_PRIME
    Bs = int(b.sum(axis=1, dtype=np.uint64).astype(object).sum()) % ADLER64_PRIME
    return (Bs << 32) + (A + 1) % ADLER64_PRIME


def bmw_file_bytes(U, shape, full):
    """The complete file SaveConfigBMW / SaveConfigBMWFull write for U: the 4096-byte
    header ("#BMW Nx Ny Nz Nt <hex checksum>", a comment line, the NUL that snprintf
    leaves, newlines) and the big-endian link data."""
    nt, nx, ny, nz = shape
    W = bmw_words(U, shape, full)
    cs = "%x" % bmw_checksum(W)
    first = "#%sBMW %d %d %d %d %s\n" % ("18" if full else "", nx, ny, nz, nt, cs)
    head = (first + "\nGenerated with Lettuce\n").encode() + b"\0"
    head = head + b"\n" * (4096 - len(head))
    return head + W.astype(">u8").tobytes(), cs


def bias_file_text(cv_min, cv_max, bins, weight, width, wt_param, threshold, grid, exceeded):
    """MetaBiasPotential::SaveParameters followed by SavePotential (metadynamics.hpp:
    304-346): values in the default stream format (six significant digits, %g)."""
    g = lambda v: "%g" % v
    lines = ["SU(3)_version_1.3", "Metadynamics Potential", "CV_min: " + g(cv_min), "CV_max: " + g(cv_max),
             "bin_number: %d" % bins, "weight: " + g(weight), "width: " + g(width),
             "well_tempered_parameter: " + g(wt_param), "threshold_weight: " + g(threshold), "END_METADYN_PARAMS",
             ",".join(g(v) for v in grid), "exceeded_count: %d" % exceeded]
    return "\n".join(lines) + "\n"


# ---------------------------------------------------------------------------
# the rest of Observables() (SU3_main.cpp:89-136)

def rect_sums(U, shape):
    """sum Re tr P_munu and sum Re tr (R12 + R21) over x, mu < nu (rectangular_action.hpp:61-95)."""
    p = ctypes.c_double()
    r = ctypes.c_double()
    lib().lto_rect_sums(_p(U), dims(shape), ctypes.byref(p), ctypes.byref(r))
    return p.value, r.value


def rect_action_normalized(U, shape, c_plaq, c_rect):
    """Rectangular<2>::ActionNormalized = Action / (6 beta V) (rectangular_action.hpp:88-100)."""
    V = volume(shape)
    sp, sr = rect_sums(U, shape)
    return (c_plaq * (6 * V - sp / 3.0) + c_rect * (12 * V - sr / 3.0)) / (6.0 * V)


def topcharge_plaquette(U, shape):
    return lib().lto_topcharge_plaquette(_p(U), dims(shape))


def wilson_loop(U, chain, shape, n0, n1, reset):
    """WilsonLoop<n0, n1, reset>(U, chain); chain (V, 4, 18) is updated in place."""
    return lib().lto_wilson_loop(_p(U), _p(chain), dims(shape), n0, n1, 1 if reset else 0)


def polyakov_loop(U, shape):
    re = ctypes.c_double()
    im = ctypes.c_double()
    lib().lto_polyakov_loop(_p(U), dims(shape), ctypes.byref(re), ctypes.byref(im))
    return complex(re.value, im.value)


def parity_update(U, shape):
    """ParityUpdate (updates/parity_update.hpp:22-47): a new field."""
    out = new_field(shape)
    lib().lto_parity_update(_p(U), _p(out), dims(shape))
    return out


# improved gauge actions Rectangular<2> (rectangular_action.hpp:29-300)
LUESCHER_WEISZ = (1.0 + 8.0 / 12.0, -1.0 / 12.0)
IWASAKI = (1.0 + 8.0 * 0.331, -0.331)
DBW2 = (1.0 + 8.0 * 1.4088, -1.4088)


def staple_action_field(U, shape, c_plaq, c_rect):
    ST = np.zeros_like(U)
    lib().lto_staple_action_field(_p(U), dims(shape), c_plaq, c_rect, _p(ST))
    return ST


def or_sweep_action(U, shape, c_plaq, c_rect, n=1, schedule=4):
    lib().lto_or_sweep_action(_p(U), dims(shape), c_plaq, c_rect, n, schedule)
    return U


def hb_sweep_action(U, shape, beta, c_plaq, c_rect, n=1, schedule=4, seed=0, counter=0):
    c = ctypes.c_uint64(counter)
    lib().lto_hb_sweep_action(_p(U), dims(shape), beta, c_plaq, c_rect, n, schedule, seed, ctypes.byref(c))
    return c.value
