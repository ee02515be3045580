"""Heatbath parity is statistical: the device draws Philox numbers where the reference
draws pcg64 ones, so chains agree in distribution, not link by link.  The reference's
chains (tests/golden/hbq16.npz, written by tests/golden/make_golden.py --hbq from the
compiled reference: 8 pcg64 seeds x 200 updates at 16^4, beta = 5.7, cold start, each
update one Checkerboard4 heatbath sweep then 4 Checkerboard4 overrelaxation sweeps as in
SU3_main.cpp:361-385) record the average plaquette and the clover charge Q after 4 stout
levels of rho = 0.12 (the metadynamics CV).  16 Philox chains run the same schedule on the
GPU.  After 50 updates of thermalisation the chain means are compared: <P> and <Q^2>
within 2 sigma, and the Q histogram (5 bins at the pooled quintiles) by a chi^2 of the
per-bin fractions below its 95 % point; sigma comes from the chain-to-chain spread on each
side (chains are independent; updates within a chain are not)."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FIXTURE = os.path.join(ROOT, "tests", "golden", "hbq16.npz")
SHAPE = (16, 16, 16, 16)
THERM = 50
GPU_SEEDS = range(2001, 2017)
CHI2_95_DOF5 = 11.0705


def flush(msg):
    print(msg, flush=True)


def chain_stats(P, Q, edges):
    """Per chain: mean plaquette, mean Q^2 and the fraction of updates in each Q bin."""
    P, Q = P[:, THERM:], Q[:, THERM:]
    idx = np.digitize(Q, edges)
    frac = np.stack([(idx == b).mean(axis=1) for b in range(len(edges) + 1)], axis=1)
    return P.mean(axis=1), (Q ** 2).mean(axis=1), frac


def diff_sigma(a, b):
    """Difference of the means of two sets of chain means and its standard error."""
    return a.mean(axis=0) - b.mean(axis=0), np.sqrt(a.var(axis=0, ddof=1) / len(a) + b.var(axis=0, ddof=1) / len(b))


def test_heatbath_plaquette_and_charge_distribution(gpu_ctx):
    import lettuce_amd as la
    if not os.path.exists(FIXTURE):
        pytest.skip("reference chains missing (python tests/golden/make_golden.py --hbq)")
    with np.load(FIXTURE, allow_pickle=False) as z:
        ref = z["PQ"]
        assert float(z["beta"]) == 5.7 and int(z["n_or"]) == 4 and int(z["n_stout"]) == 4
        rho = float(z["rho"])
    n_upd = ref.shape[1]
    f = la.GaugeField(gpu_ctx, SHAPE)
    g = la.GaugeField(gpu_ctx, SHAPE)
    P, Q = [], []
    for seed in GPU_SEEDS:
        flush(f"heatbath chain seed {seed}: {n_upd} updates")
        f.SetToIdentity()
        f.set_rng(seed, 0)
        hb, orx = la.HeatbathKernel(f, 5.7), la.OverrelaxationSubgroupKernel(f)
        p, q = [], []
        for _ in range(n_upd):
            la.Iterator.Checkerboard4(hb, 1)
            la.Iterator.Checkerboard4(orx, 4)
            p.append(la.PlaquetteSum(f) / (18.0 * f.volume))
            la.StoutSmearingN(f, g, 4, rho)
            q.append(la.TopChargeClover(g))
        P.append(p)
        Q.append(q)
    P, Q = np.array(P), np.array(Q)
    pooled = np.concatenate([Q[:, THERM:].ravel(), ref[:, THERM:, 1].ravel()])
    edges = np.quantile(pooled, [0.2, 0.4, 0.6, 0.8])
    gp, gq2, gf = chain_stats(P, Q, edges)
    rp, rq2, rf = chain_stats(ref[..., 0], ref[..., 1], edges)
    dp, sp = diff_sigma(gp, rp)
    dq2, sq2 = diff_sigma(gq2, rq2)
    df, sf = diff_sigma(gf, rf)
    chi2 = float(np.sum(df ** 2 / sf ** 2))
    flush(f"<P>   GPU {gp.mean():.6f}  reference {rp.mean():.6f}  diff {dp:+.2e} (sigma {sp:.1e})")
    flush(f"<Q^2> GPU {gq2.mean():.4f}  reference {rq2.mean():.4f}  diff {dq2:+.2e} (sigma {sq2:.1e})")
    flush(f"Q histogram fractions GPU {np.round(gf.mean(axis=0), 3)} reference {np.round(rf.mean(axis=0), 3)}"
          f" chi2 {chi2:.2f} (5 bins)")
    assert abs(dp) <= 2.0 * sp
    assert abs(dq2) <= 2.0 * sq2
    assert chi2 <= CHI2_95_DOF5
    f.close()
    g.close()
