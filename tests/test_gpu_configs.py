"""Every workload at its full size, against the oracle.

 - emission, 2^24 particles x 256 bins: the default method and the
   pair-by-pair kernel against the oracle's EXACT and FP64 spectra over all
   particles (1e-5 per bin, tests/_spectra.py);
 - energy histogram of the same 2^24 momenta (200 log bins, fourvel): counts
   bit for bit, sums to 1e-6;
 - 4096 cells, lognormal counts (median 4000, sigma 1, ~27M particles): the
   autotuned persistent launch replayed from its graph against 4096 calls of
   the single-population form (2e-6 per bin) and 16 cells against the oracle;
 - batched contraction, 65536 x 512 g bins -> 512 photon bins: 256 sampled
   rows of the split-FP16 (three products) and split-BF16 (six products)
   kernels, rows given or generated in the kernel, against the oracle's FP64
   FromDist at 1e-5 over every bin above 1e-12 of the row's peak;
 - differentiable fit, 2^26 particles reweighted to (p, c) = (-2, 1e3) with a
   target spectrum from (-2.2, 3e3): the loss gradient of a 4096-particle
   subset (all 256 bins) against the oracle's FP64 adjoint (itself checked
   against complex-step derivatives of the FP64 model), dL/dw to 1e-5 and
   dL/du, dL/dB to 1e-4 of the larger of the gradient's size and its size
   with |dL/dS|.
"""
import numpy as np
import pytest
import torch

from _adjoint import complex_step_grads
from _spectra import assert_spectrum
from ragnar_amd import torch_ops as T
from ragnar_amd import workloads

pytestmark = [pytest.mark.gpu, pytest.mark.slow]
DEV = "cuda"


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


@pytest.fixture(scope="module")
def table():
    return T.default_table()


@pytest.fixture(scope="module")
def emission_2p24():
    u, b = workloads.particles(1 << 24)
    return u, b


def test_emission_2p24_vs_oracle(orc, table, emission_2p24):
    u, b = emission_2p24
    bins = orc.logspace(1e-2, 1e7, 256)
    plan = T.SyncPlan(table, bins)
    U, B = dev(u), dev(b)
    S = {m: T.sync_spectrum(plan, U, B, method=m).cpu().numpy() for m in ("auto", "direct", "binned")}
    _, ex = orc.sync_spectrum(u, b, bins, mode=orc.EXACT, want_fp64=True)
    _, f64 = orc.sync_spectrum(u, b, bins, mode=orc.FP64, want_fp64=True)
    for m, s in S.items():
        assert_spectrum(s, ex, f64, what=f"2^24 x 256, {m}")


def test_histogram_2p24_vs_oracle(orc, emission_2p24):
    u, _ = emission_2p24
    bins = orc.logspace(1e-2, 1e3, 200)
    plan = T.HistPlan(bins, True)
    h, c = T.energy_hist(plan, dev(u), True)
    _, hd, cnt = orc.energy_hist(u, bins, True, True)
    assert np.array_equal(c.cpu().numpy().astype(np.uint64), cnt)
    assert int(cnt.sum()) == u.shape[1]
    h = h.double().cpu().numpy()
    live = hd != 0
    assert np.max(np.abs(h[live] / hd[live] - 1)) <= 1e-6
    assert np.all(h[~live] == 0)


def test_cells_4096_vs_single_calls_and_oracle(orc, table):
    u, b, offs, _, _ = workloads.cells(4096, 4000, 1.0)
    nb = 256
    bins = orc.logspace(1e-2, 1e7, nb)
    plan = T.SyncPlan(table, bins)
    U, B = dev(u), dev(b)
    out = torch.empty(4096, nb, device=DEV)
    T.sync_cells(plan, U, B, offs, out=out)       # tunes (first call of this shape) and captures
    S = T.sync_cells(plan, U, B, offs, out=out).clone()
    S2 = T.sync_cells(plan, U, B, offs, out=out)  # graph replay
    assert torch.equal(S, S2)
    single = torch.empty_like(S)
    for c in range(4096):
        a, z = int(offs[c]), int(offs[c + 1])
        T.sync_spectrum(plan, U[:, a:z].contiguous(), B[:, a:z].contiguous(), out=single[c])
    S, single = S.double().cpu().numpy(), single.double().cpu().numpy()
    live = single > 1e-12 * single.max(axis=1, keepdims=True)
    err = np.max(np.abs(S[live] / single[live] - 1))
    print(f"4096 cells ({int(offs[-1])} particles): graph vs single calls {err:.2e}")
    assert err <= 2e-6
    assert np.all(S[~live] == single[~live]) or np.max(np.abs(S[~live])) <= 1e-10 * S.max()
    for c in range(0, 4096, 256):
        a, z = int(offs[c]), int(offs[c + 1])
        _, ex = orc.sync_spectrum(u[:, a:z], b[:, a:z], bins, mode=orc.EXACT, want_fp64=True)
        _, f64 = orc.sync_spectrum(u[:, a:z], b[:, a:z], bins, mode=orc.FP64, want_fp64=True)
        assert_spectrum(S[c], ex, f64, what=f"cell {c} ({z - a} particles)")


@pytest.mark.parametrize("precision,generated", [(1, False), (3, False), (3, True)])
def test_contraction_65536_vs_fp64(orc, table, precision, generated):
    """Split FP16 (three products) and split BF16 (six products), rows given
    as values or generated from (index, cutoff) in the kernel."""
    nd, ng, nb = 65536, 512, 512
    f, p, gcut = workloads.distributions(nd, ng)
    g = orc.logspace(1, 1e6, ng)
    bins = orc.logspace(1e-2, 1e7, nb)
    plan = T.SyncPlan(table, bins)
    if generated:
        S = T.sync_fromdist_batched_plaw(plan, dev(g), dev(p.astype(np.float32)), dev(gcut.astype(np.float32)),
                                         True, precision)
    else:
        S = T.sync_fromdist_batched(plan, dev(g), dev(f), True, precision)
    S = S.cpu().numpy().astype(np.float64)
    gd = g.astype(np.float64)
    worst = 0.0
    for d in range(0, nd, nd // 256):
        fd = f[d]
        if generated:  # the row as the kernel defines it
            v = np.exp(np.float64(np.float32(p[d])) * np.log(gd) - gd / np.float64(np.float32(gcut[d])))
            fd = (v / np.trapezoid(v, gd)).astype(np.float32)
        _, f64 = orc.sync_fromdist(g, fd, True, bins, mode=orc.FP64, want_fp64=True)
        live = f64 > 1e-12 * f64.max()
        worst = max(worst, float(np.max(np.abs(S[d][live] / f64[live] - 1))))
        assert np.all(np.abs(S[d][~live]) <= 1e-10 * f64.max())
    print(f"precision {precision}{' generated' if generated else ''}, 65536 x 512 x 512: "
          f"max rel err vs FP64 over 256 rows {worst:.2e}")
    assert worst <= 1e-5


def test_fit_2p26_gradient_subset(orc, table):
    n, nb, sub = 1 << 26, 256, 4096
    u, b = workloads.particles(n, seed=321)
    bins = orc.logspace(1e-2, 1e7, nb)
    plan = T.SyncPlan(table, bins)
    U, B = dev(u), dev(b)
    w_cur = workloads.fit_weights(u, -2.0, 1e3)
    Wc = dev(w_cur)
    target = T.sync_spectrum(plan, U, B, dev(workloads.fit_weights(u, -2.2, 3e3))).double()
    S = T.sync_spectrum(plan, U, B, Wc).double()
    # the factorised forward at this size (items spanning several LDS windows
    # of runs) against the pair-by-pair kernel
    Sd = T.sync_spectrum(plan, U, B, Wc, method="direct").double()
    fl = Sd > 1e-12 * Sd.max()
    ferr = torch.max(torch.abs(S - Sd)[fl] / Sd[fl]).item()
    print(f"forward at 2^26, default vs pair by pair: {ferr:.2e}")
    assert ferr < 5e-6
    live = (S > 0) & (target > 0)
    r = torch.where(live, S.clamp_min(1e-300).log() - target.clamp_min(1e-300).log(), torch.zeros_like(S))
    gbar = torch.where(live, 2.0 * r / S.clamp_min(1e-300), torch.zeros_like(S)).float()
    gw, gu, gb, _ = T.sync_adjoint(plan, U, B, Wc, None, gbar, need=("w", "u", "b"))
    got = {"w": gw[:sub].double().cpu().numpy(), "u": gu[:, :sub].double().cpu().numpy(),
           "b": gb[:, :sub].double().cpu().numpy()}
    del U, B, gu, gb
    g64 = gbar.double().cpu().numpy()
    us, bs, ws = u[:, :sub], b[:, :sub], w_cur[:sub]
    rw, ru, rb, _, dn = orc.sync_adjoint_fp64(us, bs, bins, g64, w=ws)
    # scale of each gradient's summed terms: the same gradient with |gbar|
    aw, au, ab, _, _ = orc.sync_adjoint_fp64(us, bs, bins, np.abs(g64), w=ws)
    fx, fy = orc.ftable()
    cw, cu, cb, _ = complex_step_grads(us, None, bs, ws, bins, fx, fy, g64)
    near = dn < 1e-4
    print(f"fit gradient, 4096 of 2^26 particles: {near.sum()} within 1e-4 segment of a node")
    for name, ref, cs, ref_abs in (("w", rw[None], cw[None], aw[None]), ("u", ru, cu, au), ("b", rb, cb, ab)):
        g_ = got[name].reshape(ref.shape)
        norm = np.linalg.norm(ref, axis=0)
        err = np.linalg.norm(g_ - ref, axis=0) / np.maximum(norm, np.linalg.norm(ref_abs, axis=0))
        err_cs = (np.linalg.norm(ref - cs, axis=0) / np.maximum(norm, np.linalg.norm(ref_abs, axis=0)))[~near]
        print(f"  d{name}: max {err[~near].max():.2e} (near nodes {err[near].max() if near.any() else 0:.2e});"
              f" oracle vs complex step {err_cs.max():.1e}")
        assert err_cs.max() <= 1e-9
        assert err[~near].max() <= (1e-5 if name == "w" else 1e-4), name
        if near.any():
            assert err[near].max() <= 1e-2, name
