"""GPU parity of the hot path against the CPU oracle (through the C ABI).

Tolerances (written per test):
 - F table (reference recipe), e_peak / chiR, histogram counts: bit-exact;
 - spectra: 1e-5 per photon bin, both against the reference's float terms
   (oracle EXACT, net of the reference's own FP32 error) and against the FP64
   evaluation of the same model (tests/_spectra.py states the measure);
 - histogram sums: 1e-6 relative (FP64 sums of the same float terms).
"""
import numpy as np
import pytest
import torch

from _spectra import assert_spectrum, oracle_pair
from ragnar_amd import torch_ops as T
from ragnar_amd import workloads

pytestmark = pytest.mark.gpu

DEV = "cuda"


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


def rel_err(gpu, ref, floor=1e-12):
    gpu = np.asarray(gpu, np.float64)
    ref = np.asarray(ref, np.float64)
    live = np.abs(ref) > floor * np.abs(ref).max()
    return np.max(np.abs(gpu[live] - ref[live]) / np.abs(ref[live])), live


@pytest.fixture(scope="module")
def table():
    return T.default_table()


def test_reference_table_bit_exact(orc, table):
    x, y = table.nodes()
    xo, yo = orc.ftable()
    assert np.array_equal(x.view(np.int32), xo.view(np.int32))
    assert np.array_equal(y.view(np.int32), yo.view(np.int32))


def test_accurate_table_and_aux(orc):
    t = T.FTable(1, 200, 1e-6, 100.0)
    x, y = t.nodes()
    F = np.array([orc.F_fp64(v) for v in x])
    m = F > 1e-30
    assert np.max(np.abs(y[m] / F[m] - 1)) < 2e-7
    k, d = t.aux()
    k_ref = np.array([orc.k53(v) for v in x])
    d_ref = np.array([orc.dFdx_fp64(v) for v in x])
    mk = k_ref > 1e-300
    assert np.max(np.abs(k[mk] / k_ref[mk] - 1)) < 1e-12
    md = np.abs(d_ref) > 1e-300
    assert np.max(np.abs(d[md] - d_ref[md]) / np.abs(d_ref[md])) < 1e-9


def test_ffunc_integrand(orc):
    from ragnar_amd import ragnar as rg

    for x in (3e-6, 1e-5, 2e-4, 0.05, 0.29, 1.0, 7.5, 19.9, 20.0, 25.0):
        assert np.float32(rg.Ffunc_integrand(x)) == orc.ffunc_integrand(x), x


@pytest.mark.parametrize("seed,consts,with_e", [(1, (1.3, 2.0, 0.7), True), (2, (1.3, 2.0, 0.7), True),
                                                (3, (1.0, 1.0, 1.0), True), (4, (0.37, 3.1e-3, 41.0), True),
                                                (5, (1.0, 1.0, 1.0), False), (6, (0.37, 3.1e-3, 41.0), False)])
def test_particle_terms_bit_exact(orc, table, seed, consts, with_e):
    """e_peak and chiR of every particle equal the reference arithmetic bit for
    bit, over momenta from 1e-30 to 1e18 (including exact zeros and momenta
    whose squares overflow float), fields of both signs and several
    (B0, g_syn, e_at); with an electric field and without one (the kernels'
    shorter no-field form against the reference with E = 0)."""
    n = 4000
    rng = np.random.default_rng(seed)
    u, b = workloads.particles(n, seed=seed)
    scale = 10.0 ** rng.uniform(-30, 18, n)
    u[:, : n // 2] *= scale[: n // 2].astype(np.float32)
    u[:, 100:110] = 0.0
    u[0, 110:120] = -0.0
    e = rng.normal(0, 0.05, (3, n)).astype(np.float32)
    e[:, : n // 4] *= -3.0
    bins = orc.logspace(1e-2, 1e7, 256)
    if not with_e:
        e[:] = 0.0
    plan = T.SyncPlan(table, bins, *consts)
    terms = T.particle_terms(plan, dev(u), dev(b), None, dev(e) if with_e else None).cpu().numpy()
    for p in range(n):
        ep, chi = orc.omega_sync(u[:, p], e[:, p], b[:, p], *consts)
        assert np.array_equal(terms[p, 0], ep, equal_nan=True), (p, terms[p, 0], ep)
        assert np.array_equal(terms[p, 1], chi, equal_nan=True), (p, terms[p, 1], chi)


def test_particle_terms_electric_closed_forms(table):
    """chiR with E != 0 in closed forms (synchrotron.hpp:190-231), on the GPU:
    E parallel to u with B = 0 or B parallel to u gives |E| / (gamma B0); E
    perpendicular to u with B = 0 gives |E| / B0; crossed fields (u along x, E
    along y, B along z) give |Ey - beta Bz| / B0, zero for the force-free drift
    Ey = beta Bz.  Inputs are exact floats (integer momenta, so u^2 is exact);
    the float chiR is held to 1e-6 relative, the drift to 1e-6 |beta B|."""
    B0 = 0.7
    cases, refs = [], []
    for umag in (3.0, 40.0, 2000.0):
        g = np.sqrt(1.0 + umag * umag)
        beta = umag / g
        for em in (np.float32(0.2), np.float32(1.5)):
            em = float(em)
            cases.append(((0, 0, umag), (0, 0, -em), (0, 0, 0)))
            refs.append(em / (g * B0))
            cases.append(((0, 0, umag), (0, 0, em), (0, 0, 0.8)))
            refs.append(em / (g * B0))
            cases.append(((umag, 0, 0), (0, em, 0), (0, 0, 0)))
            refs.append(em / B0)
        bz = float(np.float32(1.3))
        for f in (-1.0, 0.5, 3.0, 1.0):
            ey = float(np.float32(f * beta * bz))
            cases.append(((umag, 0, 0), (0, ey, 0), (0, 0, bz)))
            refs.append(abs(ey - beta * bz) / B0)
    u = np.array([c[0] for c in cases], np.float32).T.copy()
    e = np.array([c[1] for c in cases], np.float32).T.copy()
    b = np.array([c[2] for c in cases], np.float32).T.copy()
    plan = T.SyncPlan(table, np.logspace(-2, 7, 256).astype(np.float32), B0, 1.0, 1.0)
    chi = T.particle_terms(plan, dev(u), dev(b), None, dev(e)).cpu().numpy()[:, 1].astype(np.float64)
    refs = np.array(refs)
    scale = np.linalg.norm(b, axis=0) * np.linalg.norm(u, axis=0) / np.sqrt(1 + (u * u).sum(0)) / B0
    drift = refs < 1e-6 * np.maximum(scale, 1e-30)
    assert drift.sum() == 3
    assert np.all(np.abs(chi[~drift] / refs[~drift] - 1) < 1e-6), (chi, refs)
    assert np.all(chi[drift] <= 1e-6 * scale[drift])


METHODS = ["direct", "binned"]


def _spectrum_case(orc, table, u, b, bins, e=None, w=None, B0=1.0, g=1.0, eat=1.0, method="direct"):
    plan = T.SyncPlan(table, bins, B0, g, eat)
    S = T.sync_spectrum(plan, dev(u), dev(b), None if w is None else dev(w),
                        None if e is None else dev(e), method=method).cpu().numpy()
    ex, f64 = oracle_pair(orc, u, b, bins, e=e, w=w, B0=B0, g_syn=g, e_at=eat)
    return S, ex, f64


@pytest.mark.parametrize("method", METHODS)
@pytest.mark.parametrize("n,nbins", [(1, 256), (1000, 200), (5000, 256), (777, 64), (3001, 65),
                                     (4096, 500), (2048, 1)])
def test_spectrum_vs_oracle(orc, table, n, nbins, method):
    u, b = workloads.particles(n, seed=n)
    bins = orc.logspace(1e-2, 1e7, nbins) if nbins > 1 else np.array([1.0], np.float32)
    if nbins > 512:
        pytest.skip("plans hold at most 512 bins; larger sets are split by the host layer")
    S, ex, f64 = _spectrum_case(orc, table, u, b, bins, method=method)
    assert_spectrum(S, ex, f64, what=f"{method} n={n} nbins={nbins}")


@pytest.mark.parametrize("method", METHODS)
def test_spectrum_fields_weights_scalars(orc, table, method):
    n = 4000
    u, b = workloads.particles(n, seed=9)
    rng = np.random.default_rng(9)
    e = rng.normal(0, 0.1, (3, n)).astype(np.float32)
    w = rng.uniform(0.1, 3.0, n).astype(np.float32)
    bins = orc.logspace(1e-3, 1e8, 300)
    S, ex, f64 = _spectrum_case(orc, table, u, b, bins, e=e, w=w, B0=1.7, g=3.0, eat=0.4,
                                method=method)
    assert_spectrum(S, ex, f64, what="E, w, B0, g_syn, e_at")


@pytest.mark.parametrize("method", METHODS)
def test_spectrum_edge_cases(orc, table, method):
    bins = np.concatenate([[0.0, -1.0], orc.logspace(1e-2, 1e6, 100)]).astype(np.float32)
    u = np.array([[0, 0, 40, 1e-30, 1e5], [0, 0, 0, 0, 0], [0, 30, 0, 0, 0]], np.float32)
    b = np.array([[0, 0, 0, 0, 0], [0, 0, 0.5, 0.5, 1], [1, 1, 1, 1, 0]], np.float32)
    S, ex, f64 = _spectrum_case(orc, table, u, b, bins, method=method)
    assert S[0] == 0 and S[1] == 0
    assert_spectrum(S, ex, f64, what="edge cases")
    # empty input
    plan = T.SyncPlan(table, orc.logspace(1e-2, 1e7, 256))
    z = torch.zeros(3, 0, device=DEV)
    assert torch.count_nonzero(T.sync_spectrum(plan, z, z, method=method)) == 0


@pytest.mark.parametrize("extra", [0, 100, 8191])
def test_binned_one_bucket_item_boundaries(orc, table, extra):
    """Every particle in one bucket: each 8192-particle chunk is one run, so
    the factorised form's work items (16384 sources) start and end exactly on
    run starts, with a partial last item or none."""
    n = 6 * 8192 + extra
    rng = np.random.default_rng(extra)
    u = np.zeros((3, n), np.float32)
    u[0] = 10.0
    b = np.zeros((3, n), np.float32)
    b[2] = 1.0
    w = rng.uniform(0.5, 2.0, n).astype(np.float32)
    bins = orc.logspace(1e-2, 1e7, 256)
    S, ex, f64 = _spectrum_case(orc, table, u, b, bins, w=w, method="binned")
    assert_spectrum(S, ex, f64, what=f"one bucket, n={n}")


@pytest.mark.parametrize("method", METHODS)
def test_spectrum_invariants_large(orc, table, method):
    n = 1 << 21
    u, b = workloads.particles(n)
    bins = orc.logspace(1e-2, 1e7, 256)
    plan = T.SyncPlan(table, bins)
    U, B = dev(u), dev(b)
    S = T.sync_spectrum(plan, U, B, method=method).double()
    S2 = T.sync_spectrum(plan, U, B, method=method).double()
    if method == "direct":
        assert torch.equal(S, S2)  # deterministic
    else:  # FP64 LDS atomics: order-dependent at the 1e-16 level
        assert torch.max(torch.abs(S - S2) / S.clamp_min(1e-300)) < 1e-12
    h = n // 3
    Sa = T.sync_spectrum(plan, U[:, :h].contiguous(), B[:, :h].contiguous(), method=method).double()
    Sb = T.sync_spectrum(plan, U[:, h:].contiguous(), B[:, h:].contiguous(), method=method).double()
    live = S > 1e-12 * S.max()
    assert torch.max(torch.abs(Sa + Sb - S)[live] / S[live]) < 2e-6
    w = torch.full((n,), 2.5, device=DEV)
    Sw = T.sync_spectrum(plan, U, B, w, method=method).double()
    assert torch.max(torch.abs(Sw - 2.5 * S)[live] / S[live]) < 1e-6
    # oracle on a subset
    sub = 1 << 16
    Ssub = T.sync_spectrum(plan, U[:, :sub].contiguous(), B[:, :sub].contiguous(),
                           method=method).cpu().numpy()
    ex, f64 = oracle_pair(orc, u[:, :sub], b[:, :sub], bins)
    assert_spectrum(Ssub, ex, f64, what="2^16 subset")


@pytest.mark.parametrize("log_bins,fourvel", [(True, True), (False, True), (True, False)])
def test_hist_vs_oracle(orc, log_bins, fourvel):
    u, _ = workloads.particles(200_003, seed=4)
    rng = np.random.default_rng(4)
    u[:, :1000] = rng.normal(0, 1e-3, (3, 1000))  # below range
    u[0, 1000:1010] = 0.0
    u[1:, 1000:1010] = 0.0                         # |u| = 0 -> 1/E = inf
    bins = orc.logspace(1e-2, 1e3, 200) if log_bins else orc.linspace(1e-2, 1e3, 200)
    plan = T.HistPlan(bins, log_bins)
    h, c = T.energy_hist(plan, dev(u), fourvel)
    h, c = h.cpu().numpy(), c.cpu().numpy()
    ho, hd, co = orc.energy_hist(u, bins, log_bins, fourvel)
    assert np.array_equal(c.astype(np.uint64), co)
    fin = np.isfinite(hd)
    assert np.array_equal(np.isfinite(h), fin)
    assert np.allclose(h[fin], hd[fin], rtol=1e-6, atol=0)


@pytest.mark.parametrize("nb", [1, 900, 2741])
def test_hist_bin_counts(orc, nb):
    """Histogram plans whose LDS holds 16, 8 or 4 private copies per block (and a
    single bin): counts bit-exact, sums 1e-6, against the oracle."""
    u, _ = workloads.particles(100_000, seed=nb)
    bins = orc.logspace(1e-2, 1e3, nb) if nb > 1 else np.array([1.0], np.float32)
    plan = T.HistPlan(bins, True)
    h, c = T.energy_hist(plan, dev(u), True)
    h, c = h.cpu().numpy(), c.cpu().numpy()
    ho, hd, co = orc.energy_hist(u, bins, True, True)
    assert np.array_equal(c.astype(np.uint64), co)
    assert int(c.sum()) == u.shape[1]
    fin = np.isfinite(hd)
    assert np.allclose(h[fin], hd[fin], rtol=1e-6, atol=0)


def test_hist_exact_edges(orc):
    bins = orc.logspace(1e-2, 1e3, 200)
    plan = T.HistPlan(bins, True)
    edges = plan.edges()
    # energies at, just below and just above every jump point (and every bin value)
    pts = []
    for v in list(edges[1:-1]) + list(bins.astype(np.float64)):
        f = np.float32(v)
        pts += [np.nextafter(f, np.float32(0)), f, np.nextafter(f, np.float32(1e30))]
    pts = np.array(pts, np.float32)
    u = np.zeros((3, len(pts)), np.float32)
    u[2] = pts
    _, c = T.energy_hist(plan, dev(u), True)
    _, _, co = orc.energy_hist(u, bins, True, True)
    assert np.array_equal(c.cpu().numpy().astype(np.uint64), co)


def test_fromdist_config1(orc, table):
    # TabulatedDistribution(Logbins(1,100,200), PlawGenerator(-2,1,100)) -> Logbins(0.01,1e7,200)
    g = orc.logspace(1, 100, 200)
    f = orc.generator("plaw", [-2.0, 1.0, 100.0], g)
    bins = orc.logspace(0.01, 1e7, 200)
    plan = T.SyncPlan(table, bins)
    S = T.sync_fromdist(plan, dev(g), dev(f), True).cpu().numpy()
    _, ex = orc.sync_fromdist(g, f, True, bins, want_fp64=True)
    _, f64 = orc.sync_fromdist(g, f, True, bins, mode=orc.FP64, want_fp64=True)
    assert_spectrum(S, ex, f64, what="config 1")
    # linear distribution bins (no gbeta factor)
    g2 = orc.linspace(1, 1000, 3000)
    f2 = orc.generator("plaw", [-2.5, 1.0, 1000.0], g2)
    S2 = T.sync_fromdist(plan, dev(g2), dev(f2), False).cpu().numpy()
    _, ex2 = orc.sync_fromdist(g2, f2, False, bins, want_fp64=True)
    _, f642 = orc.sync_fromdist(g2, f2, False, bins, mode=orc.FP64, want_fp64=True)
    assert_spectrum(S2, ex2, f642, what="linear distribution bins")


def test_methods_agree_full_size(orc, table):
    n = 1 << 24
    u, b = workloads.particles(n)
    bins = orc.logspace(1e-2, 1e7, 256)
    plan = T.SyncPlan(table, bins)
    U, B = dev(u), dev(b)
    Sd = T.sync_spectrum(plan, U, B, method="direct").double()
    Sb = T.sync_spectrum(plan, U, B, method="binned").double()
    live = Sd > 1e-12 * Sd.max()
    err = torch.max(torch.abs(Sb - Sd)[live] / Sd[live]).item()
    print("binned vs direct at 2^24 x 256:", err)
    assert err < 5e-6


@pytest.mark.parametrize("n", [1 << 21, 1 << 23])
def test_methods_agree_with_fields_and_weights(orc, table, n):
    """The factorised form's E-field bucket pass at scale: particles with E,
    weights and non-default scalars, against the pair-by-pair kernel, and a
    subset against the oracle.  2^23 particles (369 MB of inputs, beyond the
    256 MiB last-level cache) take the bucket pass's non-temporal reads, 2^21
    its plain ones."""
    u, b = workloads.particles(n, seed=77)
    rng = np.random.default_rng(77)
    e = rng.normal(0.0, 0.1, (3, n)).astype(np.float32)
    w = rng.uniform(0.1, 3.0, n).astype(np.float32)
    bins = orc.logspace(1e-3, 1e8, 300)
    plan = T.SyncPlan(table, bins, 1.7, 3.0, 0.4)
    U, B, E, W = dev(u), dev(b), dev(e), dev(w)
    Sd = T.sync_spectrum(plan, U, B, W, E, method="direct").double()
    Sb = T.sync_spectrum(plan, U, B, W, E, method="binned").double()
    live = Sd > 1e-12 * Sd.max()
    err = torch.max(torch.abs(Sb - Sd)[live] / Sd[live]).item()
    print(f"binned vs direct, E and w, {n} x 300:", err)
    assert err < 5e-6
    sub = 1 << 14
    Ssub = T.sync_spectrum(plan, U[:, :sub].contiguous(), B[:, :sub].contiguous(), W[:sub].contiguous(),
                           E[:, :sub].contiguous(), method="binned").cpu().numpy()
    ex, f64 = oracle_pair(orc, u[:, :sub], b[:, :sub], bins, e=e[:, :sub], w=w[:sub], B0=1.7, g_syn=3.0, e_at=0.4)
    assert_spectrum(Ssub, ex, f64, what="E, w, scalars, 2^14 subset")
