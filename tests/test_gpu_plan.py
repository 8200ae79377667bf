"""Replayable batch (s2g_plan_*): the whole batch captured once in a hipGraph
and replayed.  Each replay is checked against the eager s2g_deposit_batch on
the same inputs (the large-footprint part is the same computation; the
small-footprint additions may come in another order, hence 1e-12), with
eager work and synchronisations between replays, with the inputs changed in
place between replays, and against the FP64 oracle on small maps.  The
capacity path: inputs that outgrow the plan give S2G_ECAPACITY, not a fault."""
import numpy as np
import pytest

from conftest import rel_l1

pytestmark = pytest.mark.gpu


def _device_maps(s2g, nmaps=6, scale=1.0 / 400, seed_scale=1.0):
    import torch
    from sph2grid import workloads
    from sph2grid.mapping import _prepare
    dev = torch.device("cuda:0")
    out = []
    for d in workloads.batched(nmaps=nmaps, scale=scale):
        par = s2g.mappingParameters(x_lim=d["lims"][0], y_lim=d["lims"][1], z_lim=d["lims"][2],
                                    Npixels=d["Npixels"])
        pm = _prepare(d["Pos"].copy(), d["HSML"], d["M"], d["Rho"], d["Bin_Quant"],
                      d["Weights"], param=par, kernel=s2g.WendlandC6(2), show_progress=False,
                      reduce_image=True, calc_mean=True)
        t = [torch.from_numpy(np.ascontiguousarray(a)).to(dev) for a in pm.arrays]
        out.append(dict(t=t, pm=pm, npix=pm.npix, nq=pm.nq, host=[a.copy() for a in pm.arrays]))
    return out


def _descs(maps, key):
    import torch
    descs = []
    for m in maps:
        if key not in m:
            m[key] = torch.empty(m["npix"] ** 2 * m["nq"], dtype=torch.float64,
                                 device=m["t"][0].device)
        x, h, mm, rho, bq, w = m["t"]
        descs.append(dict(pos=x, hsml=h, m=mm, rho=rho, bin_q=bq, weights=w, out=m[key],
                          n=m["pm"].n, nq=m["nq"], npix=m["npix"],
                          len2pix=m["pm"].params.len2pix, kernel=m["pm"].kid,
                          calc_mean=True, reduce_image=True))
    return descs


def _max_rel(a, b):
    return float((a - b).abs().max().item()) / max(float(b.abs().max().item()), 1e-300)


def test_plan_replays_match_the_eager_batch(s2g, orc):
    import torch
    maps = _device_maps(s2g)
    eager, plan_d = _descs(maps, "o_eager"), _descs(maps, "o_plan")
    st = {}
    with s2g.BatchPlan(plan_d, dtype="f32") as plan:
        big = torch.randn(2048, 2048, device="cuda:0")
        for rep in range(4):
            if rep == 2:  # inputs changed in place: quantities doubled, hsml 10 % smaller
                for m in maps:
                    m["t"][4].mul_(2.0)
                    m["t"][1].mul_(0.9)
            s2g.deposit_batch_device(eager, dtype="f32")
            for _ in range(3):
                big = (big @ big).mul_(1e-3)  # eager work between replays
            torch.cuda.synchronize()
            plan.run(stats=st)
            torch.cuda.synchronize()
            errs = [_max_rel(m["o_plan"], m["o_eager"]) for m in maps]
            print(f"replay {rep}: max rel diff vs eager per map {['%.1e' % e for e in errs]}",
                  {k: st[k] for k in ("n_small", "n_large", "n_pairs")}, flush=True)
            assert max(errs) <= 1e-12, (rep, errs)
            assert st["n_small"] > 0 and st["n_large"] > 0 and st["n_pairs"] > 0
    # the two smallest maps of the last replay against the oracle
    for m in sorted(maps, key=lambda m: m["pm"].n)[:2]:
        x, h, mm, rho, bq, w = [a.double().cpu().numpy() for a in m["t"]]
        ref = orc.reduce_image_2D(orc.cic_mapping_2D(x.reshape(-1, 3), h, mm, rho,
                                                     bq.reshape(-1, m["nq"]), w, m["npix"],
                                                     m["pm"].params.len2pix, "wendlandc6", 2,
                                                     True), m["npix"], True)
        got = m["o_plan"].cpu().numpy().reshape(ref.shape, order="F")
        e = rel_l1(got, ref)
        print(f"map n={m['pm'].n} npix={m['npix']}: rel_l1 vs oracle {e:.2e}", flush=True)
        assert e <= 1e-5


def test_plan_capacity_error_and_rebuild(s2g):
    from sph2grid import _lib
    maps = _device_maps(s2g, nmaps=3, scale=1.0 / 800)
    plan = s2g.BatchPlan(_descs(maps, "o_plan"), dtype="f32", headroom=1.0)
    plan.run()
    for m in maps:
        m["t"][1].mul_(3.0)  # every footprint grows: more large particles and pairs
    with pytest.raises(_lib.LibraryError) as e:
        plan.run()
    print(e.value, flush=True)
    assert e.value.code == _lib.S2G_ECAPACITY and "capacit" in str(e.value)
    plan.close()
    # a new plan on the grown inputs matches the eager batch
    eager = _descs(maps, "o_eager")
    s2g.deposit_batch_device(eager, dtype="f32")
    with s2g.BatchPlan(_descs(maps, "o_plan"), dtype="f32") as plan:
        plan.run()
    for m in maps:
        assert _max_rel(m["o_plan"], m["o_eager"]) <= 1e-12


def test_plan_argument_errors(s2g):
    from sph2grid import _lib
    maps = _device_maps(s2g, nmaps=1, scale=1.0 / 800)
    for bad in (0.5, 100.0):
        with pytest.raises(_lib.LibraryError, match="headroom"):
            s2g.BatchPlan(_descs(maps, "o"), headroom=bad)
    with pytest.raises(_lib.LibraryError, match="at least one map"):
        s2g.BatchPlan([])


def test_plan_with_several_passes_and_dtypes(s2g, orc):
    """Maps of different kernels, quantity counts and accumulator flags make
    several passes in one graph; FP64 inputs; flat (unreduced) outputs."""
    import torch
    dev = torch.device("cuda:0")
    r = np.random.default_rng(31)
    specs = [("cubic", 1, 96, 900, False), ("wendlandc6", 3, 128, 1500, False),
             ("wendlandc4", 2, 80, 700, True), ("cubic", 1, 64, 400, False)]
    kid = {"cubic": 0, "wendlandc4": 3, "wendlandc6": 4}
    maps, ref = [], []
    for kname, nq, npix, n, fp32 in specs:
        x = r.uniform(-0.5, 0.5, (n, 3))
        h = np.exp(r.uniform(np.log(0.3 / npix), np.log(60.0 / npix), n))
        m, rho, w = (r.uniform(0.5, 2.0, n) for _ in range(3))
        q = r.uniform(0.1, 5.0, (n, nq))
        T = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(dev)
        d = dict(pos=T(x), hsml=T(h), m=T(m), rho=T(rho), bin_q=T(q), weights=T(w),
                 n=n, nq=nq, npix=npix, len2pix=float(npix), kernel=kid[kname],
                 calc_mean=True, fp32_accum=fp32)
        maps.append(d)
        ref.append(orc.cic_mapping_2D(x, h, m, rho, q, w, npix, float(npix), kname, 2, True))
    eager = [dict(d, out=torch.empty(d["npix"] ** 2 * (d["nq"] + 1), dtype=torch.float64,
                                     device=dev)) for d in maps]
    planned = [dict(d, out=torch.empty_like(e["out"])) for d, e in zip(maps, eager)]
    s2g.deposit_batch_device(eager, dtype="f64")
    with s2g.BatchPlan(planned, dtype="f64") as plan:
        for rep in range(2):
            plan.run()
            torch.cuda.synchronize()
            for k, (e, p, rf) in enumerate(zip(eager, planned, ref)):
                assert _max_rel(p["out"], e["out"]) <= 1e-12, (rep, k)
                got = p["out"].cpu().numpy().reshape((-1, specs[k][1] + 1), order="F")
                for q in range(got.shape[1]):
                    assert rel_l1(got[:, q], rf[:, q]) <= 1e-5, (rep, k, q)


def test_plan_item_budget_with_capped_items(s2g):
    """A plan built with tiny gather items (gather_item_max 64, one item per
    CU aimed at, so the items come from the cap): its item allocation is the
    budget pair_cap / 64, and a replay whose inputs grew within the plan's
    headroom (hsml 5 % larger: more pairs than at creation) must stay within
    it and match the eager batch run with the same split."""
    ctx = s2g.default_context()
    maps = _device_maps(s2g, nmaps=3, scale=1.0 / 800)
    try:
        ctx.set_tuning(dict(gather_item_min=8, gather_item_max=64, gather_items_per_cu=1))
        st = {}
        with s2g.BatchPlan(_descs(maps, "o_plan"), dtype="f32", headroom=1.5) as plan:
            plan.run(stats=st)
            pairs0 = st["n_pairs"]
            for m in maps:
                m["t"][1].mul_(1.05)
            eager = _descs(maps, "o_eager")
            for rep in range(2):
                s2g.deposit_batch_device(eager, dtype="f32")
                plan.run(stats=st)
                errs = [_max_rel(m["o_plan"], m["o_eager"]) for m in maps]
                print(f"replay {rep}: pairs {pairs0} -> {st['n_pairs']}, max rel diff "
                      f"{['%.1e' % e for e in errs]}", flush=True)
                assert st["n_pairs"] > pairs0
                assert max(errs) <= 1e-12
    finally:
        ctx.set_tuning(None)
