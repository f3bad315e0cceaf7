"""The request server (continuous batching) against solves run alone.

Bars: every served request's E_c equals the same system solved alone from
the MP2 guess to 1e-12 Ha with the same iteration count, its E(2) and E_HF
equal the standalone DrvMBPT / init_heg values; admission never holds more
solves than max_active nor more bytes than the budget at once; a request
larger than the budget is refused with UEGCCD_ERR_BUSY.
"""
import pytest

from tests.helpers import info_for
from uegccd_amd import lib as L

pytestmark = pytest.mark.gpu

MIX = [
    ("n14_cut5", {}, "fp64"),
    ("n14_cut5_tw", {}, "fp64"),
    ("n38_cut9", {}, "fp64"),
    ("n24_rs5", {}, "fp64"),
    ("n14_cut4", {"rs": 5.0}, "fp64"),
    ("n54_cut9", {}, "mixed"),
    ("n14_cut5", {"twist": (0.25, 0.0, -0.125)}, "fp64"),
]


def _alone(info, prec):
    from uegccd_amd.driver import CCDContext, init_heg, options
    heg = init_heg(info)
    c = CCDContext(heg, info, options(mode="diis", ndiis=8, precision=prec))
    e2, _, tab, tba = c.mp2()
    r = c.solve(tab, tba)
    c.close()
    return heg, e2, r


def _info(name, over):
    from uegccd_amd.inputfile import parse_input, standard_input
    from tests.helpers import SYSTEMS
    nel, cut, rs, tw = SYSTEMS[name]
    rs = over.get("rs", rs)
    tw = over.get("twist", tw)
    extra = {"fcut": 2.0} if name == "n24_rs5" else {}
    return parse_input(standard_input(nel, cut, rs, tw, **extra))


def _overlaps(results):
    """Largest number of solves, and of bytes, resident at one time."""
    ev = []
    for r in results:
        ev.append((r.t_start_ms, 1, r.footprint_bytes))
        ev.append((r.t_done_ms, -1, -r.footprint_bytes))
    ev.sort(key=lambda e: (e[0], e[1]))
    n = b = nmax = bmax = 0
    for _, dn, db in ev:
        n += dn
        b += db
        nmax, bmax = max(nmax, n), max(bmax, b)
    return nmax, bmax


@pytest.mark.parametrize("planners", [0, -1], ids=["planner_threads", "inline_planning"])
def test_served_equals_alone(planners):
    from uegccd_amd.driver import Server, options
    print(f"[server] alone solves (planners {planners})", flush=True)
    ref = []
    for name, over, prec in MIX:
        ref.append(_alone(_info(name, over), prec))
    print("[server] served mix", flush=True)
    srv = Server(max_active=4, iters_per_step=3, planner_threads=planners)
    ids = [srv.submit(_info(name, over), options(mode="diis", ndiis=8, precision=prec))
           for name, over, prec in MIX]
    srv.drain()
    out = [srv.poll(i) for i in ids]
    srv.close()
    for (name, over, prec), (heg, e2, r), s in zip(MIX, ref, out):
        assert s.state == L.REQ_DONE and s.status == 0, (name, s)
        assert s.niter == r.niter, (name, s.niter, r.niter)
        assert abs(s.ecorr - r.ecorr) < 1e-12, (name, s.ecorr, r.ecorr)
        assert abs(s.e2 - e2) < 1e-12
        assert abs(s.ehf - heg.EHF) < 1e-12
        assert s.nao == heg.NAO and s.nocc == heg.NOcc
        assert s.iters_fp32 == r.iters_fp32
        assert s.t_submit_ms <= s.t_start_ms <= s.t_done_ms
    nmax, _ = _overlaps(out)
    assert nmax <= 4


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_twist_stream_batch_equals_alone(precision):
    """Sixteen requests of the seeded twist stream (N 14/38/54, r_s 0.5-5,
    cutoffs 9/16/25) admitted together: each equals its solve alone.  Uneven
    sizes in one launch are what exposes any dependence of a solve's
    arithmetic on what else shares the kernels (LDS, tables, tiles)."""
    from uegccd_amd.driver import CCDContext, Server, init_heg, options
    from uegccd_amd.twist_stream import request_info, stream_requests
    reqs, _ = stream_requests(16, 1234)
    opts = options(mode="diis", ndiis=8, precision=precision)
    alone = []
    for q in reqs:
        info = request_info(q)
        c = CCDContext(init_heg(info), info, opts)
        _, _, tab, tba = c.mp2()
        alone.append(c.solve(tab, tba))
        c.close()
    print(f"[server] stream batch {precision}", flush=True)
    srv = Server(max_active=16, iters_per_step=4)
    ids = [srv.submit(request_info(q), opts) for q in reqs]
    srv.drain()
    out = [srv.poll(i) for i in ids]
    srv.close()
    for q, a, s in zip(reqs, alone, out):
        assert a.status == 0, q
        assert s.status == 0 and s.niter == a.niter, (q, s, a.niter)
        assert abs(s.ecorr - a.ecorr) < 1e-12, (q, s.ecorr, a.ecorr)


def test_admission_respects_budget_and_refuses_oversize():
    from uegccd_amd.driver import Server, footprint, init_heg, options
    opts = options(mode="diis", ndiis=8)
    small = [_info("n14_cut5", {"twist": (0.05 * k, 0.0, 0.0)}) for k in range(6)]
    big = _info("n38_cut9", {})
    fs = footprint(init_heg(small[0]), small[0], opts)
    fb = footprint(init_heg(big), big, opts)
    assert fb > 2 * fs
    budget = 2 * fs + fs // 2      # two small solves at a time, never the big one
    print(f"[server] budget {budget} small {fs} big {fb}", flush=True)
    srv = Server(hbm_budget_bytes=budget, max_active=16, iters_per_step=2)
    ids = [srv.submit(i, opts) for i in small]
    bid = srv.submit(big, opts)
    srv.drain()
    out = [srv.poll(i) for i in ids]
    rb = srv.poll(bid)
    srv.close()
    assert rb.state == L.REQ_DONE and rb.status == L.ERR_BUSY
    for s in out:
        assert s.state == L.REQ_DONE and s.status == 0
    nmax, bmax = _overlaps(out)
    assert bmax <= budget and nmax <= 2
    with pytest.raises(L.UegccdError) as e:
        Server().poll(12345)
    assert e.value.status == L.ERR_NOT_FOUND


def test_full_twist_stream_poisson():
    """The whole served twist-averaging workload: 256 requests of the seeded
    stream (seed 1234; N 14/38/54, r_s 0.5-5, cutoffs 9/16/25, twists in
    [-0.5, 0.5)^3) submitted to the server at their Poisson arrival times
    (700/s), mixed precision.  Every request converges, each served energy
    equals the same request solved alone in mixed precision (1e-12 Ha, same
    NIter), and each is within 1e-6 Ha of the request's plain FP64 solve."""
    import time
    from uegccd_amd.driver import CCDContext, Server, init_heg, options
    from uegccd_amd.twist_stream import request_info, stream_requests
    reqs, arrivals = stream_requests(256, 1234, 700.0)
    infos = [request_info(q) for q in reqs]
    omx = options(mode="diis", ndiis=8, precision="mixed")
    o64 = options(mode="diis", ndiis=8)
    alone_mx, alone_64 = [], []
    print("[stream256] alone solves", flush=True)
    for info in infos:
        heg = init_heg(info)
        for opts, out in ((omx, alone_mx), (o64, alone_64)):
            c = CCDContext(heg, info, opts)
            _, _, tab, tba = c.mp2()
            out.append(c.solve(tab, tba))
            c.close()
    print("[stream256] served at Poisson arrivals", flush=True)
    srv = Server(max_active=64, iters_per_step=4)
    ids, i, t0 = [], 0, time.perf_counter()
    while True:
        now = time.perf_counter() - t0
        while i < len(infos) and arrivals[i] <= now:
            ids.append(srv.submit(infos[i], omx))
            i += 1
        na, nq = srv.step()
        if i == len(infos) and na == 0 and nq == 0:
            break
        if na == 0 and nq == 0:
            time.sleep(max(0.0, min(arrivals[i] - (time.perf_counter() - t0), 0.001)))
    out = [srv.poll(k) for k in ids]
    srv.close()
    assert len(out) == 256
    overlap, _ = _overlaps(out)
    assert overlap > 1  # requests that arrived while others ran shared the device
    for q, mx, e64, s in zip(reqs, alone_mx, alone_64, out):
        assert mx.status == 0 and e64.status == 0, q
        assert s.state == L.REQ_DONE and s.status == 0, (q, s)
        assert s.niter == mx.niter and s.iters_fp32 == mx.iters_fp32, (q, s.niter, mx.niter)
        assert abs(s.ecorr - mx.ecorr) < 1e-12, (q, s.ecorr, mx.ecorr)
        assert abs(s.ecorr - e64.ecorr) < 1e-6, (q, s.ecorr, e64.ecorr)


def test_request_past_the_slice_table_is_refused():
    """A request whose basis the CCD passes do not take (N=2, |n|^2 <= 1050:
    a virtual-occupied |dn|^2 past the slice blocks' Coulomb table) finishes
    with UEGCCD_ERR_UNSUPPORTED from planning; a request beside it is served."""
    from uegccd_amd.driver import Server, options
    from uegccd_amd.inputfile import parse_input, standard_input
    print("[server] refused large basis", flush=True)
    opts = options(mode="diis", ndiis=8)
    srv = Server(max_active=4, iters_per_step=4)
    big = srv.submit(parse_input(standard_input(2, 1050, 1.0, fcut=10000.0)), opts)
    ok = srv.submit(_info("n14_cut5", {}), opts)
    srv.drain()
    rb, ro = srv.poll(big), srv.poll(ok)
    srv.close()
    assert rb.state == L.REQ_DONE and rb.status == L.ERR_UNSUPPORTED
    assert ro.state == L.REQ_DONE and ro.status == 0
