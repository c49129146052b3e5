"""GPU parity of the HIP propagator against the FP64 oracle (restatement of
core/seis_forward2.py).  Bars: seismograms 1e-5 relative L2, gradients 1e-4
relative L2, dot-product test 1e-6."""
import numpy as np
import pytest
import torch

from conftest import rel
from oracle import fd2d
from seismic_amd import synth
from seismic_amd.geometry import Geometry
from seismic_amd.propagator import Propagator, vel_to_seis

pytestmark = pytest.mark.gpu

SEIS_TOL = 1e-5
GRAD_TOL = 1e-4
DOT_TOL = 1e-6


def _np(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy().astype(np.float64)


def _case(g, style="curved", seed=0):
    v = synth.make_model(style, g.nz, g.nx, seed)
    return v, synth.to_vec(v)


@pytest.mark.parametrize("style", ["flat", "curved", "faulted", "grf"])
def test_forward_small(small_geoms, style):
    g, og = small_geoms
    _, vec = _case(g, style, seed=11)
    ref = fd2d.vel_to_seis(vec, og)[0]
    P = Propagator([g])
    out = _np(P.forward(torch.tensor(vec)))
    print(style, "forward rel", rel(out, ref))
    assert rel(out, ref) < SEIS_TOL


def test_tangent_small(small_geoms):
    g, og = small_geoms
    v, vec = _case(g, "curved", 5)
    rng = np.random.default_rng(0)
    dvec = np.concatenate([rng.normal(0, 10, g.nz * g.nx), [5.0]])
    s_ref, ds_ref, _ = fd2d.vel_to_seis(vec, og, vec_diff=dvec)
    P = Propagator([g])
    s, ds = P.tangent(torch.tensor(vec), torch.tensor(dvec))
    print("tangent rel", rel(_np(ds), ds_ref))
    assert rel(_np(s), s_ref) < SEIS_TOL
    assert rel(_np(ds), ds_ref) < SEIS_TOL


@pytest.mark.parametrize("ckpt", [0, 7])
def test_gradient_small(small_geoms, ckpt):
    g, og = small_geoms
    v, vec = _case(g, "faulted", 2)
    target = fd2d.vel_to_seis(synth.to_vec(v * 1.02), og)[0]
    s_ref, _, g_ref = fd2d.vel_to_seis(vec, og, vec_adjoint=target, adjoint_on_residual=True)
    P = Propagator([g], ckpt_every=ckpt)
    s, mis, grad = P.gradient(torch.tensor(vec), torch.tensor(target), on_residual=True)
    grad = _np(grad)
    print("grad rel", rel(grad, g_ref), "minvel", grad[-1], g_ref[-1])
    assert rel(_np(s), s_ref) < SEIS_TOL
    assert rel(grad, g_ref) < GRAD_TOL
    assert abs(grad[-1] - g_ref[-1]) / abs(g_ref[-1]) < GRAD_TOL
    mis_ref = np.sum((s_ref - target) ** 2)
    assert abs(float(mis[0]) - mis_ref) / mis_ref < 1e-4


def test_gradient_direct_source(small_geoms):
    """on_residual=False: grad = J^T dd for an arbitrary data vector."""
    g, og = small_geoms
    _, vec = _case(g, "flat", 4)
    rng = np.random.default_rng(1)
    dd = rng.normal(0, 1, og.n_shots * og.nt * og.n_rcv)
    _, _, g_ref = fd2d.vel_to_seis(vec, og, vec_adjoint=dd, adjoint_on_residual=False)
    P = Propagator([g])
    _, mis, grad = P.gradient(torch.tensor(vec), torch.tensor(dd), on_residual=False)
    assert mis is None
    assert rel(_np(grad), g_ref) < GRAD_TOL


def test_checkpoint_matches_full_history(small_geoms):
    g, og = small_geoms
    v, vec = _case(g, "grf", 9)
    target = fd2d.vel_to_seis(synth.to_vec(v + 40.0), og)[0]
    outs = []
    for ckpt in (0, 2, 13, 50, 299):
        P = Propagator([g], ckpt_every=ckpt)
        s, mis, grad = P.gradient(torch.tensor(vec), torch.tensor(target))
        outs.append((_np(s), _np(grad)))
    for s, gr in outs[1:]:
        assert np.array_equal(s, outs[0][0])
        assert np.array_equal(gr, outs[0][1])


CONFIGS = [(64, 32, 1, 1), (64, 32, 3, 0), (64, 32, 3, 1), (64, 64, 2, 0), (64, 64, 5, 1), (128, 32, 2, 1),
           (128, 32, 6, 0), (128, 32, 7, 1)]


def test_tile_configs_bit_identical(small_geoms):
    """Every window shape, temporal-blocking depth and row ownership gives
    bit-identical seismograms, tangents and gradients (checkpointed and full
    history)."""
    g, og = small_geoms
    v, vec = _case(g, "grf", 21)
    rng = np.random.default_rng(3)
    dv = np.concatenate([rng.normal(0, 10, g.nz * g.nx), [2.0]])
    target = fd2d.vel_to_seis(synth.to_vec(v * 1.01), og)[0]
    ref = None
    for ckpt in (0, 17):
        P = Propagator([g], ckpt_every=ckpt)
        for cfg in CONFIGS:
            for kind in ("forward", "record", "tangent", "adjoint"):
                P.set_config(kind, *cfg)
                assert P.get_config(kind) == cfg
            s, mis, gr = P.gradient(torch.tensor(vec), torch.tensor(target))
            s2, ds = P.tangent(torch.tensor(vec), torch.tensor(dv))
            out = (_np(s), _np(gr), _np(ds), float(mis[0]))
            if ckpt == 0:   # Born pass on the gradient's saved forward
                P.set_config("born", *cfg)
                assert np.array_equal(_np(P.tangent_saved(torch.tensor(vec), torch.tensor(dv))), out[2]), cfg
            if ref is None:
                ref = out
                continue
            assert np.array_equal(out[0], ref[0]), (ckpt, cfg)
            assert np.array_equal(out[1], ref[1]), (ckpt, cfg)
            assert np.array_equal(out[2], ref[2]), (ckpt, cfg)
            assert out[3] == ref[3]
    with pytest.raises(ValueError):
        P.set_config("forward", 64, 32, 8)   # interior would be empty
    with pytest.raises(ValueError):
        P.set_config("adjoint", 128, 128, 2)  # forward-only window
    # the 8-cell adjoint window
    for ckpt in (0, 17):
        P = Propagator([g], ckpt_every=ckpt)
        P.set_config("adjoint", 128, 64, 6, 1)
        s, mis, gr = P.gradient(torch.tensor(vec), torch.tensor(target))
        assert np.array_equal(_np(gr), ref[1]) and float(mis[0]) == ref[3]
    with pytest.raises(ValueError):
        P.set_config("born", 64, 32, 2, 2)   # no such row ownership


@pytest.mark.parametrize("cfg", [(128, 64, 3, 1), (128, 64, 8, 0), (256, 32, 2, 0), (256, 32, 5, 1),
                                 (128, 96, 6, 1), (128, 128, 8, 0), (128, 128, 5, 1)])
def test_large_forward_windows_bit_identical(small_geoms, cfg):
    """Forward-only windows (larger than the 52 x 56 padded grid, so tiles wrap
    several times) give the same seismograms, and the same gradients when they
    run the saving forward and the checkpoint recompute."""
    g, og = small_geoms
    v, vec = _case(g, "curved", 22)
    target = torch.tensor(fd2d.vel_to_seis(synth.to_vec(v * 1.01), og)[0])
    for ckpt in (0, 17):
        P = Propagator([g], ckpt_every=ckpt)
        P.set_stream_threshold(0)   # stream however few strips are active
        ref = _np(P.forward(torch.tensor(vec)))
        _, _, gref = P.gradient(torch.tensor(vec), target)
        P.set_config("forward", *cfg)
        P.set_config("record", *cfg)
        assert np.array_equal(_np(P.forward(torch.tensor(vec))), ref)
        _, _, gr = P.gradient(torch.tensor(vec), target)
        assert np.array_equal(_np(gr), _np(gref)), ckpt


STREAM_CONFIGS = [(128, 0, 4, 2), (128, 0, 3, 2), (192, 0, 4, 2), (256, 0, 3, 2), (128, 8, 4, 2), (64, 0, 3, 2), (64, 13, 4, 2),
                  (192, 23, 3, 2), (192, 0, 5, 2), (128, 11, 5, 2), (64, 0, 5, 2)]


@pytest.mark.parametrize("cfg", STREAM_CONFIGS)
def test_streamed_forward_bit_identical(small_geoms, cfg):
    """The row-streamed forward (column strips, one thread per column, whole
    columns or row segments) gives the seismograms, misfits and wavefields of
    the tiled kernel: plain forward, forward + misfit, and the checkpointing
    saving forward of a gradient (whose last launches and checkpoint-bounded
    launches run the tiled kernel)."""
    g, og = small_geoms
    v, vec = _case(g, "faulted", 23)
    target = torch.tensor(fd2d.vel_to_seis(synth.to_vec(v * 1.01), og)[0])
    for ckpt in (0, 17):
        P = Propagator([g], ckpt_every=ckpt)
        P.set_stream_threshold(0)   # stream however few strips are active
        ref = _np(P.forward(torch.tensor(vec)))
        mref = float(P.forward_misfit(torch.tensor(vec), target)[0])
        _, _, gref = P.gradient(torch.tensor(vec), target)
        P.set_config("forward", *cfg)
        assert P.get_config("forward") == cfg
        out = _np(P.forward(torch.tensor(vec)))
        print(cfg, ckpt, "rel to tiled", rel(out, ref), flush=True)
        assert np.array_equal(out, ref), (cfg, ckpt)
        assert float(P.forward_misfit(torch.tensor(vec), target)[0]) == mref
        _, _, gr = P.gradient(torch.tensor(vec), target)
        assert np.array_equal(_np(gr), _np(gref)), (cfg, ckpt)
    ref_o = fd2d.vel_to_seis(vec, og)[0]
    assert rel(out, ref_o) < SEIS_TOL


def test_streamed_forward_ragged_batch():
    """Mixed grid shapes, shot counts and source / receiver rows (incl. the
    surface row) in one streamed launch, against the tiled kernel; the
    streamed config is refused for padded grids under 4k rows."""
    kw = dict(nbc=8, nt=170)
    geoms = [Geometry(nz=12, nx=16, src_x=(7,), src_z=0, rcv_z=3, rcv_x=(2, 5, 8, 11, 14), **kw),
             Geometry(nz=20, nx=10, src_x=(0, 4, 9), src_z=2, rcv_z=0, rcv_x=(0, 3, 6, 9), **kw),
             Geometry(nz=8, nx=30, n_shots=7, src_z=5, rcv_z=7, **kw),
             Geometry(nz=40, nx=170, n_shots=3, **kw)]
    rng = np.random.default_rng(4)
    vec = np.concatenate([synth.to_vec(np.linspace(1800, 3600, gg.nz)[:, None] * np.ones((1, gg.nx))
                                       + rng.uniform(-100, 100, (gg.nz, gg.nx))) for gg in geoms])
    P = Propagator(geoms, ckpt_every=0)
    P.set_stream_threshold(0)   # stream however few strips are active
    ref = _np(P.forward(torch.tensor(vec)))
    for cfg in STREAM_CONFIGS:
        P.set_config("forward", *cfg)
        assert np.array_equal(_np(P.forward(torch.tensor(vec))), ref), cfg
    Ps = Propagator([Geometry(nz=2, nx=5, nbc=6, nt=150)])
    Ps.set_stream_threshold(0)   # stream however few strips are active
    Ps.set_config("forward", 128, 0, 3, 2)      # 14 padded rows >= 12
    with pytest.raises(ValueError):
        Ps.set_config("forward", 128, 0, 4, 2)  # 14 < 16
    with pytest.raises(ValueError):
        P.set_config("forward", 128, 0, 6, 2)   # streamed forward: 3 to 5 levels
    with pytest.raises(ValueError):
        P.set_config("record", 128, 0, 5, 2)    # streamed record: 3 or 4 levels
    with pytest.raises(ValueError):
        P.set_config("born", 128, 0, 5, 2)      # streamed Born: 3 or 4 levels
    with pytest.raises(ValueError):
        P.set_config("adjoint", 128, 0, 5, 2)   # streamed adjoint: 4 or 6 levels
    with pytest.raises(ValueError):
        P.set_config("adjoint", 192, 0, 4, 2)   # streamed adjoint: 64- or 128-column strips


TAN_STREAM_CONFIGS = [(128, 0, 3, 2), (192, 0, 4, 2), (128, 9, 4, 2), (192, 13, 3, 2), (128, 160, 4, 2)]


@pytest.mark.parametrize("cfg", TAN_STREAM_CONFIGS)
def test_streamed_tangent_and_record_bit_identical(small_geoms, cfg):
    """The row-streamed fused forward + tangent (tangent kind) and saving
    forward (record kind) give the tiled kernels' seismograms, tangents,
    gradients and GN products: the drop-in tangent pass, the GN product that
    records the history on a full-history plan, the Born pass on a history
    the streamed saving forward wrote, and on a plan with checkpoints every 17
    steps the GN product and the gradient whose segment recomputes stream
    (launches shorter than k run the tiled tail window)."""
    g, og = small_geoms
    v, vec = _case(g, "faulted", 25)
    target = torch.tensor(fd2d.vel_to_seis(synth.to_vec(v * 1.01), og)[0])
    rng = np.random.default_rng(12)
    dm = torch.tensor(np.concatenate([rng.normal(0, 20, g.nz * g.nx), [3.0]]))
    v = torch.tensor(vec)
    for ckpt in (0, 17):
        P = Propagator([g], ckpt_every=ckpt)
        P.set_stream_threshold(0)   # stream however few strips are active
        s_ref, ds_ref = [_np(x) for x in P.tangent(v, dm)]
        gp_ref = [_np(x) for x in P.gn_product(v, dm)]
        _, m_ref, g_ref = P.gradient(v, target)
        g_ref, m_ref = _np(g_ref), float(m_ref[0])
        b_ref = _np(P.tangent_saved(v, dm)) if ckpt == 0 else None
        P.set_config("tangent", *cfg)
        P.set_config("record", *cfg)
        assert P.get_config("tangent") == cfg and P.get_config("record") == cfg
        s_out, ds_out = [_np(x) for x in P.tangent(v, dm)]
        print(cfg, ckpt, "tangent rel to tiled", rel(ds_out, ds_ref), flush=True)
        assert np.array_equal(s_out, s_ref) and np.array_equal(ds_out, ds_ref), (cfg, ckpt)
        gp = [_np(x) for x in P.gn_product(v, dm)]
        assert np.array_equal(gp[0], gp_ref[0]) and np.array_equal(gp[1], gp_ref[1]), (cfg, ckpt)
        _, mis, gr = P.gradient(v, target)
        assert np.array_equal(_np(gr), g_ref) and float(mis[0]) == m_ref, (cfg, ckpt)
        if ckpt == 0:   # Born on the history the streamed saving forward wrote
            assert np.array_equal(_np(P.tangent_saved(v, dm)), b_ref), cfg
    _, ds_o, _ = fd2d.vel_to_seis(vec, og, vec_diff=dm.numpy())
    assert rel(ds_out, ds_o) < SEIS_TOL


def test_streamed_tangent_ragged_batch():
    """Streamed fused tangent and saving forward over mixed grids, shot counts
    and source / receiver rows (incl. the surface row) equal the tiled
    kernels, also with models inactive."""
    kw = dict(nbc=8, nt=170)
    geoms = [Geometry(nz=12, nx=16, src_x=(7,), src_z=0, rcv_z=3, rcv_x=(2, 5, 8, 11, 14), **kw),
             Geometry(nz=20, nx=10, src_x=(0, 4, 9), src_z=2, rcv_z=0, rcv_x=(0, 3, 6, 9), **kw),
             Geometry(nz=8, nx=30, n_shots=7, src_z=5, rcv_z=7, **kw),
             Geometry(nz=40, nx=170, n_shots=3, **kw)]
    rng = np.random.default_rng(13)
    vs = [np.linspace(1800, 3600, gg.nz)[:, None] * np.ones((1, gg.nx)) + rng.uniform(-100, 100, (gg.nz, gg.nx))
          for gg in geoms]
    vec = torch.tensor(np.concatenate([synth.to_vec(v) for v in vs]))
    dv = torch.tensor(np.concatenate([np.concatenate([rng.normal(0, 15, gg.nz * gg.nx), [2.0]]) for gg in geoms]))
    P = Propagator(geoms, ckpt_every=0)
    P.set_stream_threshold(0)   # stream however few strips are active
    dd = P.forward(vec * 1.01).clone()
    ref_t = _np(P.tangent(vec, dv)[1])
    ref_g = _np(P.gradient(vec, dd)[2])
    for cfg in [(128, 0, 3, 2), (192, 8, 4, 2), (128, 11, 4, 2)]:
        P.set_config("tangent", *cfg)
        P.set_config("record", *cfg)
        assert np.array_equal(_np(P.tangent(vec, dv)[1]), ref_t), cfg
        assert np.array_equal(_np(P.gradient(vec, dd)[2]), ref_g), cfg
    P.set_active([True, False, False, True])
    out = P.tangent(vec, dv)[1]
    for i in (0, 3):
        a, b = P.seis_off[i], P.seis_off[i + 1]
        assert np.array_equal(_np(out[a:b]), ref_t[a:b])
    with pytest.raises(ValueError):
        P.set_config("tangent", 64, 0, 3, 2)    # streamed tangent and record: 128- or 192-column strips
    with pytest.raises(ValueError):
        P.set_config("record", 256, 0, 4, 2)


def test_streamed_passes_fall_back_with_few_active(small_geoms):
    """With streamed configs on every kind and the default threshold (one
    active strip segment per compute unit), a plan of 8 small models streams
    while all are active (280 segments) and runs the tiled tail window when
    one is (35); gradients, tangents and Born passes equal a tiled plan's in
    both regimes as the mask switches back and forth."""
    g, _ = small_geoms
    B = 8
    v = torch.tensor(np.concatenate([synth.to_vec(synth.make_model("curved", g.nz, g.nx, 30 + i))
                                     for i in range(B)]))
    rng = np.random.default_rng(14)
    dm = torch.tensor(np.concatenate([np.concatenate([rng.normal(0, 20, g.nz * g.nx), [3.0]]) for _ in range(B)]))
    Pt = Propagator([g] * B)
    dd = Pt.forward(v * 1.01).clone()
    ref_g = _np(Pt.gradient(v, dd)[2])
    ref_b = _np(Pt.tangent_saved(v, dm))
    ref_t = _np(Pt.tangent(v, dm)[1])
    P = Propagator([g] * B)
    for kind in ("forward", "record", "tangent", "born", "adjoint"):
        P.set_config(kind, 128, 8, 4, 2)
    for mask in ([True] * B, [True] + [False] * (B - 1), [True] * B, [False] * (B - 1) + [True]):
        P.set_active(mask)
        gr = _np(P.gradient(v, dd)[2])
        b = _np(P.tangent_saved(v, dm))
        t = _np(P.tangent(v, dm)[1])
        for i in range(B):
            if not mask[i]:
                continue
            vs, ss = slice(P.vec_off[i], P.vec_off[i + 1]), slice(P.seis_off[i], P.seis_off[i + 1])
            assert np.array_equal(gr[vs], ref_g[vs]), (mask, i)
            assert np.array_equal(b[ss], ref_b[ss]) and np.array_equal(t[ss], ref_t[ss]), (mask, i)


ADJ_STREAM_CONFIGS = [(128, 0, 4, 2), (64, 0, 4, 2), (128, 9, 4, 2), (64, 13, 6, 2), (128, 0, 6, 2)]


@pytest.mark.parametrize("cfg", ADJ_STREAM_CONFIGS)
def test_streamed_adjoint_bit_identical(small_geoms, cfg):
    """The row-streamed adjoint gives the gradient (imaging terms, source
    correlation, min_vel component) and misfit of the tiled adjoint, with the
    full history and with checkpoints every 17 steps (segments whose last
    launches are shorter than k run the tiled kernel), and the GN product on
    the saved forward; the gradient matches the oracle."""
    g, og = small_geoms
    v, vec = _case(g, "grf", 24)
    target = fd2d.vel_to_seis(synth.to_vec(v * 1.01), og)[0]
    rng = np.random.default_rng(8)
    dm = torch.tensor(np.concatenate([rng.normal(0, 20, g.nz * g.nx), [3.0]]))
    for ckpt in (0, 17):
        P = Propagator([g], ckpt_every=ckpt)
        P.set_stream_threshold(0)   # stream however few strips are active
        _, mref, gref = P.gradient(torch.tensor(vec), torch.tensor(target))
        gref, mref = _np(gref), float(mref[0])
        if ckpt == 0:
            hv_ref = _np(P.gn_product_saved(torch.tensor(vec), dm)[-1])
        P.set_config("adjoint", *cfg)
        assert P.get_config("adjoint") == cfg
        _, mis, gr = P.gradient(torch.tensor(vec), torch.tensor(target))
        print(cfg, ckpt, "rel to tiled", rel(_np(gr), gref), flush=True)
        assert np.array_equal(_np(gr), gref), (cfg, ckpt)
        assert float(mis[0]) == mref
        if ckpt == 0:
            assert np.array_equal(_np(P.gn_product_saved(torch.tensor(vec), dm)[-1]), hv_ref), cfg
    _, g_o = fd2d.vel_to_seis(vec, og, vec_adjoint=target, adjoint_on_residual=True)[::2]
    assert rel(_np(gr), g_o) < GRAD_TOL


def test_streamed_adjoint_ragged_batch():
    """Streamed adjoint over mixed grids, shot counts and source / receiver
    rows (incl. the surface row) equals the tiled adjoint, full history and
    checkpointed, also with models inactive."""
    kw = dict(nbc=8, nt=170)
    geoms = [Geometry(nz=12, nx=16, src_x=(7,), src_z=0, rcv_z=3, rcv_x=(2, 5, 8, 11, 14), **kw),
             Geometry(nz=20, nx=10, src_x=(0, 4, 9), src_z=2, rcv_z=0, rcv_x=(0, 3, 6, 9), **kw),
             Geometry(nz=8, nx=30, n_shots=7, src_z=5, rcv_z=7, **kw),
             Geometry(nz=40, nx=170, n_shots=3, **kw)]
    rng = np.random.default_rng(9)
    vs = [np.linspace(1800, 3600, gg.nz)[:, None] * np.ones((1, gg.nx)) + rng.uniform(-100, 100, (gg.nz, gg.nx))
          for gg in geoms]
    vec = torch.tensor(np.concatenate([synth.to_vec(v) for v in vs]))
    for ckpt in (0, 13):
        P = Propagator(geoms, ckpt_every=ckpt)
        P.set_stream_threshold(0)   # stream however few strips are active
        dd = P.forward(vec * 1.01).clone()
        _, mref, gref = P.gradient(vec, dd)
        gref, mref = _np(gref), _np(mref)
        for cfg in [(128, 0, 4, 2), (64, 8, 4, 2), (128, 11, 6, 2)]:
            P.set_config("adjoint", *cfg)
            _, mis, gr = P.gradient(vec, dd)
            assert np.array_equal(_np(gr), gref), (cfg, ckpt)
            assert np.array_equal(_np(mis), mref), (cfg, ckpt)
        P.set_active([False, True, False, True])
        _, mis, gr = P.gradient(vec, dd)
        for i in (1, 3):
            a, b = P.vec_off[i], P.vec_off[i + 1]
            assert np.array_equal(_np(gr[a:b]), gref[a:b]), (i, ckpt)


def test_streamed_born_ragged_batch():
    """Streamed Born over mixed grids, shot counts and source / receiver rows
    equals the tiled Born (and the fused tangent), also with models inactive."""
    kw = dict(nbc=8, nt=170)
    geoms = [Geometry(nz=12, nx=16, src_x=(7,), src_z=0, rcv_z=3, rcv_x=(2, 5, 8, 11, 14), **kw),
             Geometry(nz=20, nx=10, src_x=(0, 4, 9), src_z=2, rcv_z=0, rcv_x=(0, 3, 6, 9), **kw),
             Geometry(nz=8, nx=30, n_shots=7, src_z=5, rcv_z=7, **kw),
             Geometry(nz=40, nx=170, n_shots=3, **kw)]
    rng = np.random.default_rng(6)
    vec = np.concatenate([synth.to_vec(np.linspace(1800, 3600, gg.nz)[:, None] * np.ones((1, gg.nx))
                                       + rng.uniform(-100, 100, (gg.nz, gg.nx))) for gg in geoms])
    dv = np.concatenate([np.concatenate([rng.normal(0, 15, gg.nz * gg.nx), [2.0]]) for gg in geoms])
    P = Propagator(geoms, ckpt_every=0)
    P.set_stream_threshold(0)   # stream however few strips are active
    _, ref = P.tangent(torch.tensor(vec), torch.tensor(dv))
    ref = _np(ref)
    P.forward_save(torch.tensor(vec))
    for cfg in [(128, 0, 3, 2), (192, 17, 3, 2), (256, 0, 3, 2), (64, 0, 3, 2), (64, 30, 4, 2), (128, 0, 4, 2)]:
        P.set_config("born", *cfg)
        out = _np(P.tangent_saved(torch.tensor(vec), torch.tensor(dv)))
        assert np.array_equal(out, ref), cfg
    P.set_active([True, False, True, False])
    out = P.tangent_saved(torch.tensor(vec), torch.tensor(dv))   # inactive models' slices are not written
    for i in (0, 2):
        a, b = P.seis_off[i], P.seis_off[i + 1]
        assert np.array_equal(_np(out[a:b]), ref[a:b])


@pytest.mark.parametrize("cfg", [(64, 32, 1, 1), (128, 64, 4, 0), (256, 32, 3, 1), (64, 64, 8, 0),
                                 (128, 0, 3, 2), (192, 0, 3, 2), (256, 0, 3, 2), (128, 20, 3, 2), (64, 0, 3, 2), (64, 11, 4, 2),
                                 (192, 9, 3, 2)])
def test_born_saved_matches_fused(small_geoms, cfg):
    """Tangent and GN product from the saved forward are bit-identical to the
    passes that recompute the forward, and match the oracle."""
    g, og = small_geoms
    v, vec = _case(g, "curved", 4)
    rng = np.random.default_rng(5)
    dm = np.concatenate([rng.normal(0, 20, g.nz * g.nx), [3.0]])
    P = Propagator([g])
    P.set_stream_threshold(0)   # stream however few strips are active
    P.set_config("born", *cfg)
    _, ds_fused = P.tangent(torch.tensor(vec), torch.tensor(dm))
    dd = torch.tensor(fd2d.vel_to_seis(synth.to_vec(v * 0.99), og)[0])
    P.gradient(torch.tensor(vec), dd)
    ds_saved = P.tangent_saved(torch.tensor(vec), torch.tensor(dm))
    assert np.array_equal(_np(ds_saved), _np(ds_fused))
    d1, g1 = P.gn_product(torch.tensor(vec), torch.tensor(dm))
    d2, g2 = P.gn_product_saved(torch.tensor(vec), torch.tensor(dm))
    assert np.array_equal(_np(d1), _np(d2))
    assert np.array_equal(_np(g1), _np(g2))
    _, jdm, _ = fd2d.vel_to_seis(vec, og, vec_diff=dm)
    _, _, jtjdm = fd2d.vel_to_seis(vec, og, vec_adjoint=jdm)
    assert rel(_np(d2), jdm) < SEIS_TOL
    assert rel(_np(g2), jtjdm) < GRAD_TOL


def test_saved_passes_need_a_saved_forward(small_geoms):
    g, og = small_geoms
    _, vec = _case(g, "flat", 1)
    v = torch.tensor(vec)
    P = Propagator([g])
    with pytest.raises(RuntimeError):      # no saving forward yet (library check)
        P.tangent_saved(v, v, check=False)
    P.forward(v)                           # a plain forward saves nothing
    with pytest.raises(RuntimeError):
        P.gn_product_saved(v, v, check=False)
    P.forward_save(v)
    assert P.saved_matches(v)
    assert not P.saved_matches(v * 1.01)
    with pytest.raises(RuntimeError):      # a different point (host check)
        P.tangent_saved(v * 1.01, v)
    P.tangent_saved(v, v)
    with pytest.raises(RuntimeError):      # checkpointed plans keep no full history
        Propagator([g], ckpt_every=11).tangent_saved(v, v, check=False)
    # batch: a saving forward on model 0 alone leaves model 1 without history
    Pb = Propagator([g, g])
    vb = torch.cat([v, v])
    Pb.set_active([True, False])
    Pb.forward_save(vb)
    assert Pb.saved_matches(vb)
    Pb.set_active([True, True])
    assert not Pb.saved_matches(vb)
    with pytest.raises(RuntimeError):
        Pb.tangent_saved(vb, vb, check=False)
    Pb.set_active([False, True])
    Pb.gradient(vb, torch.zeros(Pb.n_seis))
    Pb.set_active([True, True])
    assert Pb.saved_matches(vb)
    ds = Pb.tangent_saved(vb, vb)
    _, ds_ref = Pb.tangent(vb, vb)
    assert np.array_equal(_np(ds), _np(ds_ref))


def test_dot_product(small_geoms):
    """<J dm, r> = <dm, J^T r> on the float32 GPU path, r a model-misfit residual."""
    g, og = small_geoms
    v, vec = _case(g, "curved", 3)
    rng = np.random.default_rng(7)
    dm = np.concatenate([rng.normal(0, 20, g.nz * g.nx), [3.0]])
    dd = fd2d.vel_to_seis(synth.to_vec(v * 0.98), og)[0]
    P = Propagator([g])
    s, jdm = P.tangent(torch.tensor(vec), torch.tensor(dm))
    _, _, jt = P.gradient(torch.tensor(vec), torch.tensor(dd), on_residual=True)
    r = _np(s) - dd
    a = float(np.dot(r, _np(jdm)))
    b = float(np.dot(_np(jt), dm))
    print("dot", a, b, abs(a - b) / abs(a))
    assert abs(a - b) / abs(a) < DOT_TOL


def test_gn_product(small_geoms):
    g, og = small_geoms
    v, vec = _case(g, "curved", 8)
    rng = np.random.default_rng(2)
    dv = np.concatenate([rng.normal(0, 10, g.nz * g.nx), [1.0]])
    _, jdv, _ = fd2d.vel_to_seis(vec, og, vec_diff=dv)
    _, _, jtjdv = fd2d.vel_to_seis(vec, og, vec_adjoint=jdv)
    for ckpt in (0, 11):
        P = Propagator([g], ckpt_every=ckpt)
        ds, gr = P.gn_product(torch.tensor(vec), torch.tensor(dv))
        assert rel(_np(ds), jdv) < SEIS_TOL
        assert rel(_np(gr), jtjdv) < GRAD_TOL


def test_batch_invariance_mixed_sizes():
    """Each model's outputs are bit-identical alone and inside a mixed batch."""
    kw = dict(nbc=12, nt=200)
    geoms = [Geometry(nz=20, nx=24, **kw), Geometry(nz=30, nx=30, **kw), Geometry(nz=16, nx=40, **kw)]
    vecs = [synth.to_vec(synth.make_model(s, gg.nz, gg.nx, i))
            for i, (s, gg) in enumerate(zip(["flat", "grf", "faulted"], geoms))]
    targets = [synth.to_vec(synth.make_model("flat", gg.nz, gg.nx, 100 + i)) for i, gg in enumerate(geoms)]
    Pb = Propagator(geoms, ckpt_every=9)
    dd = torch.cat([Propagator([gg]).forward(torch.tensor(t)) for gg, t in zip(geoms, targets)])
    sb, mb, gb = Pb.gradient(torch.tensor(np.concatenate(vecs)), dd)
    sb, gb = _np(sb), _np(gb)
    for i, gg in enumerate(geoms):
        P1 = Propagator([gg], ckpt_every=9)
        d1 = dd[int(Pb.seis_off[i]):int(Pb.seis_off[i + 1])]
        s1, m1, g1 = P1.gradient(torch.tensor(vecs[i]), d1)
        assert np.array_equal(_np(s1), sb[Pb.seis_off[i]:Pb.seis_off[i + 1]])
        assert np.array_equal(_np(g1), gb[Pb.vec_off[i]:Pb.vec_off[i + 1]])
        assert float(m1[0]) == float(mb[i])


def test_inactive_models_untouched():
    kw = dict(nbc=12, nt=160)
    geoms = [Geometry(nz=20, nx=20, **kw), Geometry(nz=24, nx=20, **kw)]
    vecs = np.concatenate([synth.to_vec(synth.make_model("flat", gg.nz, gg.nx, i)) for i, gg in enumerate(geoms)])
    P = Propagator(geoms)
    full = _np(P.forward(torch.tensor(vecs)))
    P.set_active([False, True])
    out = torch.full((P.n_seis,), 7.0, device="cuda")
    P.forward(torch.tensor(vecs), out=out)
    out = _np(out)
    assert np.all(out[:P.seis_off[1]] == 7.0)
    assert np.array_equal(out[P.seis_off[1]:], full[P.seis_off[1]:])


@pytest.mark.slow
def test_forward_full_size_reference_geometry():
    """Reference geometry (70x70, nbc=120, nt=999, 5 shots): shots 0 and 2."""
    g = Geometry()
    og = fd2d.Geometry()
    v = synth.curve_fault_model(70, 70, seed=1)
    vec = synth.to_vec(v)
    ref = fd2d.forward_only(vec, og, shots=[0, 2])
    out = _np(Propagator([g]).forward(torch.tensor(vec))).reshape(5, g.nt, g.n_rcv)[[0, 2]]
    print("full-size forward rel", rel(out, ref))
    assert rel(out, ref) < SEIS_TOL


def test_dropin_signature():
    g = Geometry(nz=20, nx=24, nbc=16, nt=150)
    v = synth.make_model("flat", g.nz, g.nx, 1)
    vec = torch.tensor(synth.to_vec(v)).reshape(-1, 1)
    dvec = torch.ones_like(vec)
    dd = torch.zeros(g.n_data, 1, dtype=torch.float64)
    s, sd, gr = vel_to_seis(vec, dvec, dd, adjoint_on_residual=True, geometry=g)
    assert s.shape == (g.n_data, 1) and sd.shape == (g.n_data, 1) and gr.shape == (g.n_params, 1)
    assert s.dtype == torch.float64 and gr.dtype == torch.float64
    with pytest.raises(ValueError):
        vel_to_seis(vec[:-1], geometry=g)


def test_stream_ordered_active_masks(small_geoms):
    """seis_set_active_async: masks changed back to back on one stream, each
    followed by a forward with no host synchronisation in between; every
    launch sees its own mask (active models' seismograms equal the all-active
    run bit for bit, inactive models' outputs keep what the previous launch
    wrote), and the synchronous seis_set_active gives the same lists."""
    g, _ = small_geoms
    geoms = [g] * 5
    P = Propagator(geoms)
    vec = torch.tensor(np.concatenate([synth.to_vec(synth.make_model("curved", g.nz, g.nx, 60 + i))
                                       for i in range(5)]))
    full = P.forward(vec).clone()
    masks = [[1, 0, 1, 0, 1], [0, 1, 0, 1, 0], [1, 1, 0, 0, 0], [0, 0, 0, 0, 1], [1, 1, 1, 1, 1]]
    outs = []
    for mk in masks:
        P.set_active(mk)
        out = torch.full_like(full, -1.0)
        P.forward(vec, out=out)
        outs.append(out)
    torch.cuda.synchronize()
    for mk, out in zip(masks, outs):
        for i in range(5):
            sl = slice(int(P.seis_off[i]), int(P.seis_off[i + 1]))
            if mk[i]:
                assert torch.equal(out[sl], full[sl]), (mk, i)
            else:
                assert bool((out[sl] == -1.0).all()), (mk, i)
    import ctypes
    from seismic_amd import _lib
    m = (ctypes.c_int32 * 5)(0, 1, 1, 0, 1)
    _lib.check(_lib.lib().seis_set_active(P._h, m), "seis_set_active")
    out = torch.full_like(full, -1.0)
    P._act_idx = None
    P.forward(vec, out=out)
    torch.cuda.synchronize()
    assert torch.equal(out[int(P.seis_off[1]):int(P.seis_off[3])], full[int(P.seis_off[1]):int(P.seis_off[3])])
    assert bool((out[:int(P.seis_off[1])] == -1.0).all())
