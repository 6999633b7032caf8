"""Batched solver (many systems per launch) on HIP device 0.

Every system of a batch must give what the FP64 oracle gives for it alone,
with 8^3 bricks and with the column march (column widths 8 and 16, blocks of
1, 5 and 32 x layers; the specs below include meshes smaller than a column
and not multiples of it, every order and an empty system):
FP64 batch forces rel-rms <= 1e-6, energy rel <= 1e-8; FP32-mesh batch forces
rel-rms <= 2e-5, energy rel <= 2e-6.  A system solved alone and solved in a
mix with others gives the same result up to the order of floating-point
atomic additions into the mesh (rel-rms <= 1e-12 FP64, <= 1e-5 FP32); graph
replays match eager solves to the same bound.
"""
import numpy as np
import pytest

from tests import p3m_systems as S

pytestmark = pytest.mark.gpu

# (box, n, mesh, cao, r_cut): even and non-cubic meshes, every order, one
# empty system, meshes that are and are not multiples of the 8^3 brick
SPECS = [
    ((12., 12., 12.), 600, (16, 16, 16), 5, 3.0),
    ((10., 14., 12.), 900, (12, 16, 14), 3, 3.0),
    ((20., 20., 20.), 0, (16, 16, 16), 4, 4.0),
    ((15., 15., 15.), 1500, (24, 24, 24), 7, 3.5),
    ((8., 8., 8.), 300, (8, 8, 8), 2, 2.5),
    ((16., 16., 16.), 1200, (20, 20, 20), 6, 4.0),
    ((9., 9., 9.), 400, (10, 10, 10), 1, 3.0),
]


def _systems(seed=11):
    out = []
    for k, (box, n, mesh, cao, r_cut) in enumerate(SPECS):
        rng = np.random.default_rng(seed + k)
        pos = rng.uniform(0., 1., (n, 3)) * np.asarray(box)
        q = np.where(np.arange(n) % 2 == 0, 1., -1.) * (1. + 0.5 * (k % 3))
        nq = max(n, 1)
        alpha = S.closed_form_alpha(1.3, r_cut, box, nq,
                                    float(np.sum(q * q)) or 1., 1e-4)
        out.append(dict(box=box, pos=pos, q=q, mesh=list(mesh), cao=cao,
                        r_cut=r_cut, alpha=alpha, prefactor=1.3))
    return out


def _batch(systems, single):
    from espresso_amd import P3MBatch
    b = P3MBatch(single_precision=single)
    for s in systems:
        sid = b.add_system(box_l=s["box"], n=len(s["q"]), prefactor=s["prefactor"],
                           mesh=s["mesh"], cao=s["cao"], alpha=s["alpha"],
                           r_cut=s["r_cut"])
        s["id"] = sid
    return b


def _device_arrays(systems):
    import torch
    pos = np.concatenate([s["pos"].reshape(-1, 3) for s in systems])
    q = np.concatenate([s["q"] for s in systems])
    return (torch.tensor(pos, dtype=torch.float64, device="cuda"),
            torch.tensor(q, dtype=torch.float64, device="cuda"))


def _oracle(s):
    from oracle.oracle import OracleP3M
    return OracleP3M(s["box"], s["mesh"], s["cao"], s["alpha"], s["r_cut"],
                     s["prefactor"])


@pytest.mark.parametrize("shape", [
    None,                                   # default: every mesh here takes the fused y/x pass
    dict(fused_lds=0),                      # split y, x and inverse-y passes
    dict(spread="plane8", gather="layer", threads=256, chunk=32),  # small blocks
    dict(column=8, seg=32, col_lanes=8),    # column march, 8 x 8 columns
    dict(column=16, seg=5, col_lanes=16, fused_lds=0),  # 16 x 16, 5 layers per block
    dict(column=8, seg=1, col_lanes=16, threads=256, chunk=7),  # one layer per block
    dict(group=4),                          # spread: 2 x 2 x 1 bricks per tile
    dict(group=4, threads=256, chunk=7),    # the same with small chunks
    dict(threads=256, gather_threads=512),  # gather in larger blocks than the spread
    dict(threads=512, gather_threads=256, gather="plane8"),
], ids=["fused", "split", "small_blocks", "col8", "col16_seg5", "col8_seg1", "group4",
        "group4_chunk7", "gather512", "gather256"])
@pytest.mark.parametrize("single,f_tol,e_tol", [(False, 1e-6, 1e-8),
                                                (True, 2e-5, 2e-6)])
def test_batch_matches_oracle_per_system(single, f_tol, e_tol, shape):
    import torch
    systems = _systems()
    b = _batch(systems, single)
    if shape is not None:
        if shape.get("column"):
            b.set_deterministic(False)  # FP64 default: 8^3 bricks
        b.set_launch_shape(**shape)
    pos, q = _device_arrays(systems)
    force = torch.zeros_like(pos)
    energy, virial = b.solve(pos, q, force, energy=True, pressure=True)
    f = force.cpu().numpy()
    for k, s in enumerate(systems):
        off, n = b.offset(s["id"]), len(s["q"])
        assert b.system_bytes(s["id"]) > 0
        if n == 0:
            assert energy[k] == 0. and np.all(virial[k] == 0.)
            continue
        ora = _oracle(s)
        f_ref = ora.forces(s["pos"], s["q"])
        err = S.rel_rms(f[off:off + n], f_ref)
        assert err <= f_tol, (k, err)
        e_ref = ora.energy(s["pos"], s["q"])
        assert abs(energy[k] - e_ref) <= e_tol * abs(e_ref), (k, energy[k], e_ref)
        p_ref = ora.pressure(s["pos"], s["q"])
        scale = np.max(np.abs(p_ref))
        assert np.max(np.abs(virial[k] - p_ref)) <= f_tol * scale, k


@pytest.mark.parametrize("single,tol", [(False, 1e-12), (True, 1e-5)])
def test_alone_equals_in_mix(single, tol):
    import torch
    systems = _systems(seed=5)
    b = _batch(systems, single)
    pos, q = _device_arrays(systems)
    mix = torch.zeros_like(pos)
    e_mix, _ = b.solve(pos, q, mix, energy=True)
    for k in (0, 3, 6):
        s = systems[k]
        off, n = b.offset(s["id"]), len(s["q"])
        alone = torch.zeros_like(pos)
        e_alone, _ = b.solve(pos, q, alone, systems=[s["id"]], energy=True)
        a = alone.cpu().numpy()
        assert S.rel_rms(a[off:off + n], mix.cpu().numpy()[off:off + n]) <= tol
        assert abs(e_alone[0] - e_mix[k]) <= tol * abs(e_mix[k])
        # particles of unselected systems are untouched
        a[off:off + n] = 0.
        assert np.all(a == 0.)
    # a selection in another order returns results indexed like the selection
    sel = [5, 1, 3]
    f2 = torch.zeros_like(pos)
    e_sel, _ = b.solve(pos, q, f2, systems=sel, energy=True)
    for i, k in enumerate(sel):
        assert abs(e_sel[i] - e_mix[k]) <= tol * abs(e_mix[k])


def test_single_system_batch_equals_actor():
    import torch
    from espresso_amd.electrostatics import P3M
    pos, q, box, params = S.config_c1()
    s = dict(box=box, pos=pos, q=q, mesh=[32] * 3, cao=5, r_cut=3.0,
             alpha=params["alpha"], prefactor=2.0)
    b = _batch([s], single=False)
    dpos, dq = _device_arrays([s])
    force = torch.zeros_like(dpos)
    e, _ = b.solve(dpos, dq, force, energy=True)
    p3m = P3M(tune=False, **params)
    p3m.activate(pos, q, box_l=box)
    f_ref = np.zeros_like(pos)
    p3m.add_long_range_forces(pos, q, f_ref)
    assert S.rel_rms(force.cpu().numpy(), f_ref) <= 1e-12
    assert abs(e[0] - p3m.long_range_energy(pos, q)) <= 1e-12 * abs(e[0])


def test_graph_replays_match_eager():
    import torch
    systems = _systems(seed=21)
    b = _batch(systems, single=True)
    pos, q = _device_arrays(systems)
    ref = torch.zeros_like(pos)
    b.solve(pos, q, ref)
    ref_np = ref.cpu().numpy()
    force = torch.zeros_like(pos)
    for rep in range(4):
        torch.cuda.synchronize()
        force.zero_()
        torch.cuda.synchronize()
        b.graph_solve(pos, q, force)
        b.synchronize()
        got = force.cpu().numpy()
        assert S.rel_rms(got, ref_np) <= 1e-5, rep
        # eager work between replays: an eager solve of the same selection
        # and unrelated torch work on the device
        extra = torch.zeros_like(pos)
        b.solve(pos, q, extra)
        assert S.rel_rms(extra.cpu().numpy(), ref_np) <= 1e-5
        _ = (torch.randn(1 << 20, device="cuda") * 2.).sum().item()
    # another selection re-captures; the first one again afterwards
    sub = torch.zeros_like(pos)
    b.graph_solve(pos, q, sub, systems=[3, 0])
    b.synchronize()
    for k in (3, 0):
        off, n = b.offset(k), len(systems[k]["q"])
        assert S.rel_rms(sub.cpu().numpy()[off:off + n], ref_np[off:off + n]) <= 1e-5
    force.zero_()
    torch.cuda.synchronize()
    b.graph_solve(pos, q, force)
    b.synchronize()
    assert S.rel_rms(force.cpu().numpy(), ref_np) <= 1e-5


def test_batch_argument_errors():
    import torch
    systems = _systems()[:2]
    b = _batch(systems, single=True)
    pos, q = _device_arrays(systems)
    force = torch.zeros_like(pos)
    with pytest.raises(ValueError, match="selected twice"):
        b.solve(pos, q, force, systems=[0, 0])
    with pytest.raises(ValueError, match="invalid system id"):
        b.solve(pos, q, force, systems=[2])
    with pytest.raises(ValueError, match="expected"):
        b.solve(pos[:-1], q, force)
    with pytest.raises(TypeError):
        b.solve(pos.cpu(), q, force)
    with pytest.raises(ValueError, match="Parameter 'cao' must be >= 1 and <= 7"):
        b.add_system(box_l=[10.] * 3, n=10, prefactor=1., mesh=16, cao=8,
                     alpha=1., r_cut=2.)
    with pytest.raises(ValueError, match="fixed mesh"):
        b.add_system(box_l=[10.] * 3, n=10, prefactor=1., mesh=[16, -1, -1],
                     cao=3, alpha=1., r_cut=2.)
    with pytest.raises(RuntimeError):
        # k-space cutoff larger than the box (p3m.cpp:710-750)
        b.add_system(box_l=[4.] * 3, n=10, prefactor=1., mesh=6, cao=6,
                     alpha=1., r_cut=1.)


def test_launch_shapes_and_autotune_keep_results():
    import torch
    systems = _systems(seed=31)
    b = _batch(systems, single=True)
    pos, q = _device_arrays(systems)
    ref = torch.zeros_like(pos)
    b.solve(pos, q, ref)
    ref_np = ref.cpu().numpy()
    f = torch.zeros_like(pos)
    # particle-mesh block size and chunk, including one-particle and odd chunks
    blocks = [(512, 64), (256, 32), (256, 7), (512, 1)]
    k = 0
    for fused in (0, 96 * 1024):
        for gather, spread in (("layer", "plane"), ("plane", "plane16"),
                               ("plane16", "plane8"), ("plane8", "plane")):
            for ncol, zl in ((1, 1), (2, 4), (16, 16), (64, 64)):
                threads, chunk = blocks[k % len(blocks)]
                k += 1
                b.set_launch_shape(gather, ncol, zl, fused, spread, threads, chunk)
                assert b.launch_shape()["chunk"] == chunk
                f.zero_()
                torch.cuda.synchronize()
                b.solve(pos, q, f)
                assert S.rel_rms(f.cpu().numpy(), ref_np) <= 1e-5, (fused, gather, spread,
                                                                     ncol, zl, threads, chunk)
    with pytest.raises(ValueError, match="powers of two"):
        b.set_launch_shape("plane", 3, 4)
    with pytest.raises(ValueError, match="gather must be"):
        b.set_launch_shape(5)
    with pytest.raises(ValueError, match="threads must be"):
        b.set_launch_shape(threads=128)
    with pytest.raises(ValueError, match="chunk must be"):
        b.set_launch_shape(threads=256, chunk=90)
    b.set_launch_shape()  # autotuning starts from the default shape
    scratch = torch.zeros_like(pos)
    shape = b.autotune(pos, q, scratch, reps=1)
    assert shape["ncol"] in (2, 4, 8, 16) and shape["zlines"] in (4, 8, 16)
    assert shape["fused_lds"] in (0, 96 * 1024)
    assert shape["gather"] in ("layer", "plane", "plane16", "plane8")
    assert shape["spread"] in ("plane", "plane16", "plane8")
    assert (shape["threads"], shape["chunk"]) in ((512, 64), (256, 64), (256, 32))
    assert 0. < shape["best_ms"] <= shape["default_ms"]
    f.zero_()
    torch.cuda.synchronize()
    b.solve(pos, q, f)
    assert S.rel_rms(f.cpu().numpy(), ref_np) <= 1e-5


def test_launch_shape_validation():
    """Invalid launch shapes are refused before anything changes, with a
    message naming the field; a valid split of the block sizes is kept."""
    b = _batch(_systems()[:2], False)
    before = b.launch_shape()
    for kw, msg in [(dict(gather_threads=128), "gather_threads must be 0, 256 or 512"),
                    (dict(threads=512, chunk=64, gather_threads=0, group=2),
                     "group must be 1 or 4"),
                    (dict(threads=256, chunk=64, gather_threads=512, column=12),
                     "column must be 0"),
                    (dict(threads=128), "threads must be 256 or 512")]:
        with pytest.raises(ValueError, match=msg):
            b.set_launch_shape(**kw)
        assert b.launch_shape() == before
    b.set_launch_shape(threads=256, chunk=64, gather_threads=512)
    got = b.launch_shape()
    assert got["threads"] == 256 and got["gather_threads"] == 512
