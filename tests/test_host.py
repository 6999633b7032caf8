"""Host-side tests of the C ABI library (no GPU): the library loads and
exports every entry point of include/espresso_amd_p3m.h, parameter
validation and activation errors carry the reference's messages
(src/core/p3m/common.hpp:94-150, electrostatics.py:148-176,
coulomb_interface.py, p3m_tuning_exceptions.py, long_range_actors.py), and
the host error estimates and pair terms agree with the oracle."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from espresso_amd import _lib
from espresso_amd.electrostatics import P3M, P3MGPU
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
VALID = dict(prefactor=1., accuracy=1e-3, mesh=16, cao=5, r_cut=2.0, alpha=1.0,
             tune=False)


def test_library_exports_header_symbols():
    header = open(os.path.join(ROOT, "include", "espresso_amd_p3m.h")).read()
    declared = set(re.findall(r"^\s*(?:int|const char \*)\s*(p3m_\w+)\s*\(",
                              header, re.M))
    assert declared == set(_lib.EXPORTED_SYMBOLS)
    L = _lib.load()
    for name in declared:
        assert hasattr(L, name), name


def host(**kw):
    return P3M(device=-1, **kw)


@pytest.mark.parametrize("kw,msg", [
    (dict(epsilon=-1.), "Parameter 'epsilon' must be >= 0"),
    (dict(accuracy=0.), "Parameter 'accuracy' must be > 0"),
    (dict(r_cut=-2.), "Parameter 'r_cut' must be > 0"),
    (dict(alpha=0.), "Parameter 'alpha' must be > 0"),
    (dict(mesh=0), "Parameter 'mesh' must be > 0"),
    (dict(mesh_off=[2., 0.5, 0.5]), "Parameter 'mesh_off' must be >= 0 and <= 1"),
    (dict(cao=8), "Parameter 'cao' must be >= 1 and <= 7"),
    (dict(cao=7, mesh=6), "Parameter 'cao' cannot be larger than 'mesh'"),
    (dict(timings=0), "Parameter 'timings' must be > 0"),
    (dict(prefactor=-1.), "Parameter 'prefactor' must be > 0"),
    (dict(mesh=[16, 15, 16]),
     "P3M requires an even number of mesh points in all directions"),
])
def test_parameter_validation_messages(kw, msg):
    params = dict(VALID)
    params.update(kw)
    with pytest.raises(ValueError, match=re.escape(msg)):
        host(**params)


def test_tuning_sentinels_accepted():
    # -1 means "to be tuned" when tune=True (common.hpp:104-118)
    p = host(prefactor=1., accuracy=1e-3)
    assert p.get_params()["is_tuned"] is False
    with pytest.raises(ValueError, match="Parameter 'r_cut' must be > 0"):
        host(prefactor=1., accuracy=1e-3, tune=False, mesh=16, cao=5, alpha=1.)


def test_gpu_actor_requires_single_precision():
    with pytest.raises(ValueError, match="P3M GPU only implemented in single-precision mode"):
        P3MGPU(device=-1, single_precision=False, **VALID)
    assert P3MGPU(device=-1, **VALID).get_params()["single_precision"] is True


def test_force_convention_of_the_actors():
    """P3MGPU forces follow the reference's GPU kernels, P3M's the CPU
    kernels; other values are rejected with the library's message."""
    from espresso_amd import _lib
    assert P3MGPU(device=-1, **VALID).force_convention() == "gpu"
    p = P3M(device=-1, **VALID)
    assert p.force_convention() == "cpu"
    with pytest.raises(ValueError, match="force convention must be"):
        _lib.raise_for(_lib.load().p3m_set_force_convention(p._ctx, 2), p._ctx)
    assert p.force_convention() == "cpu"


def test_tuning_real_space_modes():
    from espresso_amd import _lib
    p = P3M(device=-1, **VALID)
    p.set_tuning_real_space("device")
    p.set_tuning_real_space("model")
    with pytest.raises(ValueError, match="tuning real space must be one of"):
        p.set_tuning_real_space("host")
    with pytest.raises(ValueError, match="P3M_TUNE_RS_MODEL or P3M_TUNE_RS_DEVICE"):
        _lib.raise_for(_lib.load().p3m_set_tuning_real_space(p._ctx, 7), p._ctx)


def test_missing_required_keys():
    with pytest.raises(ValueError, match="prefactor"):
        P3M(device=-1, accuracy=1e-3)


def _two(box=10.):
    pos = np.array([[1., 1., 1.], [box / 2, box / 2, box / 2]])
    return pos, np.array([-1., 1.])


def test_activation_errors():
    pos, q = _two()
    # charge neutrality (coulomb.cpp:286-325)
    with pytest.raises(RuntimeError, match="The system is not charge neutral"):
        host(**VALID).activate(pos, np.array([1., 1.]), box_l=[10.] * 3)
    host(check_neutrality=False, **VALID)  # accepted
    # non-cubic box with non-metallic epsilon (p3m.cpp:812-818)
    with pytest.raises(RuntimeError,
                       match="CoulombP3M: non-metallic epsilon requires cubic box"):
        host(**dict(VALID, epsilon=1.)).activate(pos, q, box_l=[10., 10., 20.])
    # no charges while tuning (p3m.cpp:768-771)
    with pytest.raises(RuntimeError,
                       match="CoulombP3M: no charged particles in the system"):
        host(prefactor=2., accuracy=1e-2).activate(pos, np.zeros(2), box_l=[10.] * 3)
    # k-space cutoff larger than half the box (p3m.cpp:795-801,
    # long_range_actors.py:226-236)
    with pytest.raises(RuntimeError,
                       match="k-space cutoff .+ is larger than half of box dimension"):
        host(prefactor=1., accuracy=1., r_cut=1., alpha=1., cao=6, mesh=6,
             tune=False).activate(pos, q, box_l=[20.] * 3)
    # a valid host-only activation stops at the device
    with pytest.raises(RuntimeError, match="host-only context has no device"):
        host(**VALID).activate(pos, q, box_l=[10.] * 3)


@pytest.mark.parametrize("mesh,cao", [([16, 16, 16], 3), ([32, 24, 20], 5),
                                      ([48, 48, 48], 7)])
def test_k_space_error_matches_oracle(mesh, cao):
    L = _lib.load()
    box = [11., 9., 10.]
    out = C.c_double()
    m = (C.c_int * 3)(*mesh)
    b = (C.c_double * 3)(*box)
    for alpha_L in (5., 10., 20.):
        _lib.raise_for(L.p3m_k_space_error(1.3, m, cao, 500, 480., alpha_L, b,
                                           C.byref(out)))
        ref = O.k_space_error(1.3, mesh, cao, 500, 480., alpha_L, box)
        assert abs(out.value - ref) <= 1e-12 * ref
        _lib.raise_for(L.p3m_real_space_error(1.3, 0.2, 500, 480., alpha_L, b,
                                              C.byref(out)))
        ref = O.real_space_error(1.3, 0.2, 500, 480., alpha_L, box)
        assert abs(out.value - ref) <= 1e-14 * ref


def test_pair_terms_match_oracle():
    p = host(**dict(VALID, prefactor=1.7, r_cut=3.0, alpha=0.8))
    pos = np.array([[1., 1., 1.], [2.3, 1.4, 0.2]])
    q = np.array([1.5, -0.5])
    d = pos[0] - pos[1]
    dist = float(np.linalg.norm(d))
    e_ref, f_ref = O.real_space(pos, q, [10.] * 3, 1.7, 0.8, 3.0)
    np.testing.assert_allclose(p.pair_force(q[0] * q[1], d, dist), f_ref[0],
                               rtol=1e-14)
    assert abs(p.pair_energy(q[0] * q[1], dist) - e_ref) <= 1e-14 * abs(e_ref)
    assert p.pair_energy(1., 3.5) == 0.
    assert np.all(p.pair_force(1., [0., 0., 3.5], 3.5) == 0.)


def test_tune_model_reaches_accuracy_within_ranges():
    # reference tuning algorithm (p3m.cpp:594-791) with a cost model: the
    # result must satisfy the requested accuracy by the oracle's estimate
    from espresso_amd import tune_model
    rng = np.random.default_rng(3)
    n = 2000
    q = np.where(rng.random(n) < 0.5, 1., -1.)
    box = [15., 15., 15.]
    for acc, mr, cr in ((1e-4, (16, 48), (3, 7)), (1e-3, (8, 24), (2, 5))):
        t = tune_model(box, n, float(np.sum(q * q)), 1.0, acc,
                       mesh_range=mr, cao_range=cr)
        assert mr[0] <= t["mesh"][0] <= mr[1] and cr[0] <= t["cao"] <= cr[1]
        assert t["mesh"][0] == t["mesh"][1] == t["mesh"][2]
        ref = O.calculate_accuracy(1.0, t["mesh"], t["cao"], t["r_cut"] / box[0],
                                   n, float(np.sum(q * q)), acc, box)
        assert ref[0] <= acc
        # the tuner's k-space estimate sums the same terms grouped by |n|^2;
        # rounding of the cancelling terms differs (<= 1e-6 relative)
        assert abs(ref[0] - t["accuracy"]) <= 1e-6 * ref[0]
        assert abs(ref[3] - t["alpha"] * box[0]) <= 1e-9 * ref[3]


def test_tune_model_failure_message():
    from espresso_amd import tune_model
    with pytest.raises(RuntimeError, match="failed to reach requested accuracy"):
        tune_model([10.] * 3, 100, 100., 1.0, 1e-9, mesh_range=(4, 8),
                   cao_range=(1, 2))
    with pytest.raises(RuntimeError, match="no charged particles"):
        tune_model([10.] * 3, 0, 0., 1.0, 1e-3)


def test_grouped_k_space_estimate_of_the_tuner():
    # the tuner's grouped sum (p3m_host.cpp k_space_error_fast, reached here
    # through the tuned accuracy) against the oracle's term-by-term sum over
    # meshes, orders and splitting parameters
    from espresso_amd import tune_model
    for n, L, acc in ((800, 9., 1e-3), (3000, 14., 1e-4), (20000, 30., 1e-5)):
        t = tune_model([L] * 3, n, float(n), 1.0, acc, mesh_range=(8, 64))
        ref = O.calculate_accuracy(1.0, t["mesh"], t["cao"], t["r_cut"] / L, n,
                                   float(n), acc, [L] * 3)
        assert abs(ref[0] - t["accuracy"]) <= 1e-6 * ref[0]
        assert ref[0] <= acc * (1. + 1e-6)


def test_elc_parameters_and_messages():
    """ELC construction checks with the reference's messages (elc_data,
    elc.cpp:1056-1102; electrostatics.py ELC.validate_params) and the
    metallic epsilon it imposes on the P3M actor (p3m.impl.hpp:112-114)."""
    from espresso_amd.electrostatics import ELC
    p = P3M(device=-1, epsilon=2.0, **VALID)
    elc = ELC(actor=p, maxPWerror=1e-6, gap_size=3.)
    g = elc.get_params()
    assert g["far_cut"] == -1. and g["neutralize"] is True and g["gap_size"] == 3.
    assert p.get_params()["epsilon"] == 0.
    cases = [
        (dict(maxPWerror=-1., gap_size=3.), ValueError, "'maxPWerror' must be > 0"),
        (dict(maxPWerror=1e-6, gap_size=0.), ValueError, "'gap_size' must be > 0"),
        (dict(maxPWerror=1e-6, gap_size=3., far_cut=-2.), ValueError, "'far_cut' must be > 0"),
        (dict(maxPWerror=1e-6, gap_size=3., pot_diff=1.), ValueError,
         "'const_pot' must be True when 'pot_diff' is non-zero"),
        (dict(maxPWerror=1e-6, gap_size=3., delta_mid_top=2.), ValueError,
         "'delta_mid_top' must be >= -1 and <= \\+1"),
        (dict(maxPWerror=1e-6, gap_size=3., delta_mid_bot=-1.5), ValueError,
         "'delta_mid_bot' must be >= -1 and <= \\+1"),
        (dict(maxPWerror=1e-6, gap_size=3., delta_mid_top=1., delta_mid_bot=1.), ValueError,
         "two parallel metallic boundaries requires the const_pot option"),
        (dict(maxPWerror="x", gap_size=3.), ValueError, "maxPWerror has to be a float"),
        (dict(maxPWerror=1e-6, gap_size=3., neutralize=1), ValueError,
         "neutralize has to be a bool"),
    ]
    for kw, exc, msg in cases:
        with pytest.raises(exc, match=msg):
            ELC(actor=P3M(device=-1, **VALID), **kw)
    with pytest.raises(ValueError, match="isn't supported by ELC"):
        ELC(actor=object(), maxPWerror=1e-6, gap_size=3.)
    with pytest.raises(ValueError, match="missing"):
        ELC(actor=p, maxPWerror=1e-6)
    elc.remove()
    # contrasts switch neutralisation off (elc_data, elc.cpp:1056-1070)
    d = ELC(actor=P3M(device=-1, **VALID), maxPWerror=1e-6, gap_size=3., delta_mid_bot=0.5,
            delta_mid_top=-2e-15).get_params()
    assert d["delta_mid_bot"] == 0.5 and d["neutralize"] is False and d["const_pot"] is False
    d = ELC(actor=P3M(device=-1, **VALID), maxPWerror=1e-6, gap_size=3., const_pot=True,
            pot_diff=1.5).get_params()
    assert d["const_pot"] is False and d["pot_diff"] == 1.5


def test_dipolar_p3m_parameters_and_messages():
    """DipolarP3M keyword checks (magnetostatics.py DipolarP3M,
    dp3m.impl.hpp:79-92) on a host-only context."""
    from espresso_amd.magnetostatics import DipolarP3M
    d = DipolarP3M(device=-1, prefactor=1.1, accuracy=1e-6, mesh=32, epsilon="metallic")
    g = d.get_params()
    assert g["mesh"] == [32, 32, 32] and g["epsilon"] == 0. and g["mesh_off"] == [0.5] * 3
    with pytest.raises(ValueError, match="missing"):
        DipolarP3M(device=-1, prefactor=1.)
    with pytest.raises(ValueError, match="requires a cubic mesh"):
        DipolarP3M(device=-1, prefactor=1., accuracy=1e-4, mesh=[16, 16, 18])
    with pytest.raises(ValueError, match="'epsilon' has to be a float or 'metallic'"):
        DipolarP3M(device=-1, prefactor=1., accuracy=1e-4, epsilon="vacuum")
    with pytest.raises(TypeError, match="'timings' has to be an integer"):
        DipolarP3M(device=-1, prefactor=1., accuracy=1e-4, timings=2.5)
    with pytest.raises(ValueError, match="'timings' must be > 0"):
        DipolarP3M(device=-1, prefactor=1., accuracy=1e-4, timings=0)
    with pytest.raises(ValueError, match="'prefactor' must be > 0"):
        DipolarP3M(device=-1, prefactor=0., accuracy=1e-4)
    with pytest.raises(RuntimeError, match="host-only"):
        d.activate(np.zeros((2, 3)), np.ones((2, 3)), box_l=10.)
    with pytest.raises(RuntimeError, match="cubic box"):
        d.activate(np.zeros((2, 3)), np.ones((2, 3)), box_l=[10., 10., 11.])


def test_charge_neutrality_tolerance_parameter():
    """Actor::set_charge_neutrality_tolerance (Actor.hpp:71-79,
    Actor.impl.hpp:44-80): default 2e-12, an explicit tolerance, None and
    check_neutrality=False switch the check off, negative values raise."""
    assert host(**VALID).get_params()["charge_neutrality_tolerance"] == 2e-12
    g = host(charge_neutrality_tolerance=1e-10, **VALID).get_params()
    assert g["charge_neutrality_tolerance"] == 1e-10 and g["check_neutrality"] is True
    g = host(check_neutrality=False, **VALID).get_params()
    assert g["charge_neutrality_tolerance"] is None and g["check_neutrality"] is False
    # None first, then check_neutrality=True restores the default (the
    # reference applies the two parameters in this order)
    g = host(charge_neutrality_tolerance=None, **VALID).get_params()
    assert g["charge_neutrality_tolerance"] == 2e-12
    with pytest.raises(ValueError, match="'charge_neutrality_tolerance' must be >= 0"):
        host(charge_neutrality_tolerance=-1e-3, **VALID)


def test_checkpoint_round_trip():
    """Pickled actors come back with the same parameters; a tuned actor is
    rebuilt with is_tuned set, so it is not tuned again on activation
    (CoulombP3M.hpp:102-128; test_checkpoint.py:846-888)."""
    import pickle
    from espresso_amd.electrostatics import ELC, P3MGPU
    from espresso_amd.magnetostatics import DipolarP3M
    ref = dict(prefactor=1.0, accuracy=0.1, mesh=10, cao=1, alpha=1.0, r_cut=1.0,
               tune=False, timings=15, check_neutrality=True,
               check_complex_residuals=False, charge_neutrality_tolerance=1e-12)
    for cls in (P3M, P3MGPU):
        a = cls(device=-1, **ref)
        b = pickle.loads(pickle.dumps(a))
        assert type(b) is cls
        assert b.get_params() == a.get_params()
        g = b.get_params()
        assert g["mesh"] == [10] * 3 and g["cao"] == 1 and g["timings"] == 15
        assert g["check_complex_residuals"] is False and g["is_tuned"] is True
        assert g["charge_neutrality_tolerance"] == 1e-12
    # an actor tuned earlier: tune=True, concrete parameters, is_tuned
    tuned = P3M(device=-1, prefactor=2., accuracy=1e-4, mesh=[32, 32, 32], cao=5,
                alpha=1.1, r_cut=3.0, tune=True, is_tuned=True)
    g = pickle.loads(pickle.dumps(tuned)).get_params()
    assert g["tune"] is True and g["is_tuned"] is True and g["alpha"] == 1.1
    # an actor still to be tuned stays so
    fresh = P3M(device=-1, prefactor=1., accuracy=1e-3)
    g = pickle.loads(pickle.dumps(fresh)).get_params()
    assert g["is_tuned"] is False and g["mesh"] == [-1, -1, -1]
    elc = ELC(actor=P3M(device=-1, **VALID), maxPWerror=0.1, gap_size=6.0,
              delta_mid_top=0.9, delta_mid_bot=0.1, charge_neutrality_tolerance=5e-12)
    e2 = pickle.loads(pickle.dumps(elc))
    g1, g2 = elc.get_params(), e2.get_params()
    for k in ("gap_size", "maxPWerror", "delta_mid_top", "delta_mid_bot",
              "check_neutrality", "charge_neutrality_tolerance", "far_cut"):
        assert g1[k] == g2[k], k
    assert g2["charge_neutrality_tolerance"] == 5e-12
    assert g2["actor"].get_params() == elc.actor.get_params()
    d = DipolarP3M(device=-1, prefactor=1.1, accuracy=1e-6, mesh=32, cao=3, alpha=0.5,
                   r_cut=2.0, tune=False)
    d2 = pickle.loads(pickle.dumps(d))
    assert d2.get_params() == d.get_params() and d2.get_params()["is_tuned"] is True
