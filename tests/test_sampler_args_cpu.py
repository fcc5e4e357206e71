"""Argument checks of the drop-in samplers (host logic, no device): every
validation of dwave/samplers/sa/sampler.py:321-426 and
dwave/samplers/sqa/sampler.py:315-500, :850-1045 raises the reference's
exception type and message, in the reference's order (each case below fails
on the first check it reaches)."""
import numpy as np
import pytest

from misamplers import BinaryQuadraticModel, PathIntegralAnnealingSampler, RotorModelAnnealingSampler
from misamplers import SimulatedAnnealingSampler
from misamplers._sampling import PATH_RULES, ROTOR_RULES, SA_RULES, anneal_schedule, break_buffers


def _bqm():
    return BinaryQuadraticModel.from_ising({0: 0.5, 1: -0.25, 2: 0.0}, {(0, 1): -1.0, (1, 2): 0.5})


COMMON = [
    (dict(seed=1.5), TypeError, "'seed' should be None or an integer between 0 and 2^32 - 1: value = 1.5"),
    (dict(seed=-1), ValueError, "'seed' should be an integer between 0 and 2^32 - 1: value = -1"),
    (dict(seed=2**31), ValueError, "between 0 and 2^32 - 1"),
    (dict(interrupt_function=3), TypeError, "'interrupt_function' should be a callable"),
    (dict(num_sweeps_per_beta=1.0), TypeError, "'num_sweeps_per_beta' should be a positive integer: value = 1.0"),
    (dict(num_sweeps_per_beta=0), ValueError, "'num_sweeps_per_beta' should be a positive integer: value = 0"),
    (dict(num_sweeps=7, num_sweeps_per_beta=2), ValueError, "divisible by 'num_sweeps_per_beta'"),
    (dict(beta_range=[1, 2, 3]), ValueError, "'beta_range' should be a 2-tuple"),
    (dict(beta_range=[-1, 2]), ValueError, "'beta_range' should be a 2-tuple"),
    (dict(beta_range=[0, 2], beta_schedule_type="geometric"), ValueError, "must contain non-zero values"),
    (dict(beta_schedule_type="cubic"), ValueError, "Beta schedule type cubic not implemented"),
]


def _custom(name):
    return [
        (dict(beta_schedule_type="custom"), ValueError,
         f"'{name}' must be provided for beta_schedule_type = 'custom': value is None"),
        (dict(beta_schedule_type="custom", **{name: [[1], [2, 3]]}), ValueError,
         f"{name} cannot be case as a numpy array of dtype=float"),
        (dict(beta_schedule_type="custom", num_sweeps=5, **{name: [1.0, 2.0]}), ValueError,
         f"'num_sweeps' should be set to None, or a value consistent with '{name}'"),
        (dict(beta_schedule_type="custom", beta_range=[0.5, 2.0], **{name: [1.0, 2.0]}), ValueError,
         f"'beta_range' should be set to None, or a value consistent with '{name}'"),
        (dict(beta_schedule_type="custom", **{name: [1.0, -2.0]}), ValueError, f"'{name}' cannot include negative"),
        (dict(beta_schedule_type="linear", **{name: [1.0, 2.0]}), ValueError,
         f"'{name}' must be set to None for 'beta_schedule_type' not equal to 'custom'"),
    ]


PATH_ONLY = [
    (dict(beta_schedule_type="custom", Hp_field=[1.0, 2.0], Hd_field=[1.0]), ValueError,
     "Hd_field incompatible with Hp_field"),
    (dict(num_reads=2, num_breaks=np.zeros((3, 3), np.intc)), ValueError,
     "Path integral initialization is inconsistent"),
    (dict(num_sweeps=4, schedule_sample_interval=5), ValueError,
     "schedule_sample_interval is incompatible with Hp_field"),
    (dict(num_sweeps=4, schedule_sample_interval=0), ValueError,
     "schedule_sample_interval is incompatible with Hp_field"),
]


def _escape(msg):
    for ch in "\\()^.+[]*?$|{}":
        msg = msg.replace(ch, "\\" + ch)
    return msg


@pytest.mark.parametrize("kw,exc,msg", COMMON + _custom("beta_schedule"))
def test_simulated_annealing_sampler_arguments(kw, exc, msg):
    with pytest.raises(exc, match=_escape(msg)):
        SimulatedAnnealingSampler().sample(_bqm(), **kw)


@pytest.mark.parametrize("kw,exc,msg", COMMON + _custom("Hp_field") + PATH_ONLY)
def test_path_integral_sampler_arguments(kw, exc, msg):
    kw = dict(kw)
    if "num_sweeps" in kw and kw.get("beta_schedule_type") != "custom" and "num_sweeps_per_beta" not in kw:
        kw.setdefault("beta_range", [0.1, 1.0])
    with pytest.raises(exc, match=_escape(msg)):
        PathIntegralAnnealingSampler().sample(_bqm(), **kw)


ROTOR_ONLY = [
    (dict(trans_fields=[1.0, 2.0]), ValueError, "bqm.num_variables and trans_fields lengths incompatible"),
    (dict(initial_states=np.zeros((2, 2), np.uint8)), ValueError,
     "Shape of initial_states specified asndarray should match (num_reads, num_vars)"),
    (dict(initial_states=(np.array([[0, 1, 1]], np.int8), [0, 1, 2])), ValueError,
     "expected initial_states to match the bqm vartype"),
    (dict(beta_schedule_type="custom", num_sweeps=5, Hp_field=[1.0, 2.0]), ValueError,
     "value = 5 != 2x1"),
    (dict(beta_schedule_type="custom", Hp_field=[1.0, 2.0], Hd_field=[1.0]), ValueError,
     "Hd_field incompatible with Hp_field"),
]


@pytest.mark.parametrize("kw,exc,msg", COMMON + _custom("Hp_field") + ROTOR_ONLY)
def test_rotor_sampler_arguments(kw, exc, msg):
    with pytest.raises(exc, match=_escape(msg)):
        RotorModelAnnealingSampler().sample(_bqm(), **kw)


def test_schedule_rules():
    lin, quad = {0: 1.0}, {}
    # defaults: 1000 sweeps (simulated annealing), 100 (path integral, rotor)
    assert len(anneal_schedule(SA_RULES, [0.1, 1.0], None, 1, "geometric", None, lin, quad)[0]) == 1000
    assert len(anneal_schedule(PATH_RULES, [0.1, 1.0], None, 1, "geometric", None, lin, quad)[0]) == 100
    # one sweep in the target model; linear and geometric interpolation
    np.testing.assert_array_equal(anneal_schedule(SA_RULES, [0.5, 2.0], 1, 1, "linear", None, lin, quad)[0], [2.0])
    np.testing.assert_allclose(anneal_schedule(PATH_RULES, [0.5, 2.0], 6, 2, "linear", None, lin, quad)[0],
                               [0.5, 1.25, 2.0])
    np.testing.assert_allclose(anneal_schedule(PATH_RULES, [0.5, 2.0], 3, 1, "geometric", None, lin, quad)[0],
                               [0.5, 1.0, 2.0])
    # the rotor sampler takes num_sweeps = 0 as an empty schedule before other checks
    hp, br = anneal_schedule(ROTOR_RULES, None, 0, 1, "cubic", None, lin, quad)
    assert len(hp) == 0 and br is None
    with pytest.raises(ValueError, match="not implemented"):
        anneal_schedule(PATH_RULES, None, 0, 1, "cubic", None, lin, quad)
    # the default range of the spin-form biases when none is given
    hp, br = anneal_schedule(SA_RULES, None, 2, 1, "geometric", None, {0: 1.0, 1: 1.0}, {(0, 1): 1.0})
    np.testing.assert_allclose(br, [np.log(2) / 4, np.log(2 / 0.01) / 2])  # two variables at the minimal gap


def test_break_buffer_sizing():
    hd = np.array([2.0, 1.5])
    nb, bi, bo = break_buffers((4, 10), None, None, None, True, hd, 2)
    assert nb.shape == (4, 10) and nb.dtype == np.intc and bi.size == 1
    assert bo.size == int((1 + 1.5) * (40 + np.log(40)) + 1000)
    _, _, bo = break_buffers((4, 10), None, None, None, False, hd, 2)
    assert bo.size == 1  # worldlines projected: no break list out
    given = np.full((4, 10), 2, np.intc)
    _, bi, bo = break_buffers((4, 10), given, None, None, True, np.zeros(0), 0)
    assert bi.size == 80 and bo.size == 80  # no schedule: the input breaks pass through
