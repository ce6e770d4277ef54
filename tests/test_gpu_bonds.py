"""GPU vs oracle, bit-exact, on the rows around the core step: bonds (A14,
including breakage), the wmode-3 shear-stress boundary (A17), the adaptive
grid (A18) and the rate-state velocity switch (A20)."""

import numpy
import pytest

import oracle
from sphere_amd import engine

from . import cases
from .test_gpu_parity import EXACT_FIELDS, _compare, _run_pair

pytestmark = pytest.mark.gpu


def _bonded(frac=0.2, seed=3, sigma_b=numpy.inf):
    s = cases.fixture("shear1k.bin")
    n = int(s.np[0])
    o = oracle.Oracle(s.tobytes())
    o.step(1)
    rows = o.get("contacts")
    pairs = sorted({(min(a, int(b)), max(a, int(b))) for a in range(n) for b in rows[a]
                    if b != n and a != int(b)})
    rng = numpy.random.default_rng(seed)
    pick = rng.choice(len(pairs), size=max(1, int(frac * len(pairs))), replace=False)
    for k in sorted(pick):
        s.bond(*pairs[k])
    s.sigma_b[0] = sigma_b
    s.tau_b[0] = sigma_b
    return s


def test_bonds_bit_exact():
    s = _bonded()
    ctx, ref = _run_pair(s, 150, chunks=3)
    bad = _compare(ctx, ref, EXACT_FIELDS + ("bonds", "bonds_delta", "bonds_omega"))
    assert bad == []
    ctx.close()


def test_bond_breakage_bit_exact():
    """Weak bonds break (first index reset to nb0, src/cohesion.cuh:205)."""
    s = _bonded(frac=0.3, seed=4, sigma_b=300.0)
    ctx, ref = _run_pair(s, 120, chunks=4)
    b = ref.get("bonds")
    assert (b[:, 0] == int(s.nb0[0])).any()  # some bonds broke
    bad = _compare(ctx, ref, EXACT_FIELDS + ("bonds", "bonds_delta", "bonds_omega"))
    assert bad == []
    ctx.close()


@pytest.mark.parametrize("bond_k,ext", [(1.0e5, 0.0), (0.0, 1.0e-4), (1.0e5, 1.0e-4)])
def test_bond_options_bit_exact(bond_k, ext):
    """The bond options (SPH_S_BOND_K, SPH_S_BOND_BREAK_EXT): an explicit
    bond stiffness in N/m and a break at a normal extension, on the engine
    and the oracle; with the extension break (1e-4 of the smaller radius)
    some bonds break within the 120 steps (about 35 of 530 on the oracle)."""
    s = _bonded(frac=0.3, seed=4)
    rng = numpy.random.default_rng(6)
    s.vel[:, :3] += rng.normal(0.0, 0.01, (int(s.np[0]), 3))
    first = s.bonds[:, 0].copy()
    image = s.tobytes()
    ref = oracle.Oracle(image)
    ref.set_bond_model(bond_k, ext)
    with engine.Context(image) as ctx:
        ctx.set_scalar("bond_k", bond_k)
        ctx.set_scalar("bond_break_ext", ext)
        assert ctx.scalar("bond_k") == bond_k and ctx.scalar("bond_break_ext") == ext
        for _ in range(4):
            ref.step(30)
            ctx.step(30)
            bad = _compare(ctx, ref, EXACT_FIELDS + ("bonds", "bonds_delta", "bonds_omega"))
            assert bad == []
    # A broken bond's first index becomes nb0 (src/cohesion.cuh:205); a
    # bond whose first particle is nb0 from the start is not counted.
    nb0 = int(s.nb0[0])
    broken = int(((ref.get("bonds")[:, 0] == nb0) & (first != nb0)).sum())
    if ext > 0.0:
        assert broken > 0
    else:
        assert broken == 0


def test_wmode3_shear_stress_bit_exact():
    s = cases.fixture("shear1k.bin")
    s.shear(shear_stress=2.0e3)
    assert int(s.wmode[0]) == 3
    ctx, ref = _run_pair(s, 150, chunks=3)
    bad = _compare(ctx, ref, EXACT_FIELDS + ("walls_nx", "walls_mvfd", "walls_acc"))
    assert bad == []
    ctx.close()


def test_adaptive_grid_bit_exact():
    s = cases.fixture("shear1k.bin")
    s.adaptiveGrid()
    ctx, ref = _run_pair(s, 150, chunks=3)
    assert _compare(ctx, ref, EXACT_FIELDS + ("walls_nx", "walls_mvfd")) == []
    assert ctx.tobytes() == ref.tobytes()
    ctx.close()


def test_rate_state_switch_bit_exact():
    """fixvel in (5, 10) particles get v_x x10 when t crosses 10 s
    (src/device.cu:2454-2462)."""
    s = cases.fixture("shear1k.bin")
    top = s.fixvel > 0
    s.fixvel[top] = 6.0
    dt = float(s.time_dt[0])
    s.time_current[0] = 10.0 - 20.5 * dt
    s.time_total[0] = 20.0
    ctx, ref = _run_pair(s, 60, chunks=2)
    assert _compare(ctx, ref) == []
    v = ref.get("vel")
    assert numpy.allclose(v[top, 0], 10.0 * s.vel[top, 0])
    ctx.close()
