"""mdc's other tabulated materials in the handle API (boxes made by
mdg_batch_create_tabulated from each material's own tables, species data and
lattice): forces, energies and a short trajectory against the compiled mdc
sources of that material (mdc/examples/<m> with the program's other files),
and the mixed path against FP64."""
import numpy as np
import pytest

import gpuhelp as H
import mdgpu
import mdref

pytestmark = pytest.mark.gpu

# lattice constants of each example's get_control_param
MATERIALS = {"NaCl_sx1": 5.63, "AgI": 5.08, "CaF2": 5.9}


def _setup(mat, seed):
    a, n = MATERIALS[mat], 3
    L = n * a
    R = mdref.Reference(n, n, n, a, 5.6 / L, 27, L / 2, 2.0, material=mat)
    info = R.ions()
    pe, fe = R.table()
    rng = np.random.default_rng(seed)
    r = np.mod(R.state()["r"] + rng.uniform(-0.1, 0.1, (3, R.N)), info["L"][:, None])
    v = rng.uniform(-0.02, 0.02, (3, R.N))
    R.set_state(r=r, v=v, f=np.zeros_like(r))
    tb = mdgpu.table_box(R.N, info["L"], (n, n, n), info["mol_unit"], info["Z"], info["m"],
                         info["lookup"], pe, fe, 5.6 / L, 27, L / 2)
    return R, tb, r, v, R.species()


@pytest.mark.parametrize("mat", sorted(MATERIALS))
def test_tabulated_boxes_match_mdc(gpu, mat):
    if not mdref.ref_available(mat):
        pytest.skip("oracle/_ref material harness not built (no reference checkout)")
    R, tb, r, v, sp = _setup(mat, 7)
    b = mdgpu.Batch.tabulated([tb], mdgpu.FP64)
    b.set_state(0, r=r, v=v, species=sp)
    b.forces()
    R.call("calc_foc")
    assert H.rel_err(b.state(0)["f"], R.state()["f"]) <= 1e-12
    assert b.scalars(0)["pot"] == pytest.approx(R.scalars()["pot"], rel=1e-12)
    # ten steps of mdc's loop from the same state
    R.set_state(r=r, v=v, f=np.zeros_like(r))
    b2 = mdgpu.Batch.tabulated([tb], mdgpu.FP64)
    b2.set_state(0, r=r, v=v, species=sp)
    b2.step(10)
    R.run(1, 10)
    st, sr = b2.state(0), R.state()
    assert H.rel_err(st["r"], sr["r"]) <= 1e-12
    assert H.rel_err(st["v"], sr["v"]) <= 1e-10
    sc, rs = b2.scalars(0), R.scalars()
    assert sc["kin"] == pytest.approx(rs["kin"], rel=1e-10)
    assert sc["pot"] == pytest.approx(rs["pot"], rel=1e-10)
    # the mixed path on the same box
    m = mdgpu.Batch.tabulated([tb], mdgpu.MIXED)
    m.set_state(0, r=r, v=v, species=sp)
    m.forces()
    assert H.rel_err(m.state(0)["f"], b.state(0)["f"]) <= 1e-5
    for x in (b, b2, m):
        x.close()


def test_materials_batch_together_bit_identically(gpu):
    """Boxes of three materials in one batch step to the same bits as each
    alone (per-box determinism across materials and species counts)."""
    mats = [m for m in sorted(MATERIALS) if mdref.ref_available(m)]
    if len(mats) < 2:
        pytest.skip("oracle/_ref material harnesses not built")
    sets = [_setup(m, 11 + k) for k, m in enumerate(mats)]
    mix = mdgpu.Batch.tabulated([s[1] for s in sets], mdgpu.MIXED)
    for k, (_, _, r, v, sp) in enumerate(sets):
        mix.set_state(k, r=r, v=v, species=sp)
    mix.step(16)
    for k, (_, tb, r, v, sp) in enumerate(sets):
        lone = mdgpu.Batch.tabulated([tb], mdgpu.MIXED)
        lone.set_state(0, r=r, v=v, species=sp)
        lone.step(16)
        for key in ("r", "v", "f"):
            assert np.array_equal(lone.state(0)[key], mix.state(k)[key]), (mats[k], key)
        lone.close()
    mix.close()
