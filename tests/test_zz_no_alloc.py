"""omega_model_compute_tendency_noalloc (the tendency evaluation with its
diagnostics in the library's static scratch array) against
omega_model_compute_tendency, bit for bit, through the C ABI: two calls in a
row, the second on a smaller model, in replace and increment mode, with the
pattern-specialised and the general kernels and every mixing term on; a
model too large for the scratch is refused; the model's own auxiliary state
is left as it was."""
import numpy as np
import pytest

from omega_playground_amd import swe

from helpers import params_dict, rich_state, same_bits

pytestmark = pytest.mark.gpu


def case(nx, L, ntr, dc=10e3, seed=0):
    mesh = swe.PlanarHexagonalMesh(nx, nx, dc, L)
    model = swe.ShallowWaterModel(mesh, swe.ShallowWaterParams(**params_dict(dc, ntr, True)))
    state = swe.ShallowWaterState(mesh, ntr)
    state.upload(*rich_state(mesh, L, ntr, seed=seed))
    return mesh, model, state


def compare(mesh, model, state, ntr, mode):
    rng = np.random.default_rng(mode + 3)
    start = tuple(rng.standard_normal(a.shape) for a in state.download())
    ref = swe.ShallowWaterState(mesh, ntr)
    got = swe.ShallowWaterState(mesh, ntr)
    ref.upload(*start)
    got.upload(*start)
    model.compute_tendency(ref, state, 0.0, mode)
    model.compute_tendency_noalloc(got, state, 0.0, mode)
    for name, a, b in zip(("h", "vn", "tr"), got.download(), ref.download()):
        assert same_bits(a, b), name


@pytest.mark.parametrize("mode", [swe.ADD_REPLACE, swe.ADD_INCREMENT])
def test_noalloc_matches_compute_tendency_twice_second_smaller(mode):
    big = case(16, 64, 3)       # pattern-specialised kernels, all mixing terms
    small = case(9, 8, 1, seed=1)  # general kernels
    print("noalloc: nx=16 L=64 ntr=3, then nx=9 L=8 ntr=1, mode", mode, flush=True)
    assert big[1].kernel_variant() == 2 and small[1].kernel_variant() == 1
    compare(*big, 3, mode)
    compare(*small, 1, mode)
    # and back to the larger one after the smaller one left its contents
    compare(*big, 3, mode)


def test_noalloc_leaves_the_model_auxiliary_state():
    mesh, model, state = case(10, 64, 2)
    other = swe.ShallowWaterState(mesh, 2)
    other.upload(*rich_state(mesh, 64, 2, seed=9))
    model.compute_aux(state)
    ke = model.aux_array("ke_cell")
    out = swe.ShallowWaterState(mesh, 2)
    model.compute_tendency_noalloc(out, other, 0.0)
    assert same_bits(model.aux_array("ke_cell"), ke)


def test_noalloc_refuses_models_beyond_the_scratch():
    mesh, model, state = case(128, 64, 4, seed=2)  # ~134 MB of diagnostics
    out = swe.ShallowWaterState(mesh, 4)
    with pytest.raises(swe.OmegaError, match="static scratch"):
        model.compute_tendency_noalloc(out, state, 0.0)
