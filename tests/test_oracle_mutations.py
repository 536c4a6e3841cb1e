"""Each pin of the oracle can fail: the oracle is rebuilt with one planted
mistake at a time (-DORACLE_MUTATION=n, on the CPU) and the closed-form checks
of test_oracle_{rebin,filter,fdk}.py are run against it; every mutant must be
rejected by at least one of them.

    1  rebinning: fan angle Theta - Beta instead of Theta + Beta
    2  rebinning: ">=" clamp of the reference's GPU path instead of the CPU ">"
    3  filter: negative taps shifted by one slot in the wrap-around layout
    5  backprojection: conjugate offset temp2 +- 0.5 with the wrong sign
    6  backprojection: conjugate weight read at the view's own index
    7  backprojection: G and Gc denominators swapped
    8  backprojection: channel from -Yr
    9  backprojection: row coordinate without the SO/sqrt(SO^2-Yr^2) factor
       (no closed form reaches it; the bit-for-bit comparison with the
       reference's own compiled CPU function in test_oracle_ref.py rejects it,
       as it rejects every other mutant)
"""
import math

import numpy as np
import pytest

import oracle

from . import test_oracle_fdk as F
from . import test_oracle_filter as T
from . import test_oracle_rebin as R


def _fails(fn, *args):
    try:
        fn(*args)
    except AssertionError:
        return True
    return False


@pytest.fixture(scope="module")
def variants(tmp_path_factory):
    d = tmp_path_factory.mktemp("oracle_mutants")
    return {m: oracle.build_variant(m, d) for m in (1, 2, 3, 5, 6, 7, 8, 9)}


def test_unmutated_build_passes(tmp_path):
    path = oracle.build_variant(0, tmp_path)
    with oracle.using(path):
        R.test_affine_input_lands_on_textbook_coordinates()
        T.test_impulse_response_is_the_kernel()
        F.test_constant_projection_gives_pi_c(1.0, 5.0)


@pytest.mark.parametrize("m", [1, 2])
def test_rebin_mutants_are_rejected(variants, m):
    with oracle.using(variants[m]):
        assert _fails(R.test_affine_input_lands_on_textbook_coordinates)


def test_filter_mutant_is_rejected(variants):
    with oracle.using(variants[3]):
        assert _fails(T.test_impulse_response_is_the_kernel)
        assert _fails(T.test_constant_row_telescoping_closed_form)


@pytest.mark.parametrize("m", [5, 6])
def test_weighting_mutants_are_rejected(variants, m):
    with oracle.using(variants[m]):
        assert _fails(F.test_constant_projection_gives_pi_c, 1.0, 5.0)


def test_cone_angle_mutant_is_rejected(variants):
    # swapping the two denominators keeps the conjugate-pair identity, but not
    # the geometric cone angles
    with oracle.using(variants[7]):
        assert not _fails(F.test_constant_projection_gives_pi_c, 1.0, 5.0)
        assert _fails(F.test_3d_weight_uses_geometric_cone_angles, 80.0, -40.0, 2.85)


def test_channel_mutant_is_rejected(variants):
    with oracle.using(variants[8]):
        assert _fails(F.test_channel_is_parallel_beam_offset, 30.0, -12.0)


def test_row_factor_mutant_escapes_the_closed_forms(variants):
    """Mutation 9 drops SO/sqrt(SO^2-Yr^2) from the row coordinate.  The
    central-ray and constant-projection pins cannot see it (Yr = 0 there, and a
    constant projection ignores the row); the comparison with the reference's
    compiled CPU function (test_oracle_ref.py::test_fdk_mutants_differ_from_reference)
    is what pins that factor."""
    with oracle.using(variants[9]):
        assert not _fails(F.test_central_ray_row_is_textbook_magnification, 120.0, -0.6)
        assert not _fails(F.test_constant_projection_gives_pi_c, 1.0, 5.0)
