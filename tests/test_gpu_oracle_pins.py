"""The oracle's pins, run again in the GPU test session.

The GPU parity tests trust the CPU oracle; the pins that make it
trustworthy (tests/test_oracle.py) live with the CPU tests, which a GPU run
deselects. This module runs the self-contained ones wherever the GPU tests
run: bit-for-bit agreement with the reference's own Fortran build
(oracle/_ref, which travels with the tree: here its absence is a failure,
not a skip), closed forms, dense solves, the reference-held goldens and the
planted-mistake checks. Nothing here uses the GPU.
"""
import os

import pytest

import test_oracle as pins

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def po(oracle_mod):
    assert oracle_mod.ref_available(), "oracle/_ref/libpyranda_ref.so is missing: build it with make -C oracle"
    return oracle_mod


@pytest.mark.parametrize("op", range(5))
def test_weights_bitexact_vs_reference_build(po, op):
    pins.test_weights_bitexact_vs_reference_build(po, op)


@pytest.mark.parametrize("op", [0, 1, 2, 3])
@pytest.mark.parametrize("n", [8, 100])
@pytest.mark.parametrize("periodic", [True, False])
def test_lu_vs_reference_build(po, op, n, periodic):
    pins.test_lu_vs_reference_build(po, op, n, periodic)


def test_periodic_closed_form(po):
    pins.test_periodic_closed_form(po, 32)


def test_filters_kill_nyquist(po):
    pins.test_filters_kill_nyquist(po)


def test_composite_closed_forms(po):
    pins.test_laplacian_closed_form(po)
    pins.test_divergence_closed_form(po)
    pins.test_ringv_closed_form(po)
    pins.test_divergence_ringv_symm_components(po)


def test_goldens(po):
    pins.test_unit_test_golden(po)
    pins.test_euler_2d_curve_golden(po, 64)
    pins.test_advection_golden_through_oracle(po, 100)
    pins.test_heat1d_golden(po, 32)


def test_planted_mistakes_are_caught(po, monkeypatch):
    pins.test_planted_mistake_is_caught(po, monkeypatch)
    pins.test_advection_planted_dt_mistake_is_caught(po)


def test_mlp_oracle_pins():
    pins.test_mlp_oracle_backward_is_the_loss_gradient()
    pins.test_mlp_oracle_bf16_rounding_is_torch_bfloat16()
    pins.test_mlp_oracle_features_layout()
