"""Operator interface on the CPU: algebra, representations, shape
propagation (the reference's tests/test_operator.py, test_shaped_operators.py,
test_radon_shapes.py) and extent-placeholder resolution against the 240
expected extents of tests/test_extent_placeholder.py (golden JSON)."""

from __future__ import annotations

import json
import os

import numpy as np
import pytest

from gratopy_amd.operator import Fanbeam, Radon
from gratopy_amd.operator.base import IDENTITY, ZERO, Operator, OperatorArithmeticOperation
from gratopy_amd.utilities import Angles, Detectors, ExtentPlaceholder, ImageDomain

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "extent_placeholder_expected.json")


def test_identity_and_zero():
    assert repr(IDENTITY) == "[Id]" and repr(ZERO) == "[0]"
    assert IDENTITY.apply_to([1, 2, 3]) == [1, 2, 3]
    assert IDENTITY * [1, 2, 3] == [1, 2, 3]
    data = np.array([1, 2, 3])
    np.testing.assert_equal(ZERO.apply_to(data), np.zeros_like(data))
    np.testing.assert_equal(ZERO * data, np.zeros_like(data))
    assert ZERO + ZERO == ZERO and ZERO - ZERO == ZERO
    assert ZERO + IDENTITY == IDENTITY and IDENTITY + ZERO == IDENTITY and IDENTITY - ZERO == IDENTITY
    assert 0 * ZERO == ZERO and 42 * ZERO == ZERO
    assert ZERO * ZERO == ZERO and ZERO * IDENTITY == ZERO and IDENTITY * ZERO == ZERO
    A = Operator(name="A")
    assert IDENTITY * A == A and A * IDENTITY == A


def test_representations():
    A, B, C = Operator(name="A"), Operator(name="B"), Operator(name="C")
    assert repr(A * B) == "A*B"
    assert repr(A + B) == "A + B"
    assert repr(5 * (A + IDENTITY) - B) == "5*A + 5*[Id] + (-1)*B"
    assert repr(5 * (A + IDENTITY) * C - B) == "(5*A + 5*[Id])*C + (-1)*B"
    assert repr(A * B * A * B * A * B) == "A*B*A*B*A*B"
    assert repr(5 * A * B * A * B * A * B) == "5*A*B*A*B*A*B"
    assert repr(5 * (A * B * B * A)) == "5*A*B*B*A"


def test_composition_structure():
    A, B = Operator(name="A"), Operator(name="B")
    op = 5 * (A + IDENTITY) - B
    assert op.is_composite()
    assert op._arithmetic_operation == OperatorArithmeticOperation.ADDITION
    assert len(op._operands) == 3
    assert 5 * (A + B + A) == 5 * A + 5 * B + 5 * A
    assert 5 * (A * B * B * A) == 5 * A * B * B * A


def test_shape_propagation():
    A = Operator(name="A", input_shape=(10,), output_shape=(20,))
    with pytest.raises(ValueError, match="Input shape mismatch"):
        _ = A + Operator(name="B", input_shape=(11,), output_shape=(20,))
    with pytest.raises(ValueError, match="Output shape mismatch"):
        _ = A + Operator(name="B", input_shape=(10,), output_shape=(21,))
    B = Operator(name="B")
    assert (A + B).input_shape == (10,) and (B + A).output_shape == (20,)
    P = Operator(name="P", input_shape=(20,), output_shape=(30,))
    assert (P * A).input_shape == (10,) and (P * A).output_shape == (30,)
    with pytest.raises(ValueError, match="Shape mismatch"):
        _ = P * Operator(name="Q", input_shape=(10,), output_shape=(25,))
    chain = P * Operator(name="X") * Operator(name="Y") * A
    assert chain.input_shape == (10,) and chain.output_shape == (30,)
    nested = (P + P) * A
    assert nested.input_shape == (10,) and nested.output_shape == (30,)
    assert (IDENTITY * A).input_shape == (10,) and (A * IDENTITY).output_shape == (20,)
    assert (A + ZERO).input_shape == (10,) and (ZERO + A).output_shape == (20,)
    assert (IDENTITY + ZERO) is IDENTITY


def test_radon_shapes_and_transpose():
    R = Radon(image_domain=128, angles=180, detectors=200)
    assert R.input_shape == (128, 128) and R.output_shape == (200, 180)
    assert R.T.input_shape == (200, 180) and R.T.output_shape == (128, 128)
    assert repr(R.T) == "Radon.T" and R.T.T.adjoint is False
    assert (R.T * R).input_shape == (128, 128) and (R * R.T).output_shape == (200, 180)
    with pytest.raises(ValueError, match="Input shape mismatch"):
        _ = Radon(image_domain=128, angles=180) + Radon(image_domain=129, angles=180)
    with pytest.raises(ValueError, match="Shape mismatch"):
        _ = Radon(image_domain=128, angles=180) * Radon(image_domain=129, angles=180)
    with pytest.raises(ValueError, match="Shape mismatch"):
        _ = Radon(image_domain=128, angles=180).T * Radon(image_domain=128, angles=100)


def test_radon_angles_defaults():
    R = Radon(image_domain=16, angles=10)
    assert R.angles.half_circle is True
    np.testing.assert_allclose(R.angles.angles, np.linspace(0, np.pi, 10, endpoint=False))
    np.testing.assert_allclose(R.angles.weights, np.full(10, np.pi / 10))
    angles = Angles.uniform(10)
    R = Radon(image_domain=16, angles=angles)
    assert R.angles is angles and R.angles.half_circle is False
    np.testing.assert_allclose(R.angles.weights, np.full(10, 2 * np.pi / 10))
    assert R.detectors.number == int(np.ceil(np.hypot(16, 16))) and R.detectors.extent == 2.0


def test_angles_constructors():
    a = Angles.uniform_interval(0.0, 1.0, 4)
    np.testing.assert_allclose(a.angles, [0.125, 0.375, 0.625, 0.875])
    np.testing.assert_allclose(a.weights, 0.25)
    b = Angles.intervals([2, 3], [0.0, 2.0], [1.0, 3.5])
    assert len(b) == 5 and np.isclose(b.weights.sum(), 2.5)
    c = Angles.from_list([0.0, np.pi / 2, np.pi], half_circle=True)
    np.testing.assert_allclose(c.weights, [np.pi / 4, np.pi / 2, np.pi / 4])
    d = Detectors(-12)
    assert d.number == 12 and d.reversed is True
    with pytest.raises(ValueError):
        Angles([0.0, 1.0], [1.0])


def test_radon_placeholders_resolve():
    r = Radon(ImageDomain(size=16, extent=2.0), 10, Detectors(number=20, extent=ExtentPlaceholder.FULL))
    assert isinstance(r.detectors.extent, float) and r.detectors.extent > 0
    r = Radon(ImageDomain(size=16, extent=ExtentPlaceholder.FULL), 10, Detectors(number=20, extent=2.0))
    assert isinstance(r.image_domain.extent, float) and r.image_domain.extent > 0
    with pytest.raises(NotImplementedError):
        Radon(ImageDomain(size=16, extent=ExtentPlaceholder.FULL), 10,
              Detectors(number=20, extent=ExtentPlaceholder.VALID))


def _extent_cases(kind):
    data = json.load(open(GOLDEN))
    for key, expected in data[f"{kind}_placeholder"].items():
        name, ratio = key.split(",")
        idx = 0
        for dc in data["detector_centers"]:
            for ic in data["image_centers"]:
                yield ExtentPlaceholder[name], float(ratio), dc, tuple(ic), expected[idx]
                idx += 1


@pytest.mark.parametrize("placeholder,ratio,dc,ic,expected", list(_extent_cases("detector")))
def test_detector_extent_placeholder(placeholder, ratio, dc, ic, expected):
    Ny = int(ratio * 400)
    make = lambda: Radon(ImageDomain(size=(400, Ny), center=ic, extent=2.0), 180,
                         Detectors(number=100, center=dc, extent=placeholder))
    if expected is None:
        with pytest.raises(ValueError):
            make()
    else:
        assert make().detectors.extent == pytest.approx(expected, abs=0.02)


@pytest.mark.parametrize("placeholder,ratio,dc,ic,expected", list(_extent_cases("image")))
def test_image_extent_placeholder(placeholder, ratio, dc, ic, expected):
    Ny = int(ratio * 400)
    make = lambda: Radon(ImageDomain(size=(400, Ny), center=ic, extent=placeholder), 180,
                         Detectors(number=100, center=dc, extent=2.0))
    if expected is None:
        with pytest.raises(ValueError):
            make()
    else:
        assert make().image_domain.extent == pytest.approx(expected, abs=0.02)


def test_apply_without_queue_raises():
    R = Radon(image_domain=16, angles=10)
    with pytest.raises(ValueError, match="No device queue available"):
        R.apply_to(np.zeros((16, 16), np.float32))


def test_fanbeam_operator_shapes():
    F = Fanbeam((800.0, 400.0), 64, 90, 100)
    assert F.input_shape == (64, 64) and F.output_shape == (100, 90)
    assert F.T.input_shape == (100, 90) and repr(F.T) == "Fanbeam.T"
    assert F.angles.half_circle is False
    G = Fanbeam(600.0, 32, 10)
    assert G.state["source_distances"] == (600.0, 300.0)


class _Matrix(Operator):
    """A leaf applying a fixed matrix (host numpy), for the algebra's semantics."""

    def __init__(self, name, M):
        super().__init__(name=name, input_shape=(M.shape[1],), output_shape=(M.shape[0],))
        self.M = M

    def apply_to(self, argument, output=None, **kwargs):
        result = self.scalar * (self.M @ argument)
        if output is None:
            return result
        output[...] = result
        return output


def test_algebra_applies_as_linear_maps():
    rng = np.random.default_rng(0)
    Ma, Mb, Mc = rng.standard_normal((4, 3)), rng.standard_normal((4, 3)), rng.standard_normal((3, 3))
    A, B, C = _Matrix("A", Ma), _Matrix("B", Mb), _Matrix("C", Mc)
    x = rng.standard_normal(3)
    np.testing.assert_allclose((2 * A - 3 * B) * x, 2 * Ma @ x - 3 * Mb @ x)
    np.testing.assert_allclose(((A + B) * C) * x, (Ma + Mb) @ (Mc @ x))
    np.testing.assert_allclose((4 * (A * C * C)) * x, 4 * Ma @ Mc @ Mc @ x)
    out = np.zeros(4)
    assert (A * C).apply_to(x, output=out) is out
    np.testing.assert_allclose(out, Ma @ Mc @ x)
    # factors of composed leaves move to the leftmost factor; the operands stay unchanged
    P = A * (2 * C)
    assert repr(P) == "2*A*C" and C.scalar == 1 and A.scalar == 1
    P.scalar = 3  # a composition multiplies its leftmost factor
    assert repr(P) == "6*A*C"
    S = A + B
    S.scalar = -1  # a sum multiplies every term
    assert repr(S) == "(-1)*A + (-1)*B" and S.scalar == 1
    assert ZERO * A == ZERO and A * ZERO == ZERO
    np.testing.assert_allclose((A + ZERO) * x, Ma @ x)
