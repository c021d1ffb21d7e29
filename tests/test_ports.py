"""Host-side checks of the N-port front end (the reference's
tests/test_dc_ports.py:215-234): port sets and correlations that the
reduction must reject, and the driven-port block it keeps."""

from __future__ import annotations

import numpy as np
import pytest

from electrical.matrix_free_mpir_fem import PortSet, reduce_correlation

PADS = (((0, 0, 0), (0, 1, 0)), ((0, 4, 0), (0, 3, 0)), ((1, 2, 7), (1, 1, 7), (1, 3, 7)))


def test_port_sets_are_validated():
    ports = PortSet(PADS, names=("A", "B", "load"), reference=2)
    assert ports.count == 3 and ports.driven == (0, 1)
    assert PortSet(PADS).names == ("port_0", "port_1", "port_2")
    with pytest.raises(ValueError, match="at least two ports"):
        PortSet(pads=(PADS[0],))
    with pytest.raises(ValueError, match="belongs to ports"):
        PortSet(pads=(PADS[0], PADS[0]))
    with pytest.raises(ValueError, match="at least one node"):
        PortSet(pads=(PADS[0], ()))
    with pytest.raises(ValueError, match="reference"):
        PortSet(pads=PADS, reference=3)
    with pytest.raises(ValueError, match="unique"):
        PortSet(pads=PADS, names=("a", "a", "b"))
    terminals = ports.voltage_terminals((1.0, 0.0, 0.0))
    assert [t.name for t in terminals] == ["A", "B", "load"]
    assert [t.voltage_v for t in terminals] == [1.0, 0.0, 0.0]
    with pytest.raises(ValueError, match="one voltage per port"):
        ports.voltage_terminals((1.0, 0.0))


def test_correlations_are_reduced_to_the_driven_ports():
    ports = PortSet(PADS, reference=1)
    i = np.array([0.7, -1.2, 0.5])
    reduced, scale = reduce_correlation(ports, np.outer(i, i))
    np.testing.assert_allclose(reduced, np.outer(i[[0, 2]], i[[0, 2]]))
    assert scale == pytest.approx(1.44)  # the largest entry, at least 1
    with pytest.raises(ValueError, match="3×3"):
        reduce_correlation(ports, np.eye(2))
    with pytest.raises(ValueError, match="symmetric"):
        reduce_correlation(ports, np.array([[1.0, -1.0, 0.0], [-0.5, 1.0, -0.5], [-0.5, 0.0, 0.5]]))
    with pytest.raises(ValueError, match="rows must sum to zero"):
        reduce_correlation(ports, np.eye(3))
