"""The piquasso connector class (piquasso_amd/connector.py): a subclass of
the connector it is given whose permanent / permanent_laplace are the
engine's (piquasso/_simulators/connectors/numpy_/connector.py:53-54)."""

import numpy as np
import pytest


def test_connector_class_binds_the_engine(monkeypatch):
    """connector_class() subclasses the given connector and routes
    permanent / permanent_laplace (keyword names matrix, rows, cols, as
    piquasso's callers pass them) to the engine's functions; without
    piquasso and without a base class it raises ImportError."""
    import piquasso_amd._math.permanent as PM
    from piquasso_amd import connector as C

    calls = []
    monkeypatch.setattr(PM, "permanent", lambda m, r, c: calls.append(("perm", r, c)) or 1.5)
    monkeypatch.setattr(PM, "permanent_laplace", lambda m, r, c: calls.append(("lap", r, c)) or [2.0])
    C._cache.clear()

    class Base:
        def permanent(self, matrix, rows, cols):
            raise AssertionError("the base method must be overridden")

        def other(self):
            return "kept"

    cls = C.connector_class(Base)
    conn = cls()
    assert isinstance(conn, Base) and conn.other() == "kept"
    assert conn.permanent(np.eye(2), cols=[1, 1], rows=[1, 1]) == 1.5
    assert conn.permanent_laplace(np.eye(2), [1, 0], [1, 1]) == [2.0]
    assert calls == [("perm", [1, 1], [1, 1]), ("lap", [1, 0], [1, 1])]
    assert C.connector_class(Base) is cls
    C._cache.clear()
    with pytest.raises(ImportError):
        C.connector_class()

