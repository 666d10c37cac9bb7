"""A piquasso connector whose permanents run on this engine.

piquasso's passive simulators reach the permanent through their connector:
``NumpyConnector`` binds ``permanent`` and ``permanent_laplace`` of the
native module (piquasso/_simulators/connectors/numpy_/connector.py:53-54),
and ``BuiltinSimulator._validate_connector`` accepts any subclass of it
(piquasso/_simulators/simulator.py:41-55).  :func:`connector_class` builds
that subclass with the two methods bound to :func:`piquasso_amd.permanent`
and :func:`piquasso_amd.permanent_laplace` (same argument names, result
types and errors); everything else stays piquasso's.  piquasso is imported
only when the class is built, so this package does not depend on it.

Usage::

    import piquasso as pq
    from piquasso_amd.connector import connector_class

    simulator = pq.SamplingSimulator(d=40, connector=connector_class()())
"""

from __future__ import annotations

import importlib

__all__ = ["connector_class"]

_cache = {}


def connector_class(base=None):
    """The ``NumpyConnector`` subclass (or a subclass of ``base``, a
    connector class of the caller's choice) whose ``permanent`` and
    ``permanent_laplace`` call the GPU engine."""
    if base is None:
        try:
            mod = importlib.import_module("piquasso._simulators.connectors.numpy_.connector")
        except ImportError as e:
            raise ImportError("piquasso is not importable: the connector class needs it "
                              "(or pass a connector class as `base`)") from e
        base = mod.NumpyConnector
    if base in _cache:
        return _cache[base]
    from ._math.permanent import permanent, permanent_laplace

    def _permanent(self, matrix, rows, cols):
        return permanent(matrix, rows, cols)

    def _permanent_laplace(self, matrix, rows, cols):
        return permanent_laplace(matrix, rows, cols)

    cls = type("MI355XConnector", (base,), {
        "__doc__": f"{base.__name__} whose permanents run on the MI355X engine.",
        "permanent": _permanent,
        "permanent_laplace": _permanent_laplace,
    })
    _cache[base] = cls
    return cls
