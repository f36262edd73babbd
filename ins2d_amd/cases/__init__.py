"""Case plugins: the reference's case-file API restated in Python.

A case module exposes the functions the reference resolves by name for a
case ``<case>`` (readme.md:9-21; boundary_conditions.m:7-8,56,70;
initialize.m:6; mesh_generation.m:7):

``parameters()``            -> dict of ``<case>_parameters.m`` variables
``BCtype()``                -> {'BC_unsteady': 0, 'u': {...}, 'v': {...}}
``mesh(options)``           -> (x, y) grid-line coordinates
``IC(t, options)``          -> (u, v, p) 2-D arrays on the u/v/p points
``uBC(x, y, t, options)``   -> u boundary values, length len(x)*len(y)
``vBC(x, y, t, options)``   -> v boundary values, length len(x)*len(y)
``dudtBC``, ``dvdtBC``      -> their time derivatives (cases with
                               BC_unsteady = 1; set_bc_vectors.m:73-79)

``load(name, **overrides)`` returns a case object whose ``parameters()``
includes the overrides.
"""
import importlib
from types import SimpleNamespace


def load(name, **overrides):
    mod = importlib.import_module(f"ins2d_amd.cases.{name}")
    base = mod.parameters

    def parameters():
        p = dict(base())
        p.update(overrides)
        return p

    return SimpleNamespace(name=name, parameters=parameters, BCtype=mod.BCtype, mesh=mod.mesh,
                           IC=mod.IC, uBC=mod.uBC, vBC=mod.vBC, dudtBC=getattr(mod, 'dudtBC', None),
                           dvdtBC=getattr(mod, 'dvdtBC', None), module=mod)
