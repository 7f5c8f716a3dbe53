"""Substrate base classes: the duck type the solver consumes.

Mirrors ``ContactMechanics/Substrates.py:34-184`` (``Substrate.Error``,
``is_periodic``, ``check``, ``ElasticSubstrate.compute``).
"""

import abc


class Substrate(object, metaclass=abc.ABCMeta):
    """Generic base class of all substrates."""

    _periodic = None
    name = "generic_halfspace"

    class Error(Exception):
        pass

    def spawn_child(self, dummy):
        raise self.Error("Only substrates with free boundaries can do this")

    @classmethod
    def is_periodic(cls):
        if cls._periodic is not None:
            return cls._periodic
        raise cls.Error(f"periodicity of Substrate type '{cls.name}' ('{cls.__name__}') is not defined")

    @property
    @abc.abstractmethod
    def nb_domain_grid_pts(self):
        pass

    @property
    @abc.abstractmethod
    def nb_subdomain_grid_pts(self):
        pass

    @property
    @abc.abstractmethod
    def subdomain_locations(self):
        pass

    @property
    @abc.abstractmethod
    def communicator(self):
        pass

    def check(self, force=None):
        """All forces are acceptable for a generic substrate."""


class ElasticSubstrate(Substrate, metaclass=abc.ABCMeta):
    """Generic base class of elastic substrates."""

    name = "generic_elastic_halfspace"

    def __init__(self):
        self.energy = None
        self.force = None
        self.force_k = None

    def compute(self, disp, pot=True, forces=False):
        self.energy, self.force = self.evaluate(disp, pot, forces)

    def evaluate(self, disp, pot=True, forces=False):
        raise NotImplementedError()
