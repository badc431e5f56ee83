"""Drop-in package for the reference's import path (``from cusingler import
cusingler``): the compiled pybind11 module ``cusingler.cusingler`` built by
``python -m cusingler_amd.build``, exposing ``run(...)``."""
