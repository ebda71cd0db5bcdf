"""Autograd operators, renderer and synthetic scenes on top of the gfx950
``cudaExtensions`` op surface (the counterpart of the reference's ``splatting``
package)."""
from . import utils, scenes  # noqa: F401
