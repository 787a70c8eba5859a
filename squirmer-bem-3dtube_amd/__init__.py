"""squirmer-bem-3dtube_amd: MI355X-native squirmer-in-tube BEM (libsqbem) and its Python binding.

The directory name is not an importable identifier; load it with ``load_package()`` from
``__graft_entry__`` or import ``sqbem`` with this directory on ``sys.path``.
"""
from .sqbem import *  # noqa: F401,F403
