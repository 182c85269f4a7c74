"""The reference extension's import name (lib/torch_extension/setup.py:28-96 installs a top-level
``pyFastFss``): ``import pyFastFss`` gives this library's ops unchanged, the module object of
``fastfunctionsecretsharing_amd.pyFastFss`` itself."""
import sys

from fastfunctionsecretsharing_amd import pyFastFss as _ops

sys.modules[__name__] = _ops
