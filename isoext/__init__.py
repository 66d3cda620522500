"""``import isoext`` compatibility package: re-exports isoext_amd under the
reference's module name so existing code and tests run unchanged."""

from isoext_amd import *  # noqa: F401,F403
from isoext_amd import __all__  # noqa: F401
from isoext_amd import sdf, utils  # noqa: F401
