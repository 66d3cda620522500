"""Alias of isoext_amd.sdf under the reference's module path (isoext.sdf)."""

from isoext_amd.sdf import *  # noqa: F401,F403
from isoext_amd.sdf import get_sdf_grad, get_sdf_normal  # noqa: F401
