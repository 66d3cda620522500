"""Alias of isoext_amd.utils under the reference's module path (isoext.utils)."""

from isoext_amd.utils import gaussian_smooth, make_grid, write_obj  # noqa: F401
