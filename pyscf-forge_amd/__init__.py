"""MI355X direct-CI sigma engine (placeholder during bring-up)."""
from . import synthetic  # noqa: F401
