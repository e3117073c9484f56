"""fasterhenry_amd: MI355X-native PEEC extraction hot path (placeholder init)."""
from .model import *  # noqa: F401,F403
