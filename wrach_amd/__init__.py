"""wrach_amd: wrach's per-frame particle step on MI355X (gfx950), behind a
C ABI (include/wrach.h) that mirrors wrach's runners/api surface."""
from .api import (PackedData, Particle, ShaderWorldSettings, World, WrachAPI,  # noqa: F401
                  WrachConfig, WrachError, WrachState)
