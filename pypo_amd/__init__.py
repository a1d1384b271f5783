"""pypo_amd: MI355X-native physical optics and ray tracing (PyPO hot path)."""
