"""evsl_amd: MI355X-native filtered-Lanczos / KPM hot path behind EVSL's C API."""
