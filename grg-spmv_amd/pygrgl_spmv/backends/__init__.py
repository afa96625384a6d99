"""Backends. The HIP backend lives in ``pygrgl_spmv.backends.hip`` (imported on demand)."""
